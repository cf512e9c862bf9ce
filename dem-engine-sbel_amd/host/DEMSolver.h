// DEMSolver: C++ facade with the reference's user-facing API (src/DEM/API.h:44-887) for the per-step hot path,
// running on the MI355X C ABI (include/deme/deme.h). Names, argument meaning and error behaviour (std::runtime_error
// on invalid input) follow the reference; entry points outside the hot path are absent.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/deme/deme.h"

namespace deme {

// Verbosity levels (src/DEM/Defines.h:104)
enum VERBOSITY { QUIET = 0, ERROR = 10, WARNING = 20, INFO = 30, STEP_STATS = 32, STEP_METRIC = 35, DEBUG = 40 };
// Stepping methods (src/DEM/Defines.h:106)
enum class TIME_INTEGRATOR { FORWARD_EULER, CENTERED_DIFFERENCE, EXTENDED_TAYLOR, CHUNG };
// Spatial directions (src/DEM/Defines.h:133)
enum class SPATIAL_DIR { X, Y, Z, NONE };
enum class FORCE_MODEL { HERTZIAN, HERTZIAN_FRICTIONLESS, CUSTOM };
// Output file formats (src/DEM/Defines.h:114): CSV is written; BINARY leaves an empty file, as the reference does;
// CHPF (a third-party binary format) is not built and throws.
enum class OUTPUT_FORMAT { CSV, BINARY, CHPF };
// Columns of the sphere and clump files beyond the position (src/DEM/Defines.h:118-131). The reference's writers
// emit ABSV, VEL and FAMILY; the other flags are accepted and write nothing, as there.
enum OUTPUT_CONTENT {
    XYZ = 0,
    QUAT = 1,
    ABSV = 2,
    VEL = 4,
    ANG_VEL = 8,
    ACC = 16,
    ANG_ACC = 32,
    FAMILY = 64,
    MAT = 128,
    EXP_FACTOR = 256
};
// Columns of the contact file beyond the owners and the type (src/DEM/Defines.h:135-143); COMPONENT writes nothing.
enum CNT_OUTPUT_CONTENT { OWNERS = 0, FORCE = 1, POINT = 2, COMPONENT = 4, NORMAL = 8, TORQUE_ONLY_FORCE = 16 };

// Column names of the output files (src/DEM/Structs.h:584-611): the on-disk format, kept so files interchange with
// the reference's tools and readers.
const std::string OUTPUT_FILE_X_COL_NAME = "X";
const std::string OUTPUT_FILE_Y_COL_NAME = "Y";
const std::string OUTPUT_FILE_Z_COL_NAME = "Z";
const std::string OUTPUT_FILE_R_COL_NAME = "r";
const std::string OUTPUT_FILE_CLUMP_TYPE_NAME = "clump_type";
const std::string OUTPUT_FILE_OWNER_1_NAME = "A";
const std::string OUTPUT_FILE_OWNER_2_NAME = "B";
const std::string OUTPUT_FILE_COMP_1_NAME = "compA";
const std::string OUTPUT_FILE_COMP_2_NAME = "compB";
const std::string OUTPUT_FILE_CNT_TYPE_NAME = "contact_type";
const std::string OUTPUT_FILE_FORCE_X_NAME = "f_x";
const std::string OUTPUT_FILE_FORCE_Y_NAME = "f_y";
const std::string OUTPUT_FILE_FORCE_Z_NAME = "f_z";
const std::string OUTPUT_FILE_TOF_X_NAME = "tof_x";
const std::string OUTPUT_FILE_TOF_Y_NAME = "tof_y";
const std::string OUTPUT_FILE_TOF_Z_NAME = "tof_z";
const std::string OUTPUT_FILE_NORMAL_X_NAME = "n_x";
const std::string OUTPUT_FILE_NORMAL_Y_NAME = "n_y";
const std::string OUTPUT_FILE_NORMAL_Z_NAME = "n_z";
const std::string OUTPUT_FILE_SPH_SPH_CONTACT_NAME = "SS";
const std::string OUTPUT_FILE_SPH_ANAL_CONTACT_NAME = "SA";
const std::string OUTPUT_FILE_SPH_MESH_CONTACT_NAME = "SM";

constexpr unsigned int RESERVED_FAMILY_NUM = 255;
constexpr unsigned int DEFAULT_CLUMP_FAMILY_NUM = 0;

// Material (src/DEM/Structs.h, DEMMaterial): named float properties; load order is the material index.
struct DEMMaterial {
    std::unordered_map<std::string, float> mat_prop;
    unsigned int load_order = 0;
    explicit DEMMaterial(const std::unordered_map<std::string, float>& props) : mat_prop(props) {}
};

// Clump template (src/DEM/Structs.h:389-477).
class DEMClumpTemplate {
  public:
    float mass = 0;
    float3 MOI = make_float3(0, 0, 0);
    std::vector<float> radii;
    std::vector<float3> relPos;
    std::vector<std::shared_ptr<DEMMaterial>> materials;
    unsigned int nComp = 0;
    unsigned int mark = 0;
    std::string m_name = "NULL";
    float volume = 0.0;
    void SetVolume(float vol) { volume = vol; }
    void AssignName(const std::string& name) { m_name = name; }
    // Scale all geometry components (Structs.h:465-474)
    void Scale(float s);
    // Re-express component positions in the centroid/principal frame (Structs.h:439-449)
    void InformCentroidPrincipal(float3 center, float4 prin_Q);
    // Rotate then move the components (Structs.h:452-461)
    void Move(float3 vec, float4 rot_Q);
};

// Batch of clumps to add (src/DEM/Structs.h:480-559).
class DEMClumpBatch {
  private:
    const size_t nClumps;
    void assertLength(size_t len, const std::string& name) const;

  public:
    bool family_isSpecified = false;
    std::vector<std::shared_ptr<DEMClumpTemplate>> types;
    std::vector<unsigned int> families;
    std::vector<float3> vel;
    std::vector<float3> angVel;
    std::vector<float3> xyz;
    std::vector<float4> oriQ;
    size_t load_order = 0;
    explicit DEMClumpBatch(size_t num);
    size_t GetNumClumps() const { return nClumps; }
    void SetTypes(const std::vector<std::shared_ptr<DEMClumpTemplate>>& input);
    void SetTypes(const std::shared_ptr<DEMClumpTemplate>& input);
    void SetType(const std::shared_ptr<DEMClumpTemplate>& input) { SetTypes(input); }
    void SetPos(const std::vector<float3>& input);
    void SetPos(float3 input) { SetPos(std::vector<float3>(nClumps, input)); }
    void SetVel(const std::vector<float3>& input);
    void SetVel(float3 input) { SetVel(std::vector<float3>(nClumps, input)); }
    void SetAngVel(const std::vector<float3>& input);
    void SetAngVel(float3 input) { SetAngVel(std::vector<float3>(nClumps, input)); }
    void SetOriQ(const std::vector<float4>& input);
    void SetOriQ(float4 input) { SetOriQ(std::vector<float4>(nClumps, input)); }
    void SetFamilies(const std::vector<unsigned int>& input);
    void SetFamilies(unsigned int input) { SetFamilies(std::vector<unsigned int>(nClumps, input)); }
    void SetFamily(unsigned int input) { SetFamilies(input); }
};

// External object made of analytical components (src/DEM/BdrsAndObjs.h:100-170); planes only on this path.
class DEMExternObj {
  public:
    struct Plane {
        float3 position;
        float3 normal;
        std::shared_ptr<DEMMaterial> material;
    };
    unsigned int family_code = RESERVED_FAMILY_NUM;
    float3 init_pos = make_float3(0, 0, 0);
    float mass = 1e15f;
    float3 MOI = make_float3(1e15f, 1e15f, 1e15f);
    unsigned int load_order = 0;
    std::vector<Plane> planes;
    void SetFamily(unsigned int code);
    void SetMass(float m) { mass = m; }
    void SetMOI(float3 m) { MOI = m; }
    void SetInitPos(float3 p) { init_pos = p; }
    void AddPlane(const float3 pos, const float3 normal, const std::shared_ptr<DEMMaterial>& material);
};

// Force model in use (AuxClasses.h:123-160): a built-in type, or a custom model body (the reference's model-string
// names; compiled at Initialize by hipRTC) with the names of its per-contact wildcards.
struct DEMForceModel {
    FORCE_MODEL type = FORCE_MODEL::HERTZIAN;
    std::string model;
    std::set<std::string> contact_wildcards = {"delta_tan_x", "delta_tan_y", "delta_tan_z", "delta_time"};
    void SetForceModelType(FORCE_MODEL t);
    void DefineCustomModel(const std::string& m);
    void SetPerContactWildcards(const std::set<std::string>& names);
};

class DEMSolver;

// Inspector (AuxClasses.h:24-74): the reduced value of a quantity over the whole domain or over a region given as C++
// statements reading the world position X, Y, Z and returning a bool. Quantities: clump_max_z, clump_min_z,
// clump_max_absv, clump_volume (the sum of template volumes, DEMClumpTemplate::SetVolume). An unknown quantity throws
// at creation; a malformed region throws at the first GetValue, which compiles it.
class DEMInspector {
  public:
    DEMInspector(DEMSolver* sys, const std::string& quantity, const std::string& region, bool all_domain);
    float GetValue();

  private:
    DEMSolver* sys;
    std::string quantity, region;
    bool all_domain;
};

// The owners a tracker follows (Structs.h:562-579): a batch of clumps or one external object.
struct DEMTrackedObj {
    enum class Kind { CLUMP_BATCH, EXTERNAL } kind = Kind::CLUMP_BATCH;
    size_t load_order = 0;
    uint32_t ownerID = std::numeric_limits<uint32_t>::max();  // assigned by Initialize
    size_t nSpanOwners = 1;
};

// Tracker (AuxClasses.h:77-121): get and set the state of the owners of a tracked batch or object; offset picks the
// owner within a batch.
class DEMTracker {
  public:
    explicit DEMTracker(DEMSolver* sys) : sys(sys) {}
    std::shared_ptr<DEMTrackedObj> obj;
    float3 Pos(size_t offset = 0);
    float3 AngVelLocal(size_t offset = 0);
    float3 Vel(size_t offset = 0);
    float4 OriQ(size_t offset = 0);
    unsigned int GetFamily(size_t offset = 0);
    float3 ContactAcc(size_t offset = 0);
    float3 ContactAngAccLocal(size_t offset = 0);
    void SetPos(float3 pos, size_t offset = 0);
    void SetAngVel(float3 angVel, size_t offset = 0);
    void SetVel(float3 vel, size_t offset = 0);
    void SetOriQ(float4 oriQ, size_t offset = 0);

  private:
    DEMSolver* sys;
    uint32_t owner(size_t offset) const;
};

class DEMSolver {
  public:
    explicit DEMSolver(unsigned int nGPUs = 1);
    ~DEMSolver();
    DEMSolver(const DEMSolver&) = delete;
    DEMSolver& operator=(const DEMSolver&) = delete;

    void SetVerbosity(VERBOSITY v) { verbosity = v; }

    // ---- world (API.h:55-92) ----
    void InstructBoxDomainDimension(float x, float y, float z, SPATIAL_DIR dir_exact = SPATIAL_DIR::NONE);
    void InstructBoxDomainNumVoxel(unsigned char x, unsigned char y, unsigned char z, float len_unit = 1e-10f);
    void InstructBoxDomainBoundingBC(const std::string& inst, const std::shared_ptr<DEMMaterial>& mat) {
        m_user_add_bounding_box = inst;
        m_bounding_box_material = mat;
    }
    void SetCoordSysOrigin(const std::string& where) { m_user_instructed_origin = where; }
    void SetCoordSysOrigin(float3 O) {
        m_boxLBF = O;
        m_user_instructed_origin = "explicit";
    }
    void SetGravitationalAcceleration(float3 g) { G = g; }
    void SetInitTimeStep(double ts_size) { m_ts_size = ts_size; }
    void SetCDUpdateFreq(int freq) { m_updateFreq = freq; }
    void SetIntegrator(TIME_INTEGRATOR intg) { m_integrator = intg; }
    void SetInitBinSize(double bin_size) {
        use_user_defined_bin_size = true;
        m_binSize = bin_size;
    }
    void SetExpandFactor(float beta, bool fix = true) {
        m_expand_factor = beta;
        use_user_defined_expand_factor = fix;
    }
    void SetMaxVelocity(float max_vel) { m_approx_max_vel = max_vel; }
    void SetExpandSafetyParam(float param) { m_expand_safety_param = param; }

    // ---- force model (API.h:120-130) ----
    std::shared_ptr<DEMForceModel> UseFrictionalHertzianModel();
    std::shared_ptr<DEMForceModel> UseFrictionlessHertzianModel();
    std::shared_ptr<DEMForceModel> DefineContactForceModel(const std::string& model);

    // ---- families (API.h:255-298) ----
    void SetFamilyFixed(unsigned int ID);
    void SetFamilyPrescribedLinVel(unsigned int ID, const std::string& velX, const std::string& velY,
                                   const std::string& velZ, bool dictate = true);
    void SetFamilyPrescribedAngVel(unsigned int ID, const std::string& velX, const std::string& velY,
                                   const std::string& velZ, bool dictate = true);
    void SetFamilyPrescribedPosition(unsigned int ID, const std::string& X, const std::string& Y, const std::string& Z);
    void DisableContactBetweenFamilies(unsigned int ID1, unsigned int ID2);
    void EnableContactBetweenFamilies(unsigned int ID1, unsigned int ID2);
    // ChangeFamilyWhen (API.h:262): each step, owners of family ID_from whose condition (C++ statements reading pos,
    // vel, acc, mass, family_code, t, h and ending in `return <bool>;`) holds move to family ID_to.
    void ChangeFamilyWhen(unsigned int ID_from, unsigned int ID_to, const std::string& condition);
    // ChangeFamily (API.h:268): all owners of family ID_from move to ID_to now.
    void ChangeFamily(unsigned int ID_from, unsigned int ID_to);
    // DisableFamilyOutput (API.h:260): owners of this family are left out of the sphere and clump files.
    void DisableFamilyOutput(unsigned int ID);

    // ---- output files (API.h:318-371, 430-436) ----
    void SetOutputFormat(OUTPUT_FORMAT format) { m_out_format = format; }
    void SetOutputContent(unsigned int content) { m_out_content = content; }
    void SetContactOutputFormat(OUTPUT_FORMAT format) { m_cnt_out_format = format; }
    void SetContactOutputContent(unsigned int content) { m_cnt_out_content = content; }
    // One row per clump: X,Y,Z,Qw,Qx,Qy,Qz,clump_type then the requested columns (dT.cpp:990-1052).
    void WriteClumpFile(const std::string& outfilename) const;
    // One row per component sphere: X,Y,Z,r then the requested columns (dT.cpp:896-986).
    void WriteSphereFile(const std::string& outfilename) const;
    // One row per pair of the contact list of the last step: A,B,contact_type then the requested columns
    // (dT.cpp:1054-1180). Contacts are written in the library's canonical order (by A's sphere; plane contacts first;
    // sphere partners by ascending id).
    void WriteContactFile(const std::string& outfilename) const;
    // Readers of the clump file (API.h:326-371): clump type name -> positions / quaternions (w, x, y, z), in file
    // order. Columns are found by name; others are ignored; blank lines are skipped; numbers are read as the
    // reference's CSV reader reads them. Throws std::runtime_error on a missing or repeated column, a row with too
    // few or too many fields, or a field that is not a number.
    static std::unordered_map<std::string, std::vector<float3>> ReadClumpXyzFromCsv(
        const std::string& infilename,
        const std::string& clump_header = OUTPUT_FILE_CLUMP_TYPE_NAME,
        const std::string& x_header = OUTPUT_FILE_X_COL_NAME,
        const std::string& y_header = OUTPUT_FILE_Y_COL_NAME,
        const std::string& z_header = OUTPUT_FILE_Z_COL_NAME);
    static std::unordered_map<std::string, std::vector<float4>> ReadClumpQuatFromCsv(
        const std::string& infilename,
        const std::string& clump_header = OUTPUT_FILE_CLUMP_TYPE_NAME,
        const std::string& qw_header = "Qw",
        const std::string& qx_header = "Qx",
        const std::string& qy_header = "Qy",
        const std::string& qz_header = "Qz");

    // ---- entities (API.h:160-310) ----
    std::shared_ptr<DEMMaterial> LoadMaterial(const std::unordered_map<std::string, float>& mat_prop);
    std::shared_ptr<DEMClumpTemplate> LoadClumpType(DEMClumpTemplate& clump);
    std::shared_ptr<DEMClumpTemplate> LoadClumpType(float mass, float3 moi, const std::vector<float>& sp_radii,
                                                    const std::vector<float3>& sp_locations_xyz,
                                                    const std::vector<std::shared_ptr<DEMMaterial>>& sp_materials);
    std::shared_ptr<DEMClumpTemplate> LoadClumpType(float mass, float3 moi, const std::vector<float>& sp_radii,
                                                    const std::vector<float3>& sp_locations_xyz,
                                                    const std::shared_ptr<DEMMaterial>& sp_material);
    std::shared_ptr<DEMClumpTemplate> LoadSphereType(float mass, float radius,
                                                     const std::shared_ptr<DEMMaterial>& material);
    std::shared_ptr<DEMClumpBatch> AddClumps(DEMClumpBatch& input_batch);
    std::shared_ptr<DEMClumpBatch> AddClumps(const std::vector<std::shared_ptr<DEMClumpTemplate>>& input_types,
                                             const std::vector<float3>& input_xyz);
    std::shared_ptr<DEMExternObj> AddExternalObject();
    std::shared_ptr<DEMExternObj> AddBCPlane(const float3 pos, const float3 normal,
                                             const std::shared_ptr<DEMMaterial>& material);

    // ---- inspectors and trackers (API.h:240-249) ----
    std::shared_ptr<DEMInspector> CreateInspector(const std::string& quantity = "clump_max_z");
    std::shared_ptr<DEMInspector> CreateInspector(const std::string& quantity, const std::string& region);
    std::shared_ptr<DEMTracker> Track(std::shared_ptr<DEMExternObj>& obj);
    std::shared_ptr<DEMTracker> Track(std::shared_ptr<DEMClumpBatch>& obj);
    bool GetInitStatus() const { return sys_initialized; }

    // ---- mid-simulation changes (API.h:120-122, 302-316) ----
    // ReadContactForceModel: the model body from a file (a path as given).
    std::shared_ptr<DEMForceModel> ReadContactForceModel(const std::string& filename);
    // ClearCache: forget the cached clump batches and external objects (trackers stay valid).
    void ClearCache();
    // UpdateClumps: add every cached clump batch to the initialised system. Its clumps take the owner numbers after
    // all existing owners (as the reference numbers them); contacts and their histories are kept. New templates,
    // materials, external objects or family prescriptions since Initialize are not allowed.
    void UpdateClumps();
    // ChangeClumpSizes: scale the component offsets and radii of the given clumps by their factors.
    void ChangeClumpSizes(const std::vector<uint32_t>& IDs, const std::vector<float>& factors);

    // ---- running (API.h:374-386) ----
    void Initialize();
    void DoDynamics(double thisCallDuration);
    void DoDynamicsThenSync(double thisCallDuration);
    void DoStepDynamics() { DoDynamics(m_ts_size); }

    // ---- state (API.h:196-214) ----
    float3 GetOwnerPosition(uint32_t ownerID);
    float3 GetOwnerVelocity(uint32_t ownerID);
    float3 GetOwnerAngVel(uint32_t ownerID);
    unsigned int GetOwnerFamily(uint32_t ownerID);
    float4 GetOwnerOriQ(uint32_t ownerID);
    float3 GetOwnerAcc(uint32_t ownerID);
    float3 GetOwnerAngAcc(uint32_t ownerID);
    void SetOwnerPosition(uint32_t ownerID, float3 pos);
    void SetOwnerVelocity(uint32_t ownerID, float3 vel);
    void SetOwnerAngVel(uint32_t ownerID, float3 angVel);
    void SetOwnerOriQ(uint32_t ownerID, float4 oriQ);

    // ---- queries and timing ----
    size_t GetNumOwners() const { return nOwnerBodies; }
    size_t GetNumClumps() const { return nOwnerClumps; }
    size_t GetNumSpheres() const { return nSpheresGM; }
    size_t GetNumContacts();
    double GetTimeElapsed();
    float GetUpdatedExpandFactor() const { return m_expand_factor; }
    double GetBinSize() const { return m_binSize; }
    const deme_sim_params& GetSimParams() const { return m_params; }
    void ShowTimingStats();
    // Raw context for callers that drive the C ABI directly (e.g. bench harnesses).
    deme_ctx* GetContext() { return m_ctx; }

  private:
    friend class DEMInspector;
    friend class DEMTracker;
    struct FamilyPrescription {
        unsigned int family;
        std::string linPosX = "none", linPosY = "none", linPosZ = "none";
        std::string linVelX = "none", linVelY = "none", linVelZ = "none";
        std::string rotVelX = "none", rotVelY = "none", rotVelZ = "none";
        bool linVelPrescribed = false, rotVelPrescribed = false, rotPosPrescribed = false, linPosPrescribed = false;
        bool used = false;
    };

    void check(int status, const char* what) const;
    // Owner numbers: the reference numbers clumps added by UpdateClumps after the external objects; the device keeps
    // all clumps first. toInternal maps the reference's number to the device's, toUser back.
    uint32_t toInternal(uint32_t u) const;
    uint32_t toUser(uint32_t i) const;
    void assignTrackers(size_t firstOwner);
    void syncHostState();
    void pushHostState();
    void addWorldBoundingBox();

    VERBOSITY verbosity = INFO;
    deme_ctx* m_ctx = nullptr;
    bool sys_initialized = false;

    // world
    float3 m_user_boxSize = make_float3(-1.f, -1.f, -1.f);
    SPATIAL_DIR m_box_dir_length_is_exact = SPATIAL_DIR::NONE;
    bool explicit_nv_override = false;
    unsigned char nvXp2 = 0, nvYp2 = 0, nvZp2 = 0;
    float l = std::numeric_limits<float>::max();
    float3 m_boxLBF = make_float3(0, 0, 0);
    std::string m_user_instructed_origin = "explicit";
    std::string m_user_add_bounding_box = "none";
    std::shared_ptr<DEMMaterial> m_bounding_box_material;
    float3 G = make_float3(0, 0, 0);
    double m_ts_size = -1.0;
    int m_updateFreq = 0;
    TIME_INTEGRATOR m_integrator = TIME_INTEGRATOR::EXTENDED_TAYLOR;
    bool use_user_defined_bin_size = false;
    double m_binSize = 0.0;
    bool use_user_defined_expand_factor = false;
    float m_expand_factor = 0.f;
    float m_approx_max_vel = -1.f;
    float m_expand_safety_param = 1.f;
    std::shared_ptr<DEMForceModel> m_force_model = std::make_shared<DEMForceModel>();

    // entities
    std::vector<std::shared_ptr<DEMMaterial>> m_loaded_materials;
    std::vector<std::shared_ptr<DEMClumpTemplate>> m_templates;
    std::vector<std::shared_ptr<DEMClumpBatch>> cached_input_clump_batches;
    std::vector<std::shared_ptr<DEMExternObj>> cached_extern_objs;
    std::vector<std::shared_ptr<DEMTrackedObj>> m_tracked_objs;
    std::vector<FamilyPrescription> m_input_family_prescription;
    struct FamilyChange {
        unsigned int from, to;
        std::string condition;
    };
    std::vector<FamilyChange> m_family_changes, m_pending_family_changes;
    std::vector<std::pair<unsigned int, unsigned int>> m_input_no_contact_pairs;
    std::vector<uint8_t> m_family_mask;
    bool m_mask_dirty = false;
    std::set<unsigned int> m_no_output_families;

    // output settings (API.h:474-479 defaults)
    OUTPUT_FORMAT m_out_format = OUTPUT_FORMAT::CSV;
    OUTPUT_FORMAT m_cnt_out_format = OUTPUT_FORMAT::CSV;
    unsigned int m_out_content = OUTPUT_CONTENT::QUAT | OUTPUT_CONTENT::ABSV;
    unsigned int m_cnt_out_content = CNT_OUTPUT_CONTENT::FORCE | CNT_OUTPUT_CONTENT::POINT;

    // derived
    deme_sim_params m_params{};
    size_t nOwnerClumps = 0, nOwnerBodies = 0, nSpheresGM = 0, nExtObj = 0, nAnalGM = 0;
    // geometry kept for the output files: component spheres (owner, radius, offset in the owner frame), the template
    // of each clump, the owner of each analytical component
    std::vector<uint32_t> m_sph_owner;
    std::vector<float> m_sph_r;
    std::vector<float3> m_sph_rel;
    std::vector<uint32_t> m_clump_template;
    std::vector<uint32_t> m_anal_owner;
    // counts at Initialize: clumps (owner numbers below it are the same on both sides), and what UpdateClumps must
    // find unchanged
    size_t m_init_clumps = 0, m_init_templates = 0, m_init_materials = 0, m_init_prescriptions = 0;
    size_t m_ext_loads = 0, m_init_ext_loads = 0;

    void writeSpheresCsv(std::string& out);
    void writeClumpsCsv(std::string& out);
    void writeContactsCsv(std::string& out);

    // host mirror of the owner state (refreshed lazily after dynamics)
    bool m_mirror_valid = false, m_mirror_dirty = false;
    std::vector<uint64_t> h_voxel;
    std::vector<uint16_t> h_lx, h_ly, h_lz;
    std::vector<float> h_qw, h_qx, h_qy, h_qz, h_vx, h_vy, h_vz, h_wx, h_wy, h_wz, h_ax, h_ay, h_az, h_alx, h_aly,
        h_alz;
    std::vector<uint8_t> h_fam;
};

}  // namespace deme
