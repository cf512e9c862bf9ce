// DEMSolver facade (see DEMSolver.h). Host-side set-up restates DEMSolver::Initialize
// (src/DEM/APIPublic.cpp:659-693, APIPrivate.cpp) and hands every per-step operation to the C ABI.
#include "DEMSolver.h"
#include "host_math.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace deme {

namespace {

using hostmath::host_rotate;
using hostmath::normalize3;

float3 f3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }

uint64_t sat_u64(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)x;
}
uint16_t sat_u16(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65535.0) return 65535;
    return (uint16_t)(uint32_t)x;
}

// hostPositionToVoxelID (HostSideHelpers.hpp:327-349) with the exact l / voxelSize
void host_pos_to_voxel(const deme_sim_params& p, double X, double Y, double Z, uint64_t& id, uint16_t& sx, uint16_t& sy,
                       uint16_t& sz) {
    const uint64_t nx = sat_u64(X / p.voxelSize_exact), ny = sat_u64(Y / p.voxelSize_exact),
                   nz = sat_u64(Z / p.voxelSize_exact);
    sx = sat_u16((X - (double)nx * p.voxelSize_exact) / p.l_exact);
    sy = sat_u16((Y - (double)ny * p.voxelSize_exact) / p.l_exact);
    sz = sat_u16((Z - (double)nz * p.voxelSize_exact) / p.l_exact);
    id = nx;
    id += ny << p.nvXp2;
    id += nz << (p.nvXp2 + p.nvYp2);
}

// hostVoxelIDToPosition (HostSideHelpers.hpp:307-324)
void host_voxel_to_pos(const deme_sim_params& p, uint64_t id, uint16_t sx, uint16_t sy, uint16_t sz, double& X,
                       double& Y, double& Z) {
    const uint64_t vx = id & (((uint64_t)1 << p.nvXp2) - 1);
    const uint64_t vy = (id >> p.nvXp2) & (((uint64_t)1 << p.nvYp2) - 1);
    const uint64_t vz = id >> (p.nvXp2 + p.nvYp2);
    X = (double)vx * p.voxelSize_exact + (double)sx * p.l_exact;
    Y = (double)vy * p.voxelSize_exact + (double)sy * p.l_exact;
    Z = (double)vz * p.voxelSize_exact + (double)sz * p.l_exact;
}

uint32_t locate_mask_pair(uint32_t i, uint32_t j) {
    if (i > j) std::swap(i, j);
    return (1 + j) * j / 2 + i;
}

void check_family(unsigned int ID, const char* what) {
    if (ID > 255) {
        std::stringstream ss;
        ss << what << ": family number " << ID << " is larger than the max allowance 255";
        throw std::runtime_error(ss.str());
    }
}

// A prescription string of the reference is a C expression in t; numeric constants go to the rule table.
bool parse_numeric(const std::string& s, double& v) {
    const char* b = s.c_str();
    char* e = nullptr;
    v = std::strtod(b, &e);
    if (e == b) return false;
    while (*e == ' ' || *e == '\t' || *e == 'f') e++;
    return *e == '\0';
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// templates / batches / objects
// ---------------------------------------------------------------------------------------------------------------
void DEMClumpTemplate::Scale(float s) {
    for (auto& p : relPos) {
        p.x *= s;
        p.y *= s;
        p.z *= s;
    }
    for (auto& r : radii) r *= s;
}
void DEMClumpTemplate::InformCentroidPrincipal(float3 center, float4 prin_Q) {
    for (auto& p : relPos) {
        host_rotate(p.x, p.y, p.z, prin_Q.w, -prin_Q.x, -prin_Q.y, -prin_Q.z);
        p.x -= center.x;
        p.y -= center.y;
        p.z -= center.z;
    }
}
void DEMClumpTemplate::Move(float3 vec, float4 rot_Q) {
    for (auto& p : relPos) {
        host_rotate(p.x, p.y, p.z, rot_Q.w, rot_Q.x, rot_Q.y, rot_Q.z);
        p.x += vec.x;
        p.y += vec.y;
        p.z += vec.z;
    }
}

DEMClumpBatch::DEMClumpBatch(size_t num) : nClumps(num) {
    types.resize(num);
    families.resize(num, DEFAULT_CLUMP_FAMILY_NUM);
    vel.resize(num, make_float3(0, 0, 0));
    angVel.resize(num, make_float3(0, 0, 0));
    xyz.resize(num);
    oriQ.resize(num, make_float4(0, 0, 0, 1));
}
void DEMClumpBatch::assertLength(size_t len, const std::string& name) const {
    if (len != nClumps) {
        std::stringstream ss;
        ss << name << " input argument must have length " << nClumps << " (not " << len
           << "), same as the number of clumps you originally added via AddClumps.";
        throw std::runtime_error(ss.str());
    }
}
void DEMClumpBatch::SetTypes(const std::vector<std::shared_ptr<DEMClumpTemplate>>& input) {
    assertLength(input.size(), "SetTypes");
    types = input;
}
void DEMClumpBatch::SetTypes(const std::shared_ptr<DEMClumpTemplate>& input) {
    SetTypes(std::vector<std::shared_ptr<DEMClumpTemplate>>(nClumps, input));
}
void DEMClumpBatch::SetPos(const std::vector<float3>& input) {
    assertLength(input.size(), "SetPos");
    xyz = input;
}
void DEMClumpBatch::SetVel(const std::vector<float3>& input) {
    assertLength(input.size(), "SetVel");
    vel = input;
}
void DEMClumpBatch::SetAngVel(const std::vector<float3>& input) {
    assertLength(input.size(), "SetAngVel");
    angVel = input;
}
void DEMClumpBatch::SetOriQ(const std::vector<float4>& input) {
    assertLength(input.size(), "SetOriQ");
    oriQ = input;
}
void DEMClumpBatch::SetFamilies(const std::vector<unsigned int>& input) {
    assertLength(input.size(), "SetFamilies");
    for (unsigned int f : input) check_family(f, "SetFamilies");
    families = input;
    family_isSpecified = true;
}

void DEMExternObj::SetFamily(unsigned int code) {
    check_family(code, "DEMExternObj::SetFamily");
    family_code = code;
}
void DEMExternObj::AddPlane(const float3 pos, const float3 normal, const std::shared_ptr<DEMMaterial>& material) {
    planes.push_back({pos, normalize3(normal), material});
}

// ---------------------------------------------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------------------------------------------
DEMSolver::DEMSolver(unsigned int) { m_family_mask.assign(DEME_NUM_MASK_ENTRIES, DEME_DONT_PREVENT_CONTACT); }

DEMSolver::~DEMSolver() {
    if (m_ctx) deme_ctx_destroy(m_ctx);
}

void DEMSolver::check(int status, const char* what) const {
    if (status != DEME_OK) {
        std::stringstream ss;
        ss << what << " failed (status " << status << ")";
        if (m_ctx) ss << ": " << deme_last_error(m_ctx);
        throw std::runtime_error(ss.str());
    }
}

void DEMSolver::InstructBoxDomainDimension(float x, float y, float z, SPATIAL_DIR dir_exact) {
    m_user_boxSize = make_float3(x, y, z);
    m_box_dir_length_is_exact = dir_exact;
    explicit_nv_override = false;
}

void DEMSolver::InstructBoxDomainNumVoxel(unsigned char x, unsigned char y, unsigned char z, float len_unit) {
    if (x + y + z != 64) throw std::runtime_error("Voxel numbers (as powers of 2) must add up to 64.");
    l = len_unit;
    nvXp2 = x;
    nvYp2 = y;
    nvZp2 = z;
    const double vs = (double)((size_t)1 << 16) * (double)l;
    m_user_boxSize = make_float3((float)(vs * (double)((size_t)1 << x)), (float)(vs * (double)((size_t)1 << y)),
                                 (float)(vs * (double)((size_t)1 << z)));
    explicit_nv_override = true;
}

void DEMForceModel::SetForceModelType(FORCE_MODEL t) {
    type = t;
    model.clear();
    if (t == FORCE_MODEL::HERTZIAN)
        contact_wildcards = {"delta_tan_x", "delta_tan_y", "delta_tan_z", "delta_time"};
    else if (t == FORCE_MODEL::HERTZIAN_FRICTIONLESS)
        contact_wildcards.clear();
}
void DEMForceModel::DefineCustomModel(const std::string& m) {
    type = FORCE_MODEL::CUSTOM;
    model = m;
}
void DEMForceModel::SetPerContactWildcards(const std::set<std::string>& names) {
    for (const auto& n : names)
        if (n.find(' ') != std::string::npos)
            throw std::runtime_error("Contact wildcard " + n + " is not valid: no spaces allowed in its name.");
    contact_wildcards = names;
}

std::shared_ptr<DEMForceModel> DEMSolver::UseFrictionalHertzianModel() {
    m_force_model->SetForceModelType(FORCE_MODEL::HERTZIAN);
    return m_force_model;
}
std::shared_ptr<DEMForceModel> DEMSolver::UseFrictionlessHertzianModel() {
    m_force_model->SetForceModelType(FORCE_MODEL::HERTZIAN_FRICTIONLESS);
    return m_force_model;
}
std::shared_ptr<DEMForceModel> DEMSolver::DefineContactForceModel(const std::string& model) {
    m_force_model->DefineCustomModel(model);
    return m_force_model;
}

void DEMSolver::ChangeFamilyWhen(unsigned int ID_from, unsigned int ID_to, const std::string& condition) {
    check_family(ID_from, "ChangeFamilyWhen");
    check_family(ID_to, "ChangeFamilyWhen");
    if (sys_initialized) throw std::runtime_error("ChangeFamilyWhen must be called before Initialize on this path.");
    m_family_changes.push_back({ID_from, ID_to, condition});
}

void DEMSolver::ChangeFamily(unsigned int ID_from, unsigned int ID_to) {
    check_family(ID_from, "ChangeFamily");
    check_family(ID_to, "ChangeFamily");
    if (!sys_initialized) {
        m_pending_family_changes.push_back({ID_from, ID_to, ""});  // applied when Initialize assigns families
        return;
    }
    check(deme_change_family(m_ctx, (int)ID_from, (int)ID_to), "deme_change_family");
    m_mirror_valid = false;
}

void DEMSolver::DisableFamilyOutput(unsigned int ID) {
    check_family(ID, "DisableFamilyOutput");
    m_no_output_families.insert(ID);
}

void DEMSolver::SetFamilyFixed(unsigned int ID) {
    check_family(ID, "SetFamilyFixed");
    FamilyPrescription p;
    p.family = ID;
    p.linVelX = p.linVelY = p.linVelZ = "0";
    p.rotVelX = p.rotVelY = p.rotVelZ = "0";
    p.linVelPrescribed = p.rotVelPrescribed = p.rotPosPrescribed = p.linPosPrescribed = true;
    p.used = true;
    m_input_family_prescription.push_back(p);
}

void DEMSolver::SetFamilyPrescribedLinVel(unsigned int ID, const std::string& vx, const std::string& vy,
                                          const std::string& vz, bool dictate) {
    check_family(ID, "SetFamilyPrescribedLinVel");
    FamilyPrescription p;
    p.family = ID;
    p.linVelX = vx;
    p.linVelY = vy;
    p.linVelZ = vz;
    p.linVelPrescribed = dictate;
    p.rotVelPrescribed = dictate;
    p.used = true;
    m_input_family_prescription.push_back(p);
}

void DEMSolver::SetFamilyPrescribedAngVel(unsigned int ID, const std::string& wx, const std::string& wy,
                                          const std::string& wz, bool dictate) {
    check_family(ID, "SetFamilyPrescribedAngVel");
    FamilyPrescription p;
    p.family = ID;
    p.rotVelX = wx;
    p.rotVelY = wy;
    p.rotVelZ = wz;
    p.linVelPrescribed = dictate;
    p.rotVelPrescribed = dictate;
    p.used = true;
    m_input_family_prescription.push_back(p);
}

void DEMSolver::SetFamilyPrescribedPosition(unsigned int ID, const std::string& X, const std::string& Y,
                                            const std::string& Z) {
    check_family(ID, "SetFamilyPrescribedPosition");
    FamilyPrescription p;
    p.family = ID;
    p.linPosX = X;
    p.linPosY = Y;
    p.linPosZ = Z;
    p.rotPosPrescribed = true;
    p.linPosPrescribed = true;
    p.used = true;
    m_input_family_prescription.push_back(p);
}

void DEMSolver::DisableContactBetweenFamilies(unsigned int ID1, unsigned int ID2) {
    check_family(ID1, "DisableContactBetweenFamilies");
    check_family(ID2, "DisableContactBetweenFamilies");
    m_family_mask[locate_mask_pair(ID1, ID2)] = DEME_PREVENT_CONTACT;
    m_mask_dirty = sys_initialized;
    m_input_no_contact_pairs.push_back({ID1, ID2});
}

void DEMSolver::EnableContactBetweenFamilies(unsigned int ID1, unsigned int ID2) {
    check_family(ID1, "EnableContactBetweenFamilies");
    check_family(ID2, "EnableContactBetweenFamilies");
    if (!sys_initialized)
        throw std::runtime_error("There is no need to call EnableContactBetweenFamilies before system initialization.");
    m_family_mask[locate_mask_pair(ID1, ID2)] = DEME_DONT_PREVENT_CONTACT;
    m_mask_dirty = true;
}

std::shared_ptr<DEMMaterial> DEMSolver::LoadMaterial(const std::unordered_map<std::string, float>& mat_prop) {
    for (const auto& kv : mat_prop)
        if (kv.first.find(' ') != std::string::npos)
            throw std::runtime_error("Material property " + kv.first + " is not valid: no spaces allowed in its name.");
    auto ptr = std::make_shared<DEMMaterial>(mat_prop);
    ptr->load_order = (unsigned int)m_loaded_materials.size();
    m_loaded_materials.push_back(ptr);
    return ptr;
}

std::shared_ptr<DEMClumpTemplate> DEMSolver::LoadClumpType(DEMClumpTemplate& clump) {
    if (clump.nComp != clump.radii.size() || clump.nComp != clump.relPos.size() ||
        clump.nComp != clump.materials.size())
        throw std::runtime_error(
            "Radii, relative positions and material arrays defining a clump topology, must all have the same length.");
    clump.mark = (unsigned int)m_templates.size();
    if (clump.m_name == "NULL") {
        char name[32];
        std::snprintf(name, sizeof(name), "%04u", clump.mark);
        clump.m_name = name;
    }
    m_templates.push_back(std::make_shared<DEMClumpTemplate>(clump));
    return m_templates.back();
}

std::shared_ptr<DEMClumpTemplate> DEMSolver::LoadClumpType(float mass, float3 moi, const std::vector<float>& r,
                                                           const std::vector<float3>& pos,
                                                           const std::vector<std::shared_ptr<DEMMaterial>>& mats) {
    DEMClumpTemplate c;
    c.mass = mass;
    c.MOI = moi;
    c.radii = r;
    c.relPos = pos;
    c.materials = mats;
    c.nComp = (unsigned int)r.size();
    return LoadClumpType(c);
}

std::shared_ptr<DEMClumpTemplate> DEMSolver::LoadClumpType(float mass, float3 moi, const std::vector<float>& r,
                                                           const std::vector<float3>& pos,
                                                           const std::shared_ptr<DEMMaterial>& mat) {
    return LoadClumpType(mass, moi, r, pos, std::vector<std::shared_ptr<DEMMaterial>>(r.size(), mat));
}

std::shared_ptr<DEMClumpTemplate> DEMSolver::LoadSphereType(float mass, float radius,
                                                            const std::shared_ptr<DEMMaterial>& material) {
    const float I = (float)(2.0 / 5.0 * mass * radius * radius);  // APIPublic.cpp:402-409
    return LoadClumpType(mass, make_float3(I, I, I), std::vector<float>(1, radius),
                         std::vector<float3>(1, make_float3(0, 0, 0)),
                         std::vector<std::shared_ptr<DEMMaterial>>(1, material));
}

std::shared_ptr<DEMClumpBatch> DEMSolver::AddClumps(DEMClumpBatch& input_batch) {
    input_batch.load_order = cached_input_clump_batches.size();
    cached_input_clump_batches.push_back(std::make_shared<DEMClumpBatch>(input_batch));
    return cached_input_clump_batches.back();
}

std::shared_ptr<DEMClumpBatch> DEMSolver::AddClumps(const std::vector<std::shared_ptr<DEMClumpTemplate>>& types,
                                                    const std::vector<float3>& xyz) {
    if (types.size() != xyz.size()) throw std::runtime_error("Arrays in the call AddClumps must all have the same length.");
    DEMClumpBatch b(types.size());
    b.SetTypes(types);
    b.SetPos(xyz);
    return AddClumps(b);
}

std::shared_ptr<DEMExternObj> DEMSolver::AddExternalObject() {
    m_ext_loads++;
    auto p = std::make_shared<DEMExternObj>();
    p->load_order = (unsigned int)cached_extern_objs.size();
    cached_extern_objs.push_back(p);
    return p;
}

std::shared_ptr<DEMExternObj> DEMSolver::AddBCPlane(const float3 pos, const float3 normal,
                                                    const std::shared_ptr<DEMMaterial>& material) {
    auto p = AddExternalObject();
    p->AddPlane(pos, normal, material);
    return p;
}

// addWorldBoundingBox (APIPrivate.cpp:660-687)
void DEMSolver::addWorldBoundingBox() {
    if (m_user_add_bounding_box == "none") return;
    if (!m_bounding_box_material) throw std::runtime_error("InstructBoxDomainBoundingBC needs a material.");
    const float3 L = m_boxLBF, S = m_user_boxSize;
    auto mat = m_bounding_box_material;
    auto box = AddExternalObject();
    box->AddPlane(make_float3(L.x + S.x / 2., L.y + S.y / 2., L.z), make_float3(0, 0, 1), mat);
    if (m_user_add_bounding_box == "only_bottom") return;
    box->AddPlane(make_float3(L.x, L.y + S.y / 2., L.z + S.z / 2.), make_float3(1, 0, 0), mat);
    box->AddPlane(make_float3(L.x + S.x, L.y + S.y / 2., L.z + S.z / 2.), make_float3(-1, 0, 0), mat);
    box->AddPlane(make_float3(L.x + S.x / 2., L.y, L.z + S.z / 2.), make_float3(0, 1, 0), mat);
    box->AddPlane(make_float3(L.x + S.x / 2., L.y + S.y, L.z + S.z / 2.), make_float3(0, -1, 0), mat);
    if (m_user_add_bounding_box == "top_open") return;
    box->AddPlane(make_float3(L.x + S.x / 2., L.y + S.y / 2., L.z + S.z), make_float3(0, 0, -1), mat);
}

void DEMSolver::Initialize() {
    if (sys_initialized) throw std::runtime_error("Re-initialization is not supported on this path.");
    // validateUserInputs (APIPrivate.cpp:878-907)
    if (m_user_boxSize.x <= 0.f || m_user_boxSize.y <= 0.f || m_user_boxSize.z <= 0.f)
        throw std::runtime_error("The size of the simulation world is set to be (or default to be) impossibly small.");
    if (!(m_ts_size > 0.0)) throw std::runtime_error("A positive time step size is required (SetInitTimeStep).");
    if (m_integrator == TIME_INTEGRATOR::CHUNG)
        throw std::runtime_error("The integration type is unknown or not implemented. Please select another via SetIntegrator.");
    SetFamilyFixed(RESERVED_FAMILY_NUM);

    // figureOutOrigin (APIPrivate.cpp:647-658)
    if (m_user_instructed_origin == "center") {
        m_boxLBF = make_float3(-m_user_boxSize.x / 2.0f, -m_user_boxSize.y / 2.0f, -m_user_boxSize.z / 2.0f);
    } else if (m_user_instructed_origin != "explicit") {
        throw std::runtime_error("Unrecognized location of system origin.");
    }
    addWorldBoundingBox();

    // smallest radius (decideBinSize, APIPrivate.cpp:335-344)
    float minR = std::numeric_limits<float>::max();
    for (const auto& t : m_templates)
        for (float r : t->radii) minR = std::min(minR, r);

    // expand factor (transferSimParams, APIPrivate.cpp:730-737)
    if (!use_user_defined_expand_factor)
        m_expand_factor = (float)(m_approx_max_vel * m_ts_size * m_updateFreq * m_expand_safety_param);
    if (m_expand_factor < 0.f) m_expand_factor = 0.f;

    int exact = -1;
    if (m_box_dir_length_is_exact == SPATIAL_DIR::X) exact = 0;
    if (m_box_dir_length_is_exact == SPATIAL_DIR::Y) exact = 1;
    if (m_box_dir_length_is_exact == SPATIAL_DIR::Z) exact = 2;
    const float LBF[3] = {m_boxLBF.x, m_boxLBF.y, m_boxLBF.z};
    const float g[3] = {G.x, G.y, G.z};
    int integ = m_integrator == TIME_INTEGRATOR::FORWARD_EULER         ? DEME_INTEGRATOR_FORWARD_EULER
                : m_integrator == TIME_INTEGRATOR::CENTERED_DIFFERENCE ? DEME_INTEGRATOR_CENTERED_DIFFERENCE
                                                                        : DEME_INTEGRATOR_EXTENDED_TAYLOR;
    uint8_t nv[3] = {nvXp2, nvYp2, nvZp2};
    int st = deme_build_sim_params(m_user_boxSize.x, m_user_boxSize.y, m_user_boxSize.z, exact,
                                   explicit_nv_override ? nv : nullptr, l, LBF,
                                   minR == std::numeric_limits<float>::max() ? 0.f : minR,
                                   use_user_defined_bin_size ? m_binSize : 0.0, g, (float)m_ts_size, m_expand_factor,
                                   integ, &m_params);
    if (st != DEME_OK) {
        if (st == DEME_ERR_INVALID_ARGUMENT)
            throw std::runtime_error("There are spheres with 0 radii, and the user did not specify the bin size.");
        throw std::runtime_error("The number of bins exceeds binID_t; please use a larger bin size.");
    }
    m_binSize = m_params.binSize;

    // entities: clumps first, then external objects (dT.cpp:436-620)
    std::vector<uint32_t> sphOwner;
    std::vector<float> sphR, sphRel;
    std::vector<uint8_t> sphMat;
    std::vector<float> mass, moi;
    size_t nClumps = 0;
    for (const auto& b : cached_input_clump_batches) nClumps += b->GetNumClumps();
    nOwnerClumps = nClumps;
    nExtObj = cached_extern_objs.size();
    nOwnerBodies = nClumps + nExtObj;
    const size_t nO = nOwnerBodies;
    h_voxel.assign(nO, 0);
    h_lx.assign(nO, 0);
    h_ly.assign(nO, 0);
    h_lz.assign(nO, 0);
    h_qw.assign(nO, 1.f);
    for (auto* v : {&h_qx, &h_qy, &h_qz, &h_vx, &h_vy, &h_vz, &h_wx, &h_wy, &h_wz, &h_ax, &h_ay, &h_az, &h_alx, &h_aly,
                    &h_alz})
        v->assign(nO, 0.f);
    h_fam.assign(nO, 0);
    m_clump_template.assign(nClumps, 0);
    mass.resize(nO);
    moi.resize(3 * nO);
    size_t o = 0;
    for (const auto& b : cached_input_clump_batches) {
        for (size_t i = 0; i < b->GetNumClumps(); i++, o++) {
            const auto& t = b->types.at(i);
            if (!t) throw std::runtime_error("A clump in an AddClumps batch has no template.");
            m_clump_template[o] = t->mark;
            mass[o] = t->mass;
            moi[3 * o] = t->MOI.x;
            moi[3 * o + 1] = t->MOI.y;
            moi[3 * o + 2] = t->MOI.z;
            for (unsigned int j = 0; j < t->nComp; j++) {
                sphOwner.push_back((uint32_t)o);
                sphR.push_back(t->radii[j]);
                sphRel.push_back(t->relPos[j].x);
                sphRel.push_back(t->relPos[j].y);
                sphRel.push_back(t->relPos[j].z);
                sphMat.push_back((uint8_t)t->materials[j]->load_order);
            }
            const float3 c = f3sub(b->xyz[i], m_boxLBF);
            host_pos_to_voxel(m_params, (double)c.x, (double)c.y, (double)c.z, h_voxel[o], h_lx[o], h_ly[o], h_lz[o]);
            h_qw[o] = b->oriQ[i].w;
            h_qx[o] = b->oriQ[i].x;
            h_qy[o] = b->oriQ[i].y;
            h_qz[o] = b->oriQ[i].z;
            h_vx[o] = b->vel[i].x;
            h_vy[o] = b->vel[i].y;
            h_vz[o] = b->vel[i].z;
            h_wx[o] = b->angVel[i].x;
            h_wy[o] = b->angVel[i].y;
            h_wz[o] = b->angVel[i].z;
            h_fam[o] = (uint8_t)b->families[i];
        }
    }
    nSpheresGM = sphOwner.size();
    m_sph_owner = sphOwner;
    m_sph_r = sphR;
    m_sph_rel.resize(nSpheresGM);
    for (size_t i = 0; i < nSpheresGM; i++)
        m_sph_rel[i] = make_float3(sphRel[3 * i], sphRel[3 * i + 1], sphRel[3 * i + 2]);
    std::vector<deme_anal_obj> anal;
    for (size_t k = 0; k < nExtObj; k++, o++) {
        const auto& e = cached_extern_objs[k];
        mass[o] = e->mass;
        moi[3 * o] = e->MOI.x;
        moi[3 * o + 1] = e->MOI.y;
        moi[3 * o + 2] = e->MOI.z;
        const float3 c = f3sub(e->init_pos, m_boxLBF);
        host_pos_to_voxel(m_params, (double)c.x, (double)c.y, (double)c.z, h_voxel[o], h_lx[o], h_ly[o], h_lz[o]);
        h_fam[o] = (uint8_t)e->family_code;
        for (const auto& pl : e->planes) {
            deme_anal_obj a{};
            a.owner = (uint32_t)o;
            a.type = DEME_ANAL_OBJ_TYPE_PLANE;
            a.material = (uint8_t)pl.material->load_order;
            a.normal = 0;
            a.relPos[0] = pl.position.x;
            a.relPos[1] = pl.position.y;
            a.relPos[2] = pl.position.z;
            a.rot[0] = pl.normal.x;
            a.rot[1] = pl.normal.y;
            a.rot[2] = pl.normal.z;
            a.mass = e->mass;
            anal.push_back(a);
        }
    }
    nAnalGM = anal.size();
    m_anal_owner.resize(nAnalGM);
    for (size_t a = 0; a < nAnalGM; a++) m_anal_owner[a] = anal[a].owner;

    // materials (equipMaterials, APIPrivate.cpp:1089-1140): missing properties default to 0
    std::vector<deme_material> mats;
    for (const auto& m : m_loaded_materials) {
        auto get = [&](const char* k) {
            auto it = m->mat_prop.find(k);
            return it == m->mat_prop.end() ? 0.f : it->second;
        };
        mats.push_back(deme_material{get("E"), get("nu"), get("CoR"), get("mu"), get("Crr")});
    }
    if (mats.empty()) mats.push_back(deme_material{0, 0, 0, 0, 0});

    for (const auto& fc : m_pending_family_changes)
        for (auto& f : h_fam)
            if (f == fc.from) f = (uint8_t)fc.to;
    // family prescriptions (figureOutFamilyMasks APIPrivate.cpp:574-645, equipFamilyPrescribedMotions :1005-1048)
    std::vector<FamilyPrescription> merged(256);
    for (unsigned int f = 0; f < 256; f++) merged[f].family = f;
    for (const auto& p : m_input_family_prescription) {
        auto& m = merged[p.family];
        m.used = true;
        auto take = [](std::string& dst, const std::string& src) {
            if (src != "none") dst = src;
        };
        take(m.linPosX, p.linPosX);
        take(m.linPosY, p.linPosY);
        take(m.linPosZ, p.linPosZ);
        take(m.linVelX, p.linVelX);
        take(m.linVelY, p.linVelY);
        take(m.linVelZ, p.linVelZ);
        take(m.rotVelX, p.rotVelX);
        take(m.rotVelY, p.rotVelY);
        take(m.rotVelZ, p.rotVelZ);
        m.linVelPrescribed = m.linVelPrescribed || p.linVelPrescribed;
        m.rotVelPrescribed = m.rotVelPrescribed || p.rotVelPrescribed;
        m.rotPosPrescribed = m.rotPosPrescribed || p.rotPosPrescribed;
        m.linPosPrescribed = m.linPosPrescribed || p.linPosPrescribed;
    }
    std::vector<deme_family_rule> rules(256);
    // families whose prescriptions are expressions (in t) go to the owner update kernel's strategies, compiled by
    // hipRTC as the reference's equipFamilyPrescribedMotions strings are; numeric ones stay in the rule table
    std::string velStrategy, posStrategy;
    for (unsigned int f = 0; f < 256; f++) {
        const auto& m = merged[f];
        deme_family_rule& r = rules[f];
        std::memset(&r, 0, sizeof(r));
        if (!m.used) continue;
        const std::string* all[9] = {&m.linVelX, &m.linVelY, &m.linVelZ, &m.rotVelX, &m.rotVelY,
                                     &m.rotVelZ, &m.linPosX, &m.linPosY, &m.linPosZ};
        bool expr = false;
        for (const std::string* e : all) {
            double v;
            if (*e != "none" && !parse_numeric(*e, v)) expr = true;
        }
        if (expr) {
            const char* vn[6] = {"vX", "vY", "vZ", "omgBarX", "omgBarY", "omgBarZ"};
            const char* pn[3] = {"X", "Y", "Z"};
            velStrategy += "case " + std::to_string(f) + ": {";
            for (int k = 0; k < 6; k++)
                if (*all[k] != "none") velStrategy += std::string(vn[k]) + " = " + *all[k] + ";";
            velStrategy += "LinPrescribed = " + std::to_string((int)m.linVelPrescribed) + "; RotPrescribed = " +
                           std::to_string((int)m.rotVelPrescribed) + "; break; }";
            posStrategy += "case " + std::to_string(f) + ": {";
            for (int k = 0; k < 3; k++)
                if (*all[6 + k] != "none") posStrategy += std::string(pn[k]) + " = " + *all[6 + k] + ";";
            posStrategy += "LinPrescribed = " + std::to_string((int)m.linPosPrescribed) + "; RotPrescribed = " +
                           std::to_string((int)m.rotPosPrescribed) + "; break; }";
            continue;
        }
        r.used = 1;
        r.linVelPrescribed = m.linVelPrescribed;
        r.rotVelPrescribed = m.rotVelPrescribed;
        r.linPosPrescribed = m.linPosPrescribed;
        r.rotPosPrescribed = m.rotPosPrescribed;
        const std::string* vels[6] = {&m.linVelX, &m.linVelY, &m.linVelZ, &m.rotVelX, &m.rotVelY, &m.rotVelZ};
        for (int k = 0; k < 6; k++) {
            if (*vels[k] == "none") continue;
            double v;
            if (!parse_numeric(*vels[k], v))
                throw std::runtime_error("Prescribed velocity expression '" + *vels[k] +
                                         "' is not a numeric constant; time-dependent expressions need runtime "
                                         "specialisation, which this path does not provide.");
            r.vel[k] = (float)v;
            r.velSet |= (uint8_t)(1u << k);
        }
        const std::string* pos[3] = {&m.linPosX, &m.linPosY, &m.linPosZ};
        for (int k = 0; k < 3; k++) {
            if (*pos[k] == "none") continue;
            double v;
            if (!parse_numeric(*pos[k], v))
                throw std::runtime_error("Prescribed position expression '" + *pos[k] + "' is not a numeric constant.");
            r.pos[k] = v;
            r.posSet |= (uint8_t)(1u << k);
        }
    }

    // device context
    int dev = 0;
    check(deme_ctx_create(dev, nullptr, &m_ctx), "deme_ctx_create");
    check(deme_set_params(m_ctx, &m_params), "deme_set_params");
    check(deme_upload_templates(m_ctx, (uint32_t)nSpheresGM, sphOwner.data(), sphR.data(), sphRel.data(), sphMat.data(),
                                (uint32_t)nO, (uint32_t)nClumps, mass.data(), moi.data()),
          "deme_upload_templates");
    check(deme_upload_materials(m_ctx, (uint32_t)mats.size(), mats.data()), "deme_upload_materials");
    check(deme_upload_analytical(m_ctx, (uint32_t)anal.size(), anal.data()), "deme_upload_analytical");
    check(deme_upload_families(m_ctx, m_family_mask.data(), rules.data()), "deme_upload_families");
    // the contact file reports the last step's per-contact forces and points, as the reference keeps them
    check(deme_set_keep_contact_records(m_ctx, 1), "deme_set_keep_contact_records");
    // clump volumes for the clump_volume inspector (the template volumes, AuxClasses.cpp:51-55)
    {
        std::vector<float> vol(nClumps);
        for (size_t c = 0; c < nClumps; c++) vol[c] = m_templates.at(m_clump_template[c])->volume;
        check(deme_set_owner_volumes(m_ctx, vol.data()), "deme_set_owner_volumes");
    }
    m_init_clumps = nClumps;
    m_init_templates = m_templates.size();
    m_init_materials = m_loaded_materials.size();
    m_init_prescriptions = m_input_family_prescription.size();
    m_init_ext_loads = m_ext_loads;
    assignTrackers(0);
    // motion strategies and family change rules (equipFamilyOnFlyChanges, APIPrivate.cpp:980-1003)
    std::string changeRules;
    for (const auto& fc : m_family_changes) {
        std::string cond = fc.condition;
        for (size_t p = cond.find("return"); p != std::string::npos; p = cond.find("return", p))
            cond.replace(p, 6, "shouldMakeChange = ");
        changeRules += "if (family_code == " + std::to_string(fc.from) + ") { bool shouldMakeChange = false; " + cond +
                       " if (shouldMakeChange) newFamily = " + std::to_string(fc.to) + "; }";
    }
    if (!velStrategy.empty() || !changeRules.empty()) {
        const char* keys[3] = {"_velPrescriptionStrategy_", "_posPrescriptionStrategy_", "_familyChangeRules_"};
        const char* vals[3] = {velStrategy.c_str(), posStrategy.c_str(), changeRules.c_str()};
        check(deme_jit_build(m_ctx, keys, vals, 3), "deme_jit_build (motion rules)");
    }
    if (m_force_model->type != FORCE_MODEL::HERTZIAN) {
        // other models are compiled now by hipRTC (the built-in frictional model ships precompiled)
        const std::string body = m_force_model->type == FORCE_MODEL::HERTZIAN_FRICTIONLESS
                                     ? std::string(deme_builtin_force_model(1))
                                     : m_force_model->model;
        std::vector<const char*> names;
        for (const auto& n : m_force_model->contact_wildcards) names.push_back(n.c_str());
        check(deme_set_force_model(m_ctx, body.c_str(), names.empty() ? nullptr : names.data(), (int)names.size()),
              "deme_set_force_model");
    }
    m_mirror_valid = true;
    m_mirror_dirty = true;
    pushHostState();
    sys_initialized = true;
}

void DEMSolver::pushHostState() {
    if (!m_mirror_dirty) return;
    deme_owner_state s{};
    s.voxelID = h_voxel.data();
    s.locX = h_lx.data();
    s.locY = h_ly.data();
    s.locZ = h_lz.data();
    s.oriQw = h_qw.data();
    s.oriQx = h_qx.data();
    s.oriQy = h_qy.data();
    s.oriQz = h_qz.data();
    s.vX = h_vx.data();
    s.vY = h_vy.data();
    s.vZ = h_vz.data();
    s.omgX = h_wx.data();
    s.omgY = h_wy.data();
    s.omgZ = h_wz.data();
    s.family = h_fam.data();
    check(deme_upload_state(m_ctx, &s), "deme_upload_state");
    m_mirror_dirty = false;
}

void DEMSolver::syncHostState() {
    if (!sys_initialized) throw std::runtime_error("The system is not initialized.");
    if (m_mirror_valid) return;
    deme_owner_state s{};
    s.voxelID = h_voxel.data();
    s.locX = h_lx.data();
    s.locY = h_ly.data();
    s.locZ = h_lz.data();
    s.oriQw = h_qw.data();
    s.oriQx = h_qx.data();
    s.oriQy = h_qy.data();
    s.oriQz = h_qz.data();
    s.vX = h_vx.data();
    s.vY = h_vy.data();
    s.vZ = h_vz.data();
    s.omgX = h_wx.data();
    s.omgY = h_wy.data();
    s.omgZ = h_wz.data();
    s.aX = h_ax.data();
    s.aY = h_ay.data();
    s.aZ = h_az.data();
    s.alphaX = h_alx.data();
    s.alphaY = h_aly.data();
    s.alphaZ = h_alz.data();
    s.family = h_fam.data();
    check(deme_download_state(m_ctx, &s), "deme_download_state");
    m_mirror_valid = true;
}

void DEMSolver::DoDynamics(double duration) {
    if (!sys_initialized) throw std::runtime_error("Please call Initialize before DoDynamics.");
    pushHostState();
    // same step count as the dT loop `for (cycle = 0; cycle < duration; cycle += h)` (dT.cpp:1527)
    const float h = m_params.h;
    int n = 0;
    for (double cyc = 0.0; cyc < duration; cyc += (double)h) n++;
    if (m_mask_dirty) {
        check(deme_upload_families(m_ctx, m_family_mask.data(), nullptr), "deme_upload_families");
        m_mask_dirty = false;
    }
    check(deme_step(m_ctx, n, m_updateFreq), "deme_step");
    m_mirror_valid = false;
}

void DEMSolver::DoDynamicsThenSync(double duration) {
    DoDynamics(duration);
    check(deme_sync(m_ctx), "deme_sync");
}

float3 DEMSolver::GetOwnerPosition(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    double X, Y, Z;
    host_voxel_to_pos(m_params, h_voxel.at(i), h_lx.at(i), h_ly.at(i), h_lz.at(i), X, Y, Z);
    return make_float3((float)(X + m_params.LBF[0]), (float)(Y + m_params.LBF[1]), (float)(Z + m_params.LBF[2]));
}
float3 DEMSolver::GetOwnerVelocity(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    return make_float3(h_vx.at(i), h_vy.at(i), h_vz.at(i));
}
unsigned int DEMSolver::GetOwnerFamily(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    return h_fam.at(i);
}
float3 DEMSolver::GetOwnerAngVel(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    return make_float3(h_wx.at(i), h_wy.at(i), h_wz.at(i));
}
float4 DEMSolver::GetOwnerOriQ(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    float4 q;
    q.w = h_qw.at(i);
    q.x = h_qx.at(i);
    q.y = h_qy.at(i);
    q.z = h_qz.at(i);
    return q;
}
float3 DEMSolver::GetOwnerAcc(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    return make_float3(h_ax.at(i), h_ay.at(i), h_az.at(i));
}
float3 DEMSolver::GetOwnerAngAcc(uint32_t i) {
    syncHostState();
    i = toInternal(i);
    return make_float3(h_alx.at(i), h_aly.at(i), h_alz.at(i));
}
void DEMSolver::SetOwnerPosition(uint32_t i, float3 pos) {
    syncHostState();
    i = toInternal(i);
    // float subtraction first, as dT::setOwnerPos does (dT.cpp:1783-1790)
    const double X = pos.x - m_params.LBF[0], Y = pos.y - m_params.LBF[1], Z = pos.z - m_params.LBF[2];
    host_pos_to_voxel(m_params, X, Y, Z, h_voxel.at(i), h_lx.at(i), h_ly.at(i), h_lz.at(i));
    m_mirror_dirty = true;
}
void DEMSolver::SetOwnerVelocity(uint32_t i, float3 v) {
    syncHostState();
    i = toInternal(i);
    h_vx.at(i) = v.x;
    h_vy.at(i) = v.y;
    h_vz.at(i) = v.z;
    m_mirror_dirty = true;
}
void DEMSolver::SetOwnerAngVel(uint32_t i, float3 w) {
    syncHostState();
    i = toInternal(i);
    h_wx.at(i) = w.x;
    h_wy.at(i) = w.y;
    h_wz.at(i) = w.z;
    m_mirror_dirty = true;
}
void DEMSolver::SetOwnerOriQ(uint32_t i, float4 q) {
    syncHostState();
    i = toInternal(i);
    h_qw.at(i) = q.w;
    h_qx.at(i) = q.x;
    h_qy.at(i) = q.y;
    h_qz.at(i) = q.z;
    m_mirror_dirty = true;
}

// ---------------------------------------------------------------------------------------------------------------
// owner numbering, mid-simulation changes
// ---------------------------------------------------------------------------------------------------------------
uint32_t DEMSolver::toInternal(uint32_t u) const {
    if (u >= nOwnerBodies) throw std::out_of_range("owner " + std::to_string(u) + " does not exist");
    if (u < m_init_clumps) return u;
    if (u < m_init_clumps + nExtObj) return (uint32_t)(nOwnerClumps + (u - m_init_clumps));
    return (uint32_t)(u - nExtObj);
}

uint32_t DEMSolver::toUser(uint32_t i) const {
    if (i < m_init_clumps) return i;
    if (i < nOwnerClumps) return (uint32_t)(i + nExtObj);
    return (uint32_t)(m_init_clumps + (i - nOwnerClumps));
}

// Owners of trackers not yet placed (dT.cpp:715-741): a batch's first owner follows firstOwner (the owners that
// existed before its clumps were added) by the clumps of the cached batches before it; an external object follows
// the clumps of the Initialize call by its load order.
void DEMSolver::assignTrackers(size_t firstOwner) {
    std::vector<size_t> batchStart(1, 0);
    for (const auto& b : cached_input_clump_batches) batchStart.push_back(batchStart.back() + b->GetNumClumps());
    for (auto& t : m_tracked_objs) {
        if (t->ownerID != std::numeric_limits<uint32_t>::max()) continue;
        if (t->kind == DEMTrackedObj::Kind::CLUMP_BATCH) {
            if (t->load_order + 1 >= batchStart.size()) continue;  // its batch was cleared before being added
            t->ownerID = (uint32_t)(firstOwner + batchStart[t->load_order]);
            t->nSpanOwners = batchStart[t->load_order + 1] - batchStart[t->load_order];
        } else {
            t->ownerID = (uint32_t)(m_init_clumps + t->load_order);
            t->nSpanOwners = 1;
        }
    }
}

std::shared_ptr<DEMForceModel> DEMSolver::ReadContactForceModel(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("The force model file " + filename + " is not found.");
    std::stringstream ss;
    ss << f.rdbuf();
    m_force_model->DefineCustomModel(ss.str());
    return m_force_model;
}

void DEMSolver::ClearCache() {
    cached_input_clump_batches.clear();
    cached_extern_objs.clear();
    m_input_no_contact_pairs.clear();
}

void DEMSolver::UpdateClumps() {
    if (!sys_initialized)
        throw std::runtime_error(
            "Please call UpdateClumps only after the system is initialized, because it is for adding additional clumps "
            "to an initialized DEM system.");
    if (m_ext_loads != m_init_ext_loads)
        throw std::runtime_error("UpdateClumps cannot be used after loading new analytical objects.");
    if (m_templates.size() != m_init_templates)
        throw std::runtime_error("UpdateClumps cannot be used after loading new clump templates.");
    if (m_loaded_materials.size() != m_init_materials || m_input_family_prescription.size() != m_init_prescriptions)
        throw std::runtime_error(
            "UpdateClumps should not be used if you introduce new material types or family prescription.");
    pushHostState();
    syncHostState();
    size_t nNew = 0;
    for (const auto& b : cached_input_clump_batches) nNew += b->GetNumClumps();
    if (nNew == 0) return;
    const size_t nOwnersBefore = nOwnerBodies;
    std::vector<uint32_t> sOwner;
    std::vector<float> sR, sRel, mass, moi;
    std::vector<uint8_t> sMat;
    std::vector<uint64_t> vox(nNew);
    std::vector<uint16_t> lx(nNew), ly(nNew), lz(nNew);
    std::vector<float> q[4], v[3], w[3];
    for (auto& a : q) a.resize(nNew);
    for (auto& a : v) a.resize(nNew);
    for (auto& a : w) a.resize(nNew);
    std::vector<uint8_t> fam(nNew);
    std::vector<uint32_t> tpl(nNew);
    size_t k = 0;
    for (const auto& b : cached_input_clump_batches) {
        for (size_t i = 0; i < b->GetNumClumps(); i++, k++) {
            const auto& t = b->types.at(i);
            if (!t) throw std::runtime_error("A clump in an AddClumps batch has no template.");
            tpl[k] = t->mark;
            mass.push_back(t->mass);
            moi.insert(moi.end(), {t->MOI.x, t->MOI.y, t->MOI.z});
            for (unsigned int j = 0; j < t->nComp; j++) {
                sOwner.push_back((uint32_t)k);
                sR.push_back(t->radii[j]);
                sRel.insert(sRel.end(), {t->relPos[j].x, t->relPos[j].y, t->relPos[j].z});
                sMat.push_back((uint8_t)t->materials[j]->load_order);
            }
            const float3 c = f3sub(b->xyz[i], m_boxLBF);
            host_pos_to_voxel(m_params, (double)c.x, (double)c.y, (double)c.z, vox[k], lx[k], ly[k], lz[k]);
            q[0][k] = b->oriQ[i].w;
            q[1][k] = b->oriQ[i].x;
            q[2][k] = b->oriQ[i].y;
            q[3][k] = b->oriQ[i].z;
            v[0][k] = b->vel[i].x;
            v[1][k] = b->vel[i].y;
            v[2][k] = b->vel[i].z;
            w[0][k] = b->angVel[i].x;
            w[1][k] = b->angVel[i].y;
            w[2][k] = b->angVel[i].z;
            fam[k] = (uint8_t)b->families[i];
        }
    }
    deme_owner_state st{};
    st.voxelID = vox.data();
    st.locX = lx.data();
    st.locY = ly.data();
    st.locZ = lz.data();
    st.oriQw = q[0].data();
    st.oriQx = q[1].data();
    st.oriQy = q[2].data();
    st.oriQz = q[3].data();
    st.vX = v[0].data();
    st.vY = v[1].data();
    st.vZ = v[2].data();
    st.omgX = w[0].data();
    st.omgY = w[1].data();
    st.omgZ = w[2].data();
    st.family = fam.data();
    check(deme_append_clumps(m_ctx, (uint32_t)sOwner.size(), sOwner.data(), sR.data(), sRel.data(), sMat.data(),
                             (uint32_t)nNew, mass.data(), moi.data(), &st),
          "deme_append_clumps");
    // host copies follow the device layout: the new clumps after the old clumps, the external objects after them
    const size_t at = nOwnerClumps;
    auto ins = [&](auto& dst, const auto& src) { dst.insert(dst.begin() + at, src.begin(), src.end()); };
    ins(h_voxel, vox);
    ins(h_lx, lx);
    ins(h_ly, ly);
    ins(h_lz, lz);
    ins(h_qw, q[0]);
    ins(h_qx, q[1]);
    ins(h_qy, q[2]);
    ins(h_qz, q[3]);
    ins(h_vx, v[0]);
    ins(h_vy, v[1]);
    ins(h_vz, v[2]);
    ins(h_wx, w[0]);
    ins(h_wy, w[1]);
    ins(h_wz, w[2]);
    ins(h_fam, fam);
    for (auto* a : {&h_ax, &h_ay, &h_az, &h_alx, &h_aly, &h_alz}) a->insert(a->begin() + at, nNew, 0.f);
    for (size_t i = 0; i < sOwner.size(); i++) {
        m_sph_owner.push_back((uint32_t)(at + sOwner[i]));
        m_sph_r.push_back(sR[i]);
        m_sph_rel.push_back(make_float3(sRel[3 * i], sRel[3 * i + 1], sRel[3 * i + 2]));
    }
    m_clump_template.insert(m_clump_template.end(), tpl.begin(), tpl.end());
    for (auto& o : m_anal_owner) o += (uint32_t)nNew;
    nOwnerClumps += nNew;
    nOwnerBodies += nNew;
    nSpheresGM += sOwner.size();
    std::vector<float> vol(nOwnerClumps);
    for (size_t c = 0; c < nOwnerClumps; c++) vol[c] = m_templates.at(m_clump_template[c])->volume;
    check(deme_set_owner_volumes(m_ctx, vol.data()), "deme_set_owner_volumes");
    assignTrackers(nOwnersBefore);
    m_mirror_valid = true;
    m_mirror_dirty = false;
}

void DEMSolver::ChangeClumpSizes(const std::vector<uint32_t>& IDs, const std::vector<float>& factors) {
    if (!sys_initialized)
        throw std::runtime_error(
            "ChangeClumpSizes operates on device-side arrays directly, so it requires the system to be initialized "
            "first.");
    if (IDs.size() != factors.size()) throw std::runtime_error("ChangeClumpSizes: IDs and factors differ in length.");
    std::vector<uint32_t> in(IDs.size());
    for (size_t k = 0; k < IDs.size(); k++) {
        if (IDs[k] >= nOwnerBodies) throw std::runtime_error("ChangeClumpSizes: owner " + std::to_string(IDs[k]) +
                                                             " does not exist.");
        in[k] = toInternal(IDs[k]);
    }
    check(deme_change_clump_sizes(m_ctx, (uint32_t)in.size(), in.data(), factors.data()), "deme_change_clump_sizes");
    // the host's geometry (output files) follows
    std::vector<float> f(nOwnerClumps, 1.f);
    std::vector<uint8_t> mark(nOwnerClumps, 0);
    for (size_t k = 0; k < in.size(); k++) {
        f[in[k]] = factors[k];
        mark[in[k]] = 1;
    }
    for (size_t i = 0; i < nSpheresGM; i++) {
        const uint32_t o = m_sph_owner[i];
        if (!mark[o]) continue;
        m_sph_rel[i] = make_float3(m_sph_rel[i].x * f[o], m_sph_rel[i].y * f[o], m_sph_rel[i].z * f[o]);
        m_sph_r[i] = m_sph_r[i] * f[o];
    }
}

// ---------------------------------------------------------------------------------------------------------------
// inspectors and trackers
// ---------------------------------------------------------------------------------------------------------------
DEMInspector::DEMInspector(DEMSolver* s, const std::string& q, const std::string& r, bool all)
    : sys(s), quantity(q), region(r), all_domain(all) {
    // DEMInspector::switch_quantity_type (AuxClasses.cpp:58-103)
    if (q != "clump_max_z" && q != "clump_min_z" && q != "clump_max_absv" && q != "clump_volume")
        throw std::runtime_error(q + " is not a known query type.");
}

float DEMInspector::GetValue() {
    if (!sys->GetInitStatus())
        throw std::runtime_error(
            "Inspector should only be initialized or used after the simulation system is initialized (because it "
            "uses device-side data)!");
    sys->pushHostState();
    float v = 0.f;
    const int st = deme_inspect(sys->GetContext(), quantity.c_str(), all_domain ? nullptr : region.c_str(), &v);
    if (st != DEME_OK) {
        std::string what = deme_last_error(sys->GetContext());
        if (st == DEME_ERR_INVALID_ARGUMENT && !all_domain)
            what = "One of your inspectors is set to query a specific region, but the domain is not properly defined. "
                   "It needs to return a bool variable that is a result of logical operations involving X, Y and Z.";
        throw std::runtime_error("Inspector " + quantity + ": " + what);
    }
    return v;
}

std::shared_ptr<DEMInspector> DEMSolver::CreateInspector(const std::string& quantity) {
    return std::make_shared<DEMInspector>(this, quantity, " ", true);
}
std::shared_ptr<DEMInspector> DEMSolver::CreateInspector(const std::string& quantity, const std::string& region) {
    return std::make_shared<DEMInspector>(this, quantity, region, false);
}

std::shared_ptr<DEMTracker> DEMSolver::Track(std::shared_ptr<DEMExternObj>& obj) {
    auto t = std::make_shared<DEMTrackedObj>();
    t->kind = DEMTrackedObj::Kind::EXTERNAL;
    t->load_order = obj->load_order;
    m_tracked_objs.push_back(t);
    auto tr = std::make_shared<DEMTracker>(this);
    tr->obj = t;
    return tr;
}
std::shared_ptr<DEMTracker> DEMSolver::Track(std::shared_ptr<DEMClumpBatch>& obj) {
    auto t = std::make_shared<DEMTrackedObj>();
    t->kind = DEMTrackedObj::Kind::CLUMP_BATCH;
    t->load_order = obj->load_order;
    m_tracked_objs.push_back(t);
    auto tr = std::make_shared<DEMTracker>(this);
    tr->obj = t;
    return tr;
}

uint32_t DEMTracker::owner(size_t offset) const {
    if (!sys->GetInitStatus()) throw std::runtime_error("A tracker is used before the system is initialized.");
    if (offset >= obj->nSpanOwners)
        throw std::runtime_error("Tracker offset " + std::to_string(offset) + " is beyond the " +
                                 std::to_string(obj->nSpanOwners) + " owners it tracks.");
    return obj->ownerID + (uint32_t)offset;
}
float3 DEMTracker::Pos(size_t offset) { return sys->GetOwnerPosition(owner(offset)); }
float3 DEMTracker::AngVelLocal(size_t offset) { return sys->GetOwnerAngVel(owner(offset)); }
float3 DEMTracker::Vel(size_t offset) { return sys->GetOwnerVelocity(owner(offset)); }
float4 DEMTracker::OriQ(size_t offset) { return sys->GetOwnerOriQ(owner(offset)); }
unsigned int DEMTracker::GetFamily(size_t offset) { return sys->GetOwnerFamily(owner(offset)); }
float3 DEMTracker::ContactAcc(size_t offset) { return sys->GetOwnerAcc(owner(offset)); }
float3 DEMTracker::ContactAngAccLocal(size_t offset) { return sys->GetOwnerAngAcc(owner(offset)); }
void DEMTracker::SetPos(float3 pos, size_t offset) { sys->SetOwnerPosition(owner(offset), pos); }
void DEMTracker::SetAngVel(float3 w, size_t offset) { sys->SetOwnerAngVel(owner(offset), w); }
void DEMTracker::SetVel(float3 v, size_t offset) { sys->SetOwnerVelocity(owner(offset), v); }
void DEMTracker::SetOriQ(float4 q, size_t offset) { sys->SetOwnerOriQ(owner(offset), q); }

size_t DEMSolver::GetNumContacts() {
    deme_counts c;
    check(deme_get_counts(m_ctx, &c), "deme_get_counts");
    return c.nContacts;
}

double DEMSolver::GetTimeElapsed() {
    deme_counts c;
    check(deme_get_counts(m_ctx, &c), "deme_get_counts");
    return c.time;
}

void DEMSolver::ShowTimingStats() {
    const char* names[64];
    double ms[64];
    uint64_t cnt[64];
    int n = 0;
    check(deme_kernel_timing(m_ctx, 64, names, ms, cnt, &n), "deme_kernel_timing");
    std::printf("\n~~ DEVICE KERNEL TIMING ~~\n");
    for (int i = 0; i < n && i < 64; i++)
        std::printf("%-16s %12.3f ms over %8llu launches\n", names[i], ms[i], (unsigned long long)cnt[i]);
}

}  // namespace deme
