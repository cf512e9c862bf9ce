/*
 * Plain-old-data types shared by the DEM hot-path C ABI (deme.h) and its callers.
 *
 * Integer widths follow the reference's type table (src/DEM/VariableTypes.h:11-49):
 *   subVoxelPos_t u16, voxelID_t u64, oriQ_t f32, bodyID_t/binID_t/contactPairs_t u32,
 *   binsSphereTouches_t/spheresBinTouches_t/geoSphereTouches_t u16,
 *   family_t/contact_t/materialsOffset_t/objID_t u8.
 * Constants follow src/DEM/Defines.h:24-100.
 *
 * No pointer in these structs is ever a torch type; device-resident arrays are owned by a deme_ctx.
 */
#ifndef DEME_TYPES_H
#define DEME_TYPES_H

#ifdef __HIPCC_RTC__
/* compiled by hipRTC (runtime specialisation of the force kernel): the fixed-width types come from its runtime */
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#else
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Contact types (src/DEM/Defines.h:64-68). */
#define DEME_NOT_A_CONTACT 0
#define DEME_SPHERE_SPHERE_CONTACT 1
#define DEME_SPHERE_MESH_CONTACT 2
#define DEME_SPHERE_PLANE_CONTACT 3
#define DEME_SPHERE_PLATE_CONTACT 4

/* Analytical component types (src/DEM/Defines.h:59-60). */
#define DEME_ANAL_OBJ_TYPE_PLANE 0
#define DEME_ANAL_OBJ_TYPE_PLATE 1

/* New-contact marker in the history map (src/DEM/Defines.h:80-81): 2^32 - 1. */
#define DEME_NULL_MAPPING_PARTNER 0xFFFFFFFFu

/* Reserved (always fixed) family (src/DEM/Defines.h:88). */
#define DEME_RESERVED_FAMILY_NUM 255u
#define DEME_NUM_FAMILIES 256u
/*
 * Family mask entries. The reference sizes its mask at 255*256/2 (src/DEM/APIPrivate.cpp:588) but indexes it with
 * locateMaskPair(i, j) = (1 + j) * j / 2 + i for j <= 255 (src/kernel/DEMHelperKernels.cu:52-56), i.e. up to 32895.
 * We size it for the full upper triangle; the extra row (j == 255) defaults to "contact allowed".
 */
#define DEME_NUM_MASK_ENTRIES (256u * 257u / 2u)
#define DEME_DONT_PREVENT_CONTACT 0
#define DEME_PREVENT_CONTACT 1

/* Maximum spheres in one bin (src/DEM/Defines.h:25). Exceeding it is a reported error, never a trap. */
#define DEME_MAX_SPHERES_PER_BIN 256u

/* Time integrators (src/DEM/Defines.h:106); the pass-on rules are in src/kernel/DEMCustomizablePolicies/
 * IntegrationVelPassOn{ForwardEuler,CenteredDiff,ExtendedTaylor}.cu. */
#define DEME_INTEGRATOR_FORWARD_EULER 0
#define DEME_INTEGRATOR_CENTERED_DIFFERENCE 1
#define DEME_INTEGRATOR_EXTENDED_TAYLOR 2

/* Status codes returned by every ABI entry point (0 == success). */
#define DEME_OK 0
#define DEME_ERR_INVALID_ARGUMENT 1
#define DEME_ERR_HIP 2
#define DEME_ERR_BIN_OVERFLOW 3      /* a bin holds more than DEME_MAX_SPHERES_PER_BIN spheres */
#define DEME_ERR_CAPACITY 4          /* a count exceeds the width of its reference integer type */
#define DEME_ERR_NOT_READY 5         /* call order violated (e.g. force before contact detection) */
#define DEME_ERR_JIT 6               /* runtime specialisation failed */
#define DEME_ERR_UNSUPPORTED 7       /* a feature outside the hot path was requested */
#define DEME_ERR_NO_DEVICE 8

/*
 * Simulation parameters (the fields of DEMSimParams the hot path reads, src/DEM/Defines.h:154-213).
 *
 * l / voxelSize are the values the reference's kernels see: the host formats them as fixed-point strings with 17
 * decimals (to_string_with_precision, src/DEM/HostSideHelpers.hpp:397-402; substituted at
 * src/DEM/APIPrivate.cpp:1358-1359), so kernels run on strtod() of that text. l_exact / voxelSize_exact are the values
 * the host uses when it first encodes positions (src/DEM/dT.cpp:572-575) and when it reads them back.
 */
typedef struct deme_sim_params {
    uint8_t nvXp2, nvYp2, nvZp2;
    uint8_t integrator;
    uint32_t nbX, nbY, nbZ;
    uint32_t pad0;
    double l;
    double voxelSize;
    double l_exact;
    double voxelSize_exact;
    double binSize;
    float LBF[3];
    float G[3];
    float h;
    float beta;
    float boxSize[3]; /* the user's domain (InstructBoxDomainDimension); 0 = the whole voxel world */
    float pad1;
} deme_sim_params;

/* Material properties used by the frictional Hertz-Mindlin model (FullHertzianForceModel.cu:5-19). */
typedef struct deme_material {
    float E, nu, CoR, mu, Crr;
} deme_material;

/* One analytical component (AnalyticalCompDefJitify.cu:1-14; built at src/DEM/APIPrivate.cpp:1062-1085). */
typedef struct deme_anal_obj {
    uint32_t owner;   /* absolute owner index (clumps first, then external objects) */
    uint8_t type;     /* DEME_ANAL_OBJ_TYPE_* */
    uint8_t material;
    uint8_t normal;   /* ENTITY_NORMAL_INWARD / OUTWARD; unused by planes */
    uint8_t pad;
    float relPos[3];  /* position in the owner frame */
    float rot[3];     /* plane normal in the owner frame */
    float size[3];
    float mass;       /* owner mass as seen by the force kernel */
} deme_anal_obj;

/*
 * Prescribed motion of one family (equipFamilyPrescribedMotions, src/DEM/APIPrivate.cpp:1005-1048). The reference
 * splices C expressions; the ahead-of-time path accepts numeric constants, carried here as values plus "assigned"
 * bits: velSet bit k (k = 0..5) assigns vX, vY, vZ, omgBarX, omgBarY, omgBarZ; posSet bit k assigns X, Y, Z.
 */
typedef struct deme_family_rule {
    uint8_t used;
    uint8_t linVelPrescribed, rotVelPrescribed, linPosPrescribed, rotPosPrescribed;
    uint8_t velSet, posSet;
    uint8_t pad;
    float vel[6];
    double pos[3];
} deme_family_rule;

/* Host-side owner state arrays (SoA). Any pointer may be NULL in a download call to skip that field. */
typedef struct deme_owner_state {
    uint64_t* voxelID;
    uint16_t *locX, *locY, *locZ;
    float *oriQw, *oriQx, *oriQy, *oriQz;
    float *vX, *vY, *vZ;
    float *omgX, *omgY, *omgZ;
    float *aX, *aY, *aZ;
    float *alphaX, *alphaY, *alphaZ;
    uint8_t* family;
} deme_owner_state;

/* Counters of the last contact-detection pass and step bookkeeping. */
typedef struct deme_counts {
    uint64_t nSpheres, nOwners, nAnal;
    uint64_t nBinSphereTouches; /* P */
    uint64_t nActiveBins;       /* bins touched by at least one sphere */
    uint64_t nContacts;         /* C (sphere-sphere + sphere-plane) */
    uint64_t nPrevContacts;
    uint64_t nSphereGeoContacts;
    uint64_t maxSpheresInBin;
    uint64_t nSteps;
    double time;
    uint64_t nCells; /* cells of the contact-detection cell grid (0 before the first cell-list pass) */
} deme_counts;

#ifdef __cplusplus
}
#endif

#endif
