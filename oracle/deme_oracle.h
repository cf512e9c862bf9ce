/*
 * CPU oracle for the DEM per-step hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A host C++17/OpenMP restatement of the reference's kernel math. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, always as the checker or as the reported CPU baseline; the product path never
 * links or calls it.
 *
 * Every function names the reference file:line it restates. The reference has no CPU solver and cannot be built here
 * (CUDA + jitify + CUB); this restatement is pinned by brute force, closed forms and invariants in tests/.
 */
#ifndef DEME_ORACLE_H
#define DEME_ORACLE_H

#include "../include/deme/deme_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

/* ---- host-side sizing restatements (src/DEM/APIPrivate.cpp) ---- */
/* figureOutNV (APIPrivate.cpp:222-333) with m_box_dir_length_is_exact == NONE. Returns l as float. */
void ora_figure_out_nv(float boxX, float boxY, float boxZ, int* nvX, int* nvY, int* nvZ, float* l);
/* strtod of to_string_with_precision(x, 17) (HostSideHelpers.hpp:397-402). */
double ora_literal17(double x);
/* locateMaskPair (DEMHelperKernels.cu:51-56) and recoverCntPair (DEMHelperKernels.cu:60-64). */
uint32_t ora_locate_mask_pair(uint32_t i, uint32_t j);
void ora_recover_cnt_pair(uint32_t ind, uint32_t n, uint32_t* i, uint32_t* j);

/* ---- system set-up ---- */
ora_ctx* ora_create(const deme_sim_params* p, uint32_t nSpheres, const uint32_t* sphOwner, const float* sphRadius,
                    const float* sphRelPos /* 3n */, const uint8_t* sphMat, uint32_t nOwners, const float* ownerMass,
                    const float* ownerMoi /* 3n */, uint32_t nAnal, const deme_anal_obj* anal, uint32_t nMat,
                    const deme_material* mats, const uint8_t* familyMask /* DEME_NUM_MASK_ENTRIES */,
                    const deme_family_rule* rules /* 256 */);
void ora_destroy(ora_ctx* c);
void ora_set_threads(int n);
/* Batch scenes: a sphere meets only the analytical components of its scene (analScene NULL: the owner's scene). */
void ora_set_scenes(ora_ctx* c, uint32_t nScenes, const uint32_t* ownerScene, const uint32_t* analScene);
void ora_set_state(ora_ctx* c, const deme_owner_state* s);
void ora_get_state(const ora_ctx* c, deme_owner_state* s);
void ora_get_counts(const ora_ctx* c, deme_counts* out);

/* ---- contact detection (src/algorithms/DEMCubContactDetection.cu:39-397) ---- */
/* Rows a6-a8: per-sphere bin counts / plane counts, then the stably sorted (binID, sphereID) table. */
int ora_cd_bins(ora_ctx* c);
void ora_get_bin_counts(const ora_ctx* c, uint16_t* numBinsSphereTouches, uint8_t* numAnalGeoSphereTouches);
void ora_get_bin_pairs(const ora_ctx* c, uint32_t* binIDs, uint32_t* sphereIDs); /* P entries, sorted */
void ora_get_active_bins(const ora_ctx* c, uint32_t* binIDs, uint16_t* counts);   /* B_a entries */
/* Row a9-a12: pair detection, canonical order, history map. Requires ora_cd_bins first. */
int ora_cd_pairs(ora_ctx* c);
/* Rows a6-a13 in sequence, including the wildcard remap. */
int ora_cd(ora_ctx* c);
void ora_get_contacts(const ora_ctx* c, uint32_t* idA, uint32_t* idB, uint8_t* type, uint32_t* map);
/* Replace the current contact list (e.g. with one produced by the device) and remap wildcards from the previous one. */
int ora_set_contacts(ora_ctx* c, uint64_t n, const uint32_t* idA, const uint32_t* idB, const uint8_t* type);
/* Brute-force O(S^2) sphere-sphere detection restating the bin rule as a predicate (pin only, tiny inputs). */
uint64_t ora_brute_force_pairs(ora_ctx* c, uint32_t* idA, uint32_t* idB, uint64_t cap);

/* ---- dynamics ---- */
int ora_force(ora_ctx* c);       /* a14-a15 */
int ora_collect(ora_ctx* c);     /* a16 */
int ora_integrate(ora_ctx* c);   /* a17, at the context's current time */
int ora_step(ora_ctx* c, int nSteps, int updateFreq); /* a18 deterministic schedule */
void ora_get_contact_forces(const ora_ctx* c, float* force /*3C*/, float* torqueOnly /*3C*/, float* cpA /*3C*/,
                            float* cpB /*3C*/);
void ora_get_wildcards(const ora_ctx* c, float* w /* 4 x C, wildcard-major */);
void ora_set_wildcards(ora_ctx* c, const float* w /* 4 x C */);
void ora_set_contact_forces(ora_ctx* c, const float* force, const float* torqueOnly, const float* cpA,
                            const float* cpB); /* 3C each, the current list's */
void ora_get_time(const ora_ctx* c, double* t, uint64_t* nSteps);
void ora_set_time(ora_ctx* c, double t);
/* Wall seconds accumulated per phase by ora_step / ora_cd (Discretize domain, Find contact pairs, Build history map,
 * Calculate contact forces, Collect contact forces, Integration); reset != 0 clears them after reading. */
void ora_phase_times(ora_ctx* c, double* out6, int reset);

/* FP64 evaluation of the same per-contact force for tolerance checks (same inputs, every op in double). */
void ora_force_fp64(const ora_ctx* c, double* force /*3C*/, double* torqueOnly /*3C*/);

#ifdef __cplusplus
}
#endif

#endif
