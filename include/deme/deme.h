/*
 * C ABI of the MI355X-native DEM hot path (libdeme.so).
 *
 * The reference drives this path from two worker threads through internal C++ seams:
 *   contactDetection(...)      src/algorithms/DEMCubBasedSubroutines.h:48-65 (impl DEMCubContactDetection.cu:39-397)
 *   collectContactForces(...)  src/algorithms/DEMCubBasedSubroutines.h:67-74 (impl DEMCubForceCollection.cu:20-158)
 *   jitify kernel launches     e.g. src/DEM/dT.cpp:1385-1388 (calculateContactForces), dT.cpp:1424-1431 (integrate)
 * This ABI replaces those seams. A context owns all device memory (SoA/AoS in HBM) and one HIP stream; every call
 * runs on that stream. All entry points return a DEME_* status (0 == DEME_OK); deme_last_error() gives the message.
 * No C++ exception and no device trap crosses this boundary.
 *
 * Index conventions follow the reference: owners are clumps [0, nClumps) followed by external objects; spheres are
 * numbered clump by clump (ownerClumpBody is non-decreasing); analytical components are numbered in load order.
 */
#ifndef DEME_H
#define DEME_H

#include "deme_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct deme_ctx deme_ctx;

/* ---------------------------------------------------------------------------------------------------------------
 * Lifecycle. Replaces DEMSolver's construction of kT/dT and their streams (src/DEM/APIPublic.cpp:22-41,
 * src/core/utils/GpuManager.cpp:8-33); one context == one device + one stream (stream may be NULL: the context then
 * creates its own default-priority stream).
 * ------------------------------------------------------------------------------------------------------------- */
int deme_ctx_create(int device, void* hip_stream, deme_ctx** out);
int deme_ctx_destroy(deme_ctx* ctx);
const char* deme_last_error(const deme_ctx* ctx);
int deme_device_count(int* n);
int deme_get_stream(const deme_ctx* ctx, void** hip_stream);
int deme_sync(deme_ctx* ctx);

/* ---------------------------------------------------------------------------------------------------------------
 * Host-side sizing (no device work). Restate DEMSolver::figureOutNV (src/DEM/APIPrivate.cpp:222-333),
 * decideBinSize (APIPrivate.cpp:335-369), transferSimParams' expand factor (APIPrivate.cpp:730-737) and the JIT
 * literal formatting of l / voxelSize (HostSideHelpers.hpp:397-402 used at APIPrivate.cpp:1358-1359).
 * exact_dir: -1 none, 0 X, 1 Y, 2 Z (SPATIAL_DIR, InstructBoxDomainDimension at APIPublic.cpp:92-99).
 * ------------------------------------------------------------------------------------------------------------- */
int deme_figure_out_nv(float boxX, float boxY, float boxZ, int exact_dir, uint8_t nv_out[3], float* l_out);
double deme_kernel_literal(double value);
/* Fill a deme_sim_params: nv/l from the box (or explicit nv + l when nv_in != NULL, InstructBoxDomainNumVoxel at
 * APIPublic.cpp:101-119), bin size = 2 * min_radius unless bin_size > 0 (SetInitBinSize), nb = floor(world/bin)+1. */
int deme_build_sim_params(float boxX, float boxY, float boxZ, int exact_dir, const uint8_t* nv_in, float l_in,
                          const float LBF[3], float min_radius, double bin_size, const float G[3], float h,
                          float beta, int integrator, deme_sim_params* out);
/* World coordinates (float xyz, n x 3) <-> voxel + sub-voxel integers with the exact l / voxelSize, as the host does
 * at initialisation (dT.cpp:528-575: coordinate - LBF in float, then hostPositionToVoxelID) and in getOwnerPos
 * (dT.cpp:1770-1781: hostVoxelIDToPosition + LBF). */
int deme_encode_positions(const deme_sim_params* p, uint64_t n, const float* xyz, uint64_t* voxelID, uint16_t* locX,
                          uint16_t* locY, uint16_t* locZ);
int deme_decode_positions(const deme_sim_params* p, uint64_t n, const uint64_t* voxelID, const uint16_t* locX,
                          const uint16_t* locY, const uint16_t* locZ, float* xyz);

/* ---------------------------------------------------------------------------------------------------------------
 * Static data. Replaces allocateManagedArrays/populateEntityArrays (src/DEM/kT.cpp:372-626, dT.cpp:214-693) and the
 * JIT definitions of materials, analytical objects and family prescriptions (APIPrivate.cpp:1005-1347).
 * ------------------------------------------------------------------------------------------------------------- */
int deme_set_params(deme_ctx* ctx, const deme_sim_params* p);
/* Per sphere: owner (clump index), radius, relPos (3 floats), material. Per owner (clumps and external objects):
 * mass and principal MOI (3 floats). nClumps <= nOwners; spheres must be grouped by owner in ascending order. */
int deme_upload_templates(deme_ctx* ctx, uint32_t nSpheres, const uint32_t* sphOwner, const float* sphRadius,
                          const float* sphRelPos, const uint8_t* sphMat, uint32_t nOwners, uint32_t nClumps,
                          const float* ownerMass, const float* ownerMoi);
int deme_upload_materials(deme_ctx* ctx, uint32_t n, const deme_material* mats);
int deme_upload_analytical(deme_ctx* ctx, uint32_t n, const deme_anal_obj* objs);
/* mask: DEME_NUM_MASK_ENTRIES entries (locateMaskPair layout); rules: DEME_NUM_FAMILIES entries, or NULL to update
 * only the mask after the first upload (DisableContactBetweenFamilies after initialisation, APIPublic.cpp:448-466). */
int deme_upload_families(deme_ctx* ctx, const uint8_t* mask, const deme_family_rule* rules);

/* Batches of independent scenes in one context (no counterpart in the reference, which runs one scene per solver):
 * ownerScene[o] in [0, nScenes) for every owner; analScene[a] for every analytical component, or NULL to use its
 * owner's scene. Spheres only meet analytical components of their owner's scene; components must come grouped by
 * scene, at most 64 per scene. Without this call everything is scene 0. Call after deme_upload_analytical. */
int deme_upload_scenes(deme_ctx* ctx, uint32_t nScenes, const uint32_t* ownerScene, const uint32_t* analScene);
/* State I/O (dT getters/setters, src/DEM/dT.cpp:1722-1806; initial encoding dT.cpp:572-575). */
int deme_upload_state(deme_ctx* ctx, const deme_owner_state* s);
int deme_download_state(deme_ctx* ctx, deme_owner_state* s);
/* Drops the contact list with its tangential histories and restarts the step count and time at zero, so the next step
 * begins with a detection pass on the uploaded state, as the first step after Initialize (API.h:620) does. */
int deme_clear_contacts(deme_ctx* ctx);

/* ---------------------------------------------------------------------------------------------------------------
 * Contact detection (contactDetection, DEMCubContactDetection.cu:39-397). The four stages may be called in order,
 * or all at once with deme_cd_run.
 * ------------------------------------------------------------------------------------------------------------- */
/* getNumberOfBinsEachSphereTouches (DEMBinSphereKernels.cu:12-115): per-sphere bin and plane counts, occupancy. */
int deme_cd_bin_count(deme_ctx* ctx);
/* prefix scan + populateBinSphereTouchingPairs + SortPairs/RLE (DEMCubContactDetection.cu:77-166). */
int deme_cd_bin_fill(deme_ctx* ctx);
/* getNumberOfContactsEachBin / populateContactPairsEachBin (DEMContactKernels_Blockwise.cu:15-309). */
int deme_cd_pairs(deme_ctx* ctx);
/* contact sort + history map (DEMCubContactDetection.cu:236-372, DEMHistoryMappingKernels.cu:5-60). */
int deme_cd_sort_remap(deme_ctx* ctx);
int deme_cd_run(deme_ctx* ctx);
int deme_download_bin_counts(deme_ctx* ctx, uint16_t* numBinsSphereTouches, uint8_t* numAnalGeoSphereTouches);
/* binStart: nBins + 1 entries; binSphere: P entries (sphere ids of each bin, unordered within a bin). */
int deme_download_bins(deme_ctx* ctx, uint32_t* binStart, uint32_t* binSphere);
/* Canonical contact list: by idA; plane contacts before sphere contacts; sphere partners by ascending idB. */
int deme_download_contacts(deme_ctx* ctx, uint32_t* idA, uint32_t* idB, uint8_t* type, uint32_t* map);
/* The device-wide exclusive prefix sum of the detection stages (cubDEMPrefixScan, DEMCubWrappers.cu:26-41, used by
 * contactDetection at DEMCubContactDetection.cu:81-93): d_out[i] = d_in[0] + ... + d_in[i-1] for i <= n (d_out holds
 * n + 1 entries; d_out[n] is the total modulo 2^32), *total (host, nullable) = the total in 64 bits. d_in and d_out
 * are device pointers (they may be the same array); the call runs on the context's stream and synchronises it. Sums
 * are formed in 64 bits in a fixed order. */
int deme_prefix_scan_u32(deme_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n, uint64_t* total);
/* The same scan without allocating anything (no counterpart in the reference): for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. Its tile sums
 * live in a fixed device workspace (about 1 MiB per GPU, looked up when the context is created and held until the
 * process ends), so n is at most 2^28 (DEME_ERR_INVALID_ARGUMENT beyond). d_out receives n + 1 entries as above;
 * d_total (device pointer, nullable) receives the 64-bit total. The call only enqueues on the context's stream and
 * returns without waiting. Every call on one GPU shares that workspace: calls that overlap on one GPU (two streams
 * or two contexts at once) race; calls ordered on one stream do not. */
int deme_prefix_scan_u32_noalloc(deme_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                                 uint64_t* d_total);

/* ---------------------------------------------------------------------------------------------------------------
 * Dynamics (DEMDynamicThread::calculateForces / integrateOwnerMotions, src/DEM/dT.cpp:1349-1432).
 * ------------------------------------------------------------------------------------------------------------- */
/* prepareForceArrays + rearrangeContactWildcards + calculateContactForces (DEMPrepForceKernels.cu:23-55,
 * DEMCalcForceKernels.cu:14-179 with FullHertzianForceModel.cu). */
int deme_force(deme_ctx* ctx);
/* collectContactForces (DEMCubForceCollection.cu:20-158): per-owner linear and angular accelerations. */
int deme_reduce_owner(deme_ctx* ctx);
/* integrateOwners (DEMIntegrationKernels.cu:170-179), preceded by the owner reduction of the last forces. */
int deme_integrate(deme_ctx* ctx);
/* integrateOwners alone (DEMIntegrationKernels.cu:170-179): one step of every owner from the accelerations the context
 * holds (those of the last deme_reduce_owner / deme_integrate, or uploaded with deme_upload_state), without collecting
 * contact forces; needs no contact list. */
int deme_integrate_owners(deme_ctx* ctx);
/* The step loop (DEMDynamicThread::workerThread, dT.cpp:1527-1585) with a deterministic contact-update schedule:
 * detection on the state at steps k*f (f = max(update_freq, 1)), force -> collect -> integrate each step. */
int deme_step(deme_ctx* ctx, int n_steps, int update_freq);
/* HIP-graph replay in deme_step (no counterpart in the reference, which launches every kernel and waits on the host
 * for each detection pass's counts). When on, each whole detection interval -- the detection pass and the update_freq
 * steps up to the next one -- is one graph launch, captured once per list parity and replayed while the buffers it
 * baked in stay valid: the kernels after the pass take the contact counts from device memory and skip themselves when
 * the pass failed or its lists outgrew the captured capacity, in which case the replay changes nothing and the
 * interval runs directly (regrowing the buffers) before the next capture. Steps outside whole intervals replay a
 * per-step graph. Off by default (direct launches). Results are identical either way. Kernel timing
 * (deme_enable_kernel_timing) runs directly. */
int deme_set_step_graphs(deme_ctx* ctx, int on);
/* Capacity a captured interval gets over the current list: factor x the contact (and mirror-row) count + slack
 * (defaults 1.25 and 1024). */
int deme_set_graph_headroom(deme_ctx* ctx, double factor, uint64_t slack);
/* Candidate lists for contact detection (no counterpart in the reference, whose bins are re-binned every pass):
 * a build finds every pair of different owners within `skin` (metres) of touching; until some sphere has moved half
 * the skin since, a detection pass evaluates the reference's predicate (expanded overlap, family mask, contact-point
 * bin rule) on those pairs only, which are all the pairs that can satisfy it -- the contact lists are the same as
 * the full cell sweep's. The pass measures the largest displacement itself; a pass that finds the lists stale
 * rebuilds them and repeats, and a displacement past 0.35 x skin rebuilds before the next pass. skin < 0 (default):
 * automatic, starting at 0.2 x the mean sphere radius and adapted at each build, within 0.05-0.8 x, to the lists it
 * retires (widened 1.5x when they died within a pass, narrowed 0.75x when they outlasted 400 passes, left alone from
 * 30 passes on; lists shorter-lived than that steer it 1.25x at a time by their candidates x (1 + 2.7 / passes
 * lasted)); when lists die within 4 passes even at the widest skin, the cell sweep runs every pass for 64 passes
 * (doubling while that repeats, up to 4096) before lists are tried again. 0: off (the cell sweep every pass). The
 * build positions are kept in fixed point (a power-of-two unit u, about 2^-31 of twice the bin grid's largest
 * extent); a skin with skin / 2 <= 1 nm + 4u also runs the cell sweep (deme_candidate_stats then reports 0). */
int deme_set_skin(deme_ctx* ctx, double skin);
/* The skin in force (0 while the automatic mode runs the cell sweep), the candidate pairs of the last build and the
 * number of builds so far (each nullable). */
int deme_candidate_stats(deme_ctx* ctx, double* skin, uint64_t* nCandidates, uint64_t* builds);
/* How the pair sweep of the cell-list detection tests its candidates, and its workgroup size (no counterpart in the
 * reference; the contact lists are the same in every configuration): mode 0 the double-precision overlap test on
 * every candidate; 1 a float pre-reject (rounding margin included) and the exact test on its survivors; 2 the float
 * test alone in the sweep, the same-owner rule and the exact test in the block's shared second phase; -1 the default
 * (mode 2). block: 64, 128 or 256 (default); 0 keeps it. */
int deme_set_pair_sweep(deme_ctx* ctx, int mode, int block);
/* Counters of the interval graphs since the context was created: out[0] captures, out[1] replays that advanced the
 * system, out[2] replays that fell back to direct launches. */
int deme_graph_stats(deme_ctx* ctx, uint64_t out[3]);
/* When on, the last step of every deme_step call also stores the per-contact records that deme_download_forces
 * returns (force, torque-only force, contact points), as the reference's force kernel stores contactForces,
 * contactTorque_convToForce and contactPointGeometryA/B each step (DEMCalcForceKernels.cu:160-175) for its contact
 * file writer (dT.cpp:1054-1180). Off by default: deme_step then keeps only the per-owner terms. */
int deme_set_keep_contact_records(deme_ctx* ctx, int on);

/* Drift margin beta (SetExpandFactor, API.h:100-104): detection uses r + beta. Takes effect at the next detection. */
int deme_set_beta(deme_ctx* ctx, float beta);
/* Cell edge of the cell-list detection as a multiple (>= 1) of the largest interaction distance 2 (r_max + beta)
 * (no counterpart in the reference, whose bin size plays this role for its per-bin pass). */
int deme_set_cell_scale(deme_ctx* ctx, double scale);
/* Autotuner (no counterpart in the reference): times trialSteps steps of the live system for every pair of
 * contact-update interval freqs[i] (drift margin maxVel * h * f * safety, as transferSimParams sets it,
 * APIPrivate.cpp:731-733; maxVel = 0 measures the largest owner speed) and cell scale cellScales[j], keeps the
 * fastest configured and returns its interval. msPerStep (optional, nFreqs x nScales) receives the timings. The trials
 * advance the simulation. */
int deme_autotune(deme_ctx* ctx, const int* freqs, int nFreqs, const double* cellScales, int nScales, int trialSteps,
                  float maxVel, float safety, int* bestFreq, double* msPerStep);
/* Second stage of the autotuner (no counterpart in the reference): times trialSteps steps at update interval
 * updateFreq (keeping the drift margin and cell scale in force) for every detection configuration skins[k] x
 * modes[i] x blocks[j] (deme_set_skin, deme_set_pair_sweep; skins may be NULL: the skin in force), keeps the fastest
 * configured and returns it. msPerStep (optional, nSkins x nModes x nBlocks) receives the timings. The trials advance
 * the simulation; contact lists are the same in every configuration. */
int deme_autotune_sweep(deme_ctx* ctx, const double* skins, int nSkins, const int* modes, int nModes,
                        const int* blocks, int nBlocks, int trialSteps, int updateFreq, double* bestSkin,
                        int* bestMode, int* bestBlock, double* msPerStep);

/* ---- Force model and runtime specialisation (hipRTC in place of jitify) ----
 * The force kernel is one template with the placeholders of the reference's JIT programs (DEMCalcForceKernels.cu:
 * _DEMForceModel_, _forceModelIngredientDefinition_, _forceModelIngredientAcqForA_/_B_, _forceModelContactWildcardAcq_
 * /_Write_/_Destroy_, _materialDefs_, _nvXp2_, _nvYp2_, _voxelSize_, _l_), substituted textually as JitHelper.cpp:44-48
 * does. The library ships it compiled with the built-in frictional Hertz-Mindlin model.
 *
 * deme_set_force_model: a force-model body written against the names of the reference's model strings (force,
 * torque_only_force, locCPA, locCPB, B2A, overlapDepth, AOwnerMass, BOwnerMass, ARadius, BRadius, AoriQ*, BoriQ*,
 * bodyAMatType, bodyBMatType, the material arrays E, nu, CoR, mu, Crr, the ingredients ts, ALinVel, BLinVel,
 * ARotVel, BRotVel, AOwnerFamily, BOwnerFamily, and its contact wildcards by name; DefineContactForceModel /
 * SetPerContactWildcards, API.h:115-122, AuxClasses.h:123-160), compiled now by hipRTC with the simulation constants
 * written in; DEME_ERR_JIT with the compiler log in deme_last_error on failure. At most 4 contact wildcards (kept as
 * one float4 per contact, in sorted name order, zero on creation). model = NULL returns to the built-in kernel.
 * deme_jit_build: the built-in model (or the model set last) with explicit placeholder values (keys / values, n
 * pairs) overriding the generated ones; with n = 0 it specialises the built-in kernel on the simulation constants.
 * deme_builtin_force_model: the source of built-in model 0 (frictional Hertz-Mindlin, wildcards delta_tan_x,
 * delta_tan_y, delta_tan_z, delta_time) or 1 (frictionless Hertz, no wildcards), NULL otherwise. */
int deme_set_force_model(deme_ctx* ctx, const char* model, const char* const* wildcards, int nWildcards);
/* deme_jit_build also takes the owner update kernel's placeholders (DEMIntegrationKernels.cu, DEMModeratorKernels.cu):
 * _velPrescriptionStrategy_ and _posPrescriptionStrategy_ (switch cases on `family`, as equipFamilyPrescribedMotions
 * writes them, APIPrivate.cpp:1005-1048: expressions in t setting vX vY vZ omgBarX omgBarY omgBarZ / X Y Z oriQw..z and
 * LinPrescribed, RotPrescribed) and _familyChangeRules_ (statements reading pos (double3, relative to the world's
 * lower corner), vel, acc, mass, family_code, t, h and setting newFamily, as equipFamilyOnFlyChanges writes them,
 * APIPrivate.cpp:980-1003); given any of them, the owner update kernel is compiled with them. Families a strategy
 * does not name keep the numeric prescriptions of deme_upload_families. */
int deme_jit_build(deme_ctx* ctx, const char* const* keys, const char* const* values, int n);
/* Drop the owner update kernel's strategies and rules (back to the shipped kernel). */
int deme_clear_motion_rules(deme_ctx* ctx);
/* ChangeFamily (API.h:268): every owner of family `from` moves to family `to` now. */
int deme_change_family(deme_ctx* ctx, int from, int to);
const char* deme_builtin_force_model(int which);
/* Context-free helpers: the specialised source for a model (model NULL: built-in model 0; p NULL: constants read
 * from the kernel parameters), written to out (cap bytes, NUL-terminated; *needed = full size + 1), and a hipRTC
 * compile of a source for gfx950 (log written to log, cap bytes). Neither needs a device. */
int deme_jit_render(const char* model, const char* const* wildcards, int nWildcards, const deme_sim_params* p,
                    char* out, uint64_t cap, uint64_t* needed);
int deme_jit_compile(const char* source, char* log, uint64_t cap);
/* The owner update kernel's source for the given strategies and rules (each nullable). */
int deme_jit_render_owner(const char* velStrategy, const char* posStrategy, const char* familyChangeRules, char* out,
                          uint64_t cap, uint64_t* needed);
/* The inspection kernels' source for a region (nullable: the whole domain), as deme_inspect compiles it. */
int deme_jit_render_inspect(const char* region, char* out, uint64_t cap, uint64_t* needed);

/* Adjoint of the step (no counterpart in the reference, which has no differentiation; for calibration callers that
 * fit material parameters or initial velocities to observed states). Systems of single-sphere owners (each sphere
 * its own, centred owner) with free families, fixed analytical walls, Crr = 0, the extended-Taylor integrator and at
 * most 64 materials; DEME_ERR_UNSUPPORTED otherwise.
 *
 * deme_adjoint_forward runs nSteps steps from the current state with the contact list rebuilt every `window` steps
 * (the drift margin set by deme_set_params / deme_set_beta must cover a window) and checkpoints the state and contact
 * list at every window start. deme_adjoint_backward then returns the gradients of a loss L on the final state:
 * dLdPos and dLdVel (nullable) hold dL/dX and dL/dv of the final owner positions and velocities, gradVel0 (nullable)
 * receives dL/dv of the initial velocities, each 3 x nOwners floats component-major (all x, then all y, then all z;
 * wall owners are ignored), and gradMat (nullable) receives dL/d(E, nu, CoR, mu) of every material (4 floats each).
 * The backward pass recomputes each window from its checkpoint, holding the contact set of the window fixed, and
 * leaves the system in the final state of the forward run. */
int deme_adjoint_forward(deme_ctx* ctx, int nSteps, int window);
int deme_adjoint_backward(deme_ctx* ctx, const float* dLdPos, const float* dLdVel, float* gradVel0, float* gradMat);

/* Per-contact outputs (contactForces, contactTorque_convToForce, contactPointGeometryA/B): 3 floats each. Kept by
 * deme_force; deme_step stores per-owner terms instead, and a download after it returns DEME_ERR_NOT_READY. */
int deme_download_forces(deme_ctx* ctx, float* force, float* torque_only_force, float* cpA, float* cpB);
/* Contact wildcards delta_tan_x, delta_tan_y, delta_tan_z, delta_time (Models.h:237-252), 4 x C floats. */
int deme_download_wildcards(deme_ctx* ctx, float* w);
int deme_upload_wildcards(deme_ctx* ctx, const float* w);
int deme_get_counts(deme_ctx* ctx, deme_counts* out);
int deme_set_time(deme_ctx* ctx, double t);

/* ---------------------------------------------------------------------------------------------------------------
 * Mid-simulation changes (UpdateClumps, APIPublic.cpp:792-842; ChangeClumpSizes, APIPublic.cpp:844-863 with
 * DEMMiscKernels.cu:4-51).
 * ------------------------------------------------------------------------------------------------------------- */
/* Appends nNewClumps clumps to an initialised system: nNewSph component spheres (sphOwner counts from the first new
 * clump, grouped in ascending order; radii, offsets as 3 floats each, material indices of the uploaded table), each
 * new clump's mass and MOI (3 floats), and its state (voxelID and locX/Y/Z required; the other arrays nullable:
 * identity orientation, zero velocities, family 0). The new spheres take the ids after the existing ones and the new
 * clumps the owner numbers after the existing clumps; the external-object owners move up by nNewClumps. The contact
 * list and its tangential histories are kept: the next step starts with a detection pass, which carries the
 * histories of surviving contacts over. Clump volumes must be set again (deme_set_owner_volumes). Not for batches of
 * scenes (DEME_ERR_UNSUPPORTED). */
int deme_append_clumps(deme_ctx* ctx, uint32_t nNewSph, const uint32_t* sphOwner, const float* sphRadius,
                       const float* sphRelPos, const uint8_t* sphMat, uint32_t nNewClumps, const float* ownerMass,
                       const float* ownerMoi, const deme_owner_state* state);
/* Scales the component offsets and radii of the given clumps (owner numbers below the clump count) by their factors
 * (finite, positive; a repeated id takes its last factor), in place on the device. Mass and MOI are unchanged, as in
 * the reference. The cell grid follows the new largest radius at the next detection pass. */
int deme_change_clump_sizes(deme_ctx* ctx, uint32_t n, const uint32_t* ids, const float* factors);

/* ---------------------------------------------------------------------------------------------------------------
 * Inspectors (DEMInspector::GetValue, AuxClasses.cpp:105-110, over dT::inspectCall, dT.cpp:1660-1720).
 * ------------------------------------------------------------------------------------------------------------- */
/* Volume of each clump (nClumps floats; the template volumes of DEMClumpTemplate::SetVolume), summed by the
 * "clump_volume" inspector. Call after deme_upload_templates. */
int deme_set_owner_volumes(deme_ctx* ctx, const float* volume);
/* Reduces a quantity over the current state into *value: "clump_max_z" (max over spheres of Z + r), "clump_min_z"
 * (min of Z - r), "clump_max_absv" (max of the reference's estimate of a sphere centre's speed) or "clump_volume" (sum
 * over clumps). region: NULL or blank for the whole domain, else a function body over the world position X, Y, Z
 * (float; sphere centres, clump centres for the volume) that returns whether the entity counts, compiled once by
 * hipRTC (AuxClasses.cpp:114-148). Unknown quantities, and regions that do not return or read none of X, Y, Z, give
 * DEME_ERR_INVALID_ARGUMENT; with nothing counted the value is 0 (sum), -FLT_MAX (max) or FLT_MAX (min). Sums are
 * formed in double in a fixed order; every value repeats bit for bit. Synchronises the context's stream. */
int deme_inspect(deme_ctx* ctx, const char* quantity, const char* region, float* value);

/* ---------------------------------------------------------------------------------------------------------------
 * Measurement: per-kernel HIP-event timing on the context's stream (no counterpart in the reference, which only has
 * wall-clock phase timers, src/core/utils/Timer.hpp:40-100). Names are the kernels' names.
 * ------------------------------------------------------------------------------------------------------------- */
int deme_enable_kernel_timing(deme_ctx* ctx, int on);
int deme_reset_kernel_timing(deme_ctx* ctx);
/* Writes up to max entries; returns the number of kernel names in *n. */
int deme_kernel_timing(deme_ctx* ctx, int max, const char** names, double* total_ms, uint64_t* launches, int* n);

#ifdef __cplusplus
}
#endif

#endif
