// Internal declarations shared by the HIP translation units of libdeme (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "deme_device.h"

namespace deme_int {

using deme_dev::BuildPos;
using deme_dev::CdRec;
using deme_dev::DevCounters;
using deme_dev::DEV_COUNTERS_HOST_BYTES;
using deme_dev::SpecGuard;
using deme_dev::spec_skip;
using deme_dev::ERRBIT_CAPACITY;
using deme_dev::ERRBIT_PAIR_OVERFLOW;
using deme_dev::ERRBIT_OUTSIDE;
using deme_dev::ERRBIT_STALE;
using deme_dev::ForceRec;
using deme_dev::KParams;
using deme_dev::OwnerRec;
using deme_dev::PlanePose;
using deme_dev::SphereRec;
using deme_dev::INSPECT_BLOCK;
using deme_dev::INSPECT_CLUMP_MAX_ABSV;
using deme_dev::INSPECT_CLUMP_MAX_Z;
using deme_dev::INSPECT_CLUMP_MIN_Z;
using deme_dev::INSPECT_CLUMP_VOLUME;
using deme_dev::INSPECT_MAX;
using deme_dev::INSPECT_MIN;
using deme_dev::INSPECT_SUM;

// Uniform cell grid over the simulation world used to generate candidate pairs (edge >= largest interaction
// distance, so partners lie in the 27 neighbouring cells).
struct CellGrid {
    double inv;  // 1 / cell edge
    uint32_t nx, ny, nz, nCells;
};

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint64_t SCAN_TILE = (uint64_t)SCAN_BLOCK * SCAN_ITEMS;
// the gathered two-component scan uses half-size tiles: twice the workgroups to cover its gathers' latency
constexpr int SCAN2_ITEMS = 8;
constexpr uint64_t SCAN2_TILE = (uint64_t)SCAN_BLOCK * SCAN2_ITEMS;
// up to this many spheres the candidate-list row scan uses one sphere per lane (tiles of SCAN_BLOCK)
constexpr uint64_t SCAN2_SMALL_N = 65536;
// tile sums the two-component scans need for n elements (both components)
constexpr uint64_t scan2_sums_elems(uint64_t n) {
    return 2 * ((n <= SCAN2_SMALL_N ? (n + SCAN_BLOCK - 1) / SCAN_BLOCK : (n + SCAN2_TILE - 1) / SCAN2_TILE) + 1);
}
constexpr uint32_t MAX_ANAL_OWNERS = 8;
// Plane-owner sums: each workgroup of k_anal_collect folds one tile of ANAL_TILE contacts (256 lanes x ANAL_ITEMS)
// into nAnalOwners x 6 partials, so the partials buffer holds anal_blocks(nC) x nAnalOwners x 6 floats.
constexpr uint32_t ANAL_ITEMS = 4;
constexpr uint32_t ANAL_TILE = 256 * ANAL_ITEMS;
inline uint32_t anal_blocks(uint64_t nC) { return (uint32_t)((nC + ANAL_TILE - 1) / ANAL_TILE); }

// scan.hip: out[0..n] = exclusive prefix sum of (a[idx ? idx[i] : i] + (b ? b[i] : 0)); out[n] = total;
// *totalDev = total (u64). scratch must hold scan_scratch_elems(n) u64 values: one tile sum per tile and level of the
// full-tile path, or one per half-size tile of the folded path, whichever is larger, plus a margin.
constexpr uint64_t scan_scratch_elems(uint64_t n) {
    uint64_t total = 0;
    uint64_t m = n;
    while (true) {
        const uint64_t nb = (m + SCAN_TILE - 1) / SCAN_TILE;
        total += nb + 1;
        if (nb <= 1) break;
        m = nb;
    }
    const uint64_t half = (n + SCAN2_TILE - 1) / SCAN2_TILE + 1;
    return (total > half ? total : half) + 16;
}
hipError_t scan_u32(const uint32_t* a, const uint32_t* b, uint64_t n, uint32_t* out, unsigned long long* totalDev,
                    uint64_t* scratch, hipStream_t st, const uint32_t* idx = nullptr);
// Largest scan of the allocation-free entry point: its tile sums live in a fixed device array of
// scan_scratch_elems(NOALLOC_SCAN_MAX) + 1 u64 values (about 1 MiB) per GPU.
constexpr uint64_t NOALLOC_SCAN_MAX = 1ull << 28;
// Device address of that array on `device`, which must be the current device on the first call for it (the address
// is looked up once per device and kept); *elems (nullable) = its length in u64 values.
hipError_t noalloc_scan_scratch(int device, uint64_t** ptr, uint64_t* elems);
// Two scans over one gather: out1 = scan(a1[idx[i]] + (b1 ? b1[i] : 0)), out2 = scan(a2[idx[i]]), totals at out[n]
// and in *total1 / *total2. tmp: n uint2; sums: scan2_sums_elems(n) u64. n must be at most
// SCAN2_TILE * SCAN_TILE (hipErrorInvalidValue otherwise).
hipError_t scan2_u32_gather(const uint32_t* a1, const uint32_t* b1, const uint32_t* a2, const uint32_t* idx,
                            uint64_t n, uint32_t* out1, uint32_t* out2, unsigned long long* total1,
                            unsigned long long* total2, uint2* tmp, uint64_t* sums, hipStream_t st);
// The candidate-list pass's two row scans with the counts formed in the first pass: out1 = scan(numAnal[i] + touching
// candidates of row i), out2 = scan(touching candidates with partner i); same sizes and limit as scan2_u32_gather.
// tRank[t] of each touching transposed slot t: its rank in its mirror row (other slots are not written).
// When ctr->err is set (the pass failed and is discarded) pass 1 does nothing and the outputs are meaningless.
hipError_t scan2_cand_rows(const uint32_t* candStart, const uint32_t* tStart, const uint32_t* tC, const uint8_t* acc,
                           const uint32_t* numAnal, uint64_t n, uint32_t* tRank, uint32_t* out1, uint32_t* out2,
                           unsigned long long* total1, unsigned long long* total2, uint2* tmp, uint64_t* sums,
                           const DevCounters* ctr, hipStream_t st);

// cd.hip
hipError_t launch_plane_pose(const KParams& P, const deme_anal_obj* anal, uint32_t nA, const OwnerRec* own,
                             PlanePose* pose, hipStream_t st);
// The plane poses and, in the same launch, nClear words of `clear` zeroed (a pass's counters).
hipError_t launch_pass_start(const KParams& P, const deme_anal_obj* anal, uint32_t nA, const OwnerRec* own,
                             PlanePose* pose, uint32_t* clear, uint32_t nClear, hipStream_t st);
hipError_t launch_sphere_bins(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                              const PlanePose* pose, uint32_t nA, const uint8_t* mask, uint32_t* binCount,
                              uint16_t* numBins, uint32_t* numAnal, CdRec* rec, DevCounters* ctr, hipStream_t st);
hipError_t launch_bin_stats(const uint32_t* binCount, uint64_t nBins, DevCounters* ctr, hipStream_t st);
hipError_t launch_bin_scatter(const KParams& P, const CdRec* rec, const SphereRec* sph, uint32_t nS,
                              const uint32_t* binStart, uint32_t* binCount, uint2* binEnt, const DevCounters* ctr,
                              hipStream_t st);
// bCnt: per-sphere count of contacts in which the sphere is body B (sizes the mirror rows)
hipError_t launch_bin_pairs(const KParams& P, const uint2* binEnt, uint64_t nEnt, const uint32_t* binStart,
                            const CdRec* rec, const SphereRec* sph, const OwnerRec* own, const uint8_t* mask,
                            uint2* pairs, uint32_t pairCap, DevCounters* ctr, uint32_t* rowCount, uint32_t* bCnt,
                            hipStream_t st);
hipError_t launch_row_fill(const uint2* pairs, uint32_t nPairs, const uint32_t* rowStart, const uint32_t* numAnal,
                           uint32_t* rowCount, uint32_t* idB, uint8_t* ctype, hipStream_t st);
hipError_t launch_geo_fill(const KParams& P, const CdRec* rec, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                           const PlanePose* pose, uint32_t nA, const uint8_t* mask, const uint32_t* numAnal,
                           const uint32_t* rowStart, uint32_t* idB, uint8_t* ctype, hipStream_t st);
hipError_t launch_row_finalize(uint32_t nS, const uint32_t* rowStart, const uint32_t* numAnal, uint32_t* idA,
                               uint32_t* idB, hipStream_t st);
// map: persistent-contact map to the previous list; mirror: slot of each sphere-sphere contact in its B sphere's
// mirror row (bRowStart / bPart: per sphere, the A ids of its contacts as body B, ascending)
hipError_t launch_history_map(uint32_t nS, const uint32_t* rowStart, const uint32_t* idB, const uint8_t* ctype,
                              const uint32_t* prevRowStart, const uint32_t* prevIdB, const uint8_t* prevType,
                              uint32_t* map, const uint32_t* bRowStart, const uint32_t* bPart, uint32_t* mirror,
                              hipStream_t st, SpecGuard guard = SpecGuard());
// mirror rows from a finished contact list (counts in bCnt, consumed)
hipError_t launch_brow_fill(uint32_t nC, uint32_t nS, const uint32_t* idA, const uint32_t* idB, const uint8_t* ctype,
                            const uint32_t* bRowStart, uint32_t* bCnt, uint32_t* bPart, hipStream_t st);
hipError_t launch_brow_sort(uint32_t nS, const uint32_t* bRowStart, uint32_t* bPart, hipStream_t st);

// cd_fast.hip
hipError_t launch_cell_assign(const KParams& P, const CellGrid& G, const SphereRec* sph, const OwnerRec* own,
                              uint32_t nS, const PlanePose* pose, uint32_t nA, const uint8_t* mask,
                              uint32_t* cellCount, uint32_t* cellOf, uint32_t* slotOf, uint16_t* numBins,
                              uint32_t* numAnal, unsigned long long* planeBits, DevCounters* ctr, hipStream_t st);
// sorted: the CD records in cell order (positions recomputed from sph / own as the assignment computed them);
// sortedMeta: (owner, family) in cell order; posOf: each sphere's sorted position; aCnt / bCnt (sorted order) are
// cleared for the pair pass
hipError_t launch_cell_scatter(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                               const uint32_t* cellOf, const uint32_t* slotOf, const uint32_t* cellStart,
                               CdRec* sorted, uint2* sortedMeta, uint32_t* posOf, uint32_t* aCnt, uint32_t* bCnt,
                               float beta, float4* sortedF, hipStream_t st);
// Slabs are column-major by sorted position (entry k of sorted sphere i at slab[k * nS + i]). surv: prefilter
// survivors; rowSlab / aCnt: higher-id partners (contact row); mirSlab / bCnt: lower-id partners (mirror row).
// sortedF: float copies of the sorted records (position, expanded radius); margin: bound on the float rounding of a
// pair's distance, so the float test only rejects pairs the exact test would reject. mode (sortedF needed for 1, 2):
// 0 the exact test on every candidate; 1 a float pre-reject, the exact test on its survivors in the sweep; 2 the float
// test alone in the sweep, the same-owner rule and the exact test in the block's shared second phase.
constexpr int PAIR_MODE_CAND = 3;
// mode 3 (PAIR_MODE_CAND) builds candidate lists: pairs of different owners within `skin` of touching (d <= r'A + r'B + skin), no mask
// and no bin rule. block: workgroup size, 64, 128 or 256 (hipErrorInvalidValue otherwise).
hipError_t launch_cell_pairs(const KParams& P, const CellGrid& G, uint32_t nS, const CdRec* sorted,
                             const float4* sortedF, float margin, int mode, uint32_t block, const uint2* sortedMeta,
                             const uint32_t* cellStart, const uint8_t* mask, uint32_t slabCap,
                             uint32_t* surv, uint32_t* rowSlab, uint32_t* mirSlab, uint32_t* aCnt, uint32_t* bCnt,
                             DevCounters* ctr, hipStream_t st,
                             double skin = 0.0);
hipError_t launch_cell_rows(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                            const uint32_t* rowSlab, const uint32_t* mirSlab, const uint32_t* posOf,
                            const uint32_t* rowStart, const unsigned long long* planeBits, const uint32_t* bRowStart,
                            uint32_t* idA, uint32_t* idB, uint8_t* ctype, uint32_t* bPart, hipStream_t st,
                            SpecGuard guard = SpecGuard());
// Candidate-list path. Per step: detection records in id order (pos) with (owner, family) (meta), the shared
// per-sphere detection outputs, and the largest squared displacement since the lists were built into ctr->disp2.
// The largest squared displacement since the build is taken relative to the translation uIn (3 doubles; any vector
// keeps the candidate lists exact); the displacements' extremes per axis are folded into ctr for the next pass's.
hipError_t launch_sphere_pos(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                             const PlanePose* pose, uint32_t nA, const uint8_t* mask, CdRec* pos,
                             const BuildPos* pos0, double unit, uint2* meta, uint16_t* numBins, uint32_t* numAnal,
                             unsigned long long* planeBits, DevCounters* ctr, const double* uIn, hipStream_t st);
// After a candidate build: per sphere its candidate row (candStart from a scan of the build's row counts; candA /
// candB the row's sphere and partners) and its position at the build (pos0, in fixed point: coordinate * scale
// rounded, scale a power of two).
hipError_t launch_cand_rows(uint32_t nS, const uint32_t* rowSlab, const uint32_t* candStart, uint32_t* candA,
                            uint32_t* candB, const CdRec* sorted, BuildPos* pos0, double scale, hipStream_t st);
// Per step, per candidate: the reference's predicate (acc[c] = 1 when c touches); ERRBIT_STALE instead when
// ctr->disp2 exceeds dispLimit2. Lane 0 also writes the next pass's common translation (3 doubles) to uOut.
hipError_t launch_cand_flags(const KParams& P, uint32_t nK, const CdRec* pos, const uint2* meta, const uint32_t* candA,
                             const uint32_t* candB, const uint8_t* mask, double dispLimit2, uint8_t* acc,
                             DevCounters* ctr, double* uOut, hipStream_t st);
// Per sphere: aCnt = touching candidates of its row, bCnt = touching candidates with it as partner; tRank as in
// scan2_cand_rows.
hipError_t launch_cand_count(uint32_t nS, const uint32_t* candStart, const uint32_t* tStart, const uint32_t* tC,
                             const uint8_t* acc, uint32_t* aCnt, uint32_t* bCnt, uint32_t* tRank,
                             const DevCounters* ctr, hipStream_t st);
// Per touching candidate: its canonical-row entry, its mirror link (mirror[contact] = bRowStart[B] + tRank[tPos[c]]),
// cidx[c] and, with prevAcc / prevCidx (the previous pass over the same candidates), its persistent-contact map entry.
hipError_t launch_cand_fill(uint32_t nK, const uint32_t* candA, const uint32_t* candB, const uint32_t* candStart,
                            const uint8_t* acc, const uint32_t* rowStart, const uint32_t* numAnal,
                            const uint32_t* tPos, const uint32_t* tRank, const uint32_t* bRowStart,
                            const uint8_t* prevAcc, const uint32_t* prevCidx, uint32_t* cidx, uint32_t* idA,
                            uint32_t* idB, uint8_t* ctype, uint32_t* map, uint32_t* mirror, hipStream_t st,
                            SpecGuard guard = SpecGuard());
// Per sphere: its plane contacts at the head of its canonical row (with their map entries when mapPlanes; no mirror
// slot).
hipError_t launch_cand_planes(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                              const uint32_t* rowStart, const unsigned long long* planeBits,
                              const uint32_t* prevRowStart, const uint32_t* prevIdB, const uint8_t* prevType,
                              bool mapPlanes, uint32_t* idA, uint32_t* idB, uint8_t* ctype, uint32_t* map,
                              uint32_t* mirror, hipStream_t st, SpecGuard guard = SpecGuard());
// After a build, the transpose from the rows: slot[c] = c's place among the candidates with its partner (tCnt counts
// them, cleared by the caller); then, with tStart scanned from tCnt, tC[tStart[B] + slot[c]] = c; then every
// transposed row sorted ascending (ascending A), and each candidate's transposed slot recorded (tPos[tC[t]] = t).
hipError_t launch_cand_tcount(uint32_t nK, const uint32_t* candB, uint32_t* tCnt, uint32_t* slot, hipStream_t st);
hipError_t launch_cand_tscatter(uint32_t nK, const uint32_t* candB, const uint32_t* slot, const uint32_t* tStart,
                                uint32_t* tC, hipStream_t st);
hipError_t launch_cand_tsort(uint32_t nS, const uint32_t* tStart, uint32_t* tC, uint32_t* tPos, hipStream_t st);

// dyn.hip
// aTerm / gTerm: 6 floats per contact; bTerm: 6 floats per mirror entry; out (optional): full contact records
hipError_t launch_force(const KParams& P, uint32_t nC, const uint32_t* idA, const uint32_t* idB, const uint8_t* ctype,
                        const SphereRec* sph, const OwnerRec* own, const float4* moi, const deme_anal_obj* anal,
                        const PlanePose* pose, const deme_material* mats, const float4* histIn, const uint32_t* map,
                        float4* histOut, const uint32_t* mirror, float* aTerm, float* bTerm, float* gTerm,
                        uint8_t* aOn, uint8_t* bOn,
                        ForceRec* out, hipStream_t st, SpecGuard guard = SpecGuard());
// blockPartials holds partialCap floats; a launch that would need more (anal_blocks(nC) * nAnalOwners * 6) returns
// hipErrorInvalidValue and writes nothing. With a guard (a step captured with its detection pass) nC is the capacity
// the grid covers and the kernels take the count from the pass's counters.
hipError_t launch_anal_collect(uint32_t nC, const uint32_t* idB, const uint8_t* ctype, const float* gTerm,
                               const deme_anal_obj* anal, uint32_t nA, uint32_t firstAnalOwner, uint32_t nAnalOwners,
                               float* blockPartials, uint64_t partialCap, uint32_t* nBlocksOut, hipStream_t st,
                               SpecGuard guard = SpecGuard());
hipError_t launch_anal_final(const float* blockPartials, uint32_t nBlocks, uint32_t nAnalOwners, float* analAcc,
                             hipStream_t st, SpecGuard guard = SpecGuard());
hipError_t launch_owner_update(const KParams& P, uint32_t nO, uint32_t nClumps, uint32_t firstAnalOwner,
                               uint32_t nAnalOwners, const uint32_t* ownerSphStart, const uint32_t* rowStart,
                               const uint32_t* bRowStart, const float* aTerm, const float* bTerm,
                               const uint8_t* aOn, const uint8_t* bOn,
                               const float* analAcc, const deme_family_rule* rules, OwnerRec* own,
                               float* acc /* 6 x nO SoA: aX aY aZ alX alY alZ */, int doIntegrate,
                               const double* clock /* device: the simulation time the motion rules read */,
                               hipStream_t st, OwnerRec* ownOut = nullptr /* integrated records; null: in place */,
                               SpecGuard guard = SpecGuard());
// clock[0] = t (one lane; the time the owner update's motion rules read, written on the stream before the update)
hipError_t launch_set_clock(double* clock, double t, hipStream_t st);
// clock[0] += (double)h: the time step the host adds after each step, done on the device inside a captured step
hipError_t launch_advance_clock(double* clock, float h, hipStream_t st);
// p[0..n) = v, and p2[0..n2) = v in the same launch (n2 <= 256 * blocks: small second ranges only)
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st, uint32_t* p2 = nullptr,
                           uint32_t n2 = 0);
hipError_t launch_change_family(OwnerRec* own, uint32_t nO, uint8_t from, uint8_t to, hipStream_t st);

// jit_rt.cpp: the force kernel specialised at run time by hipRTC
struct JitForce {
    hipModule_t mod = nullptr;
    hipFunction_t full = nullptr, terms = nullptr;
    void release();
    hipError_t load(const std::vector<char>& code);
    hipError_t launch(bool wantFull, const KParams& P, uint32_t nC, const uint32_t* idA, const uint32_t* idB,
                      const uint8_t* ctype, const SphereRec* sph, const OwnerRec* own, const float4* moi,
                      const deme_anal_obj* anal, const PlanePose* pose, const deme_material* mats,
                      const float4* histIn, const uint32_t* map, float4* histOut, const uint32_t* mirror,
                      float* aTerm, float* bTerm, float* gTerm, uint8_t* aOn, uint8_t* bOn, ForceRec* out,
                      hipStream_t st, SpecGuard guard = SpecGuard()) const;
};
// the owner update kernel specialised with motion strategies / family change rules
struct JitOwner {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    void release();
    hipError_t load(const std::vector<char>& code);
    hipError_t launch(const KParams& P, uint32_t nO, uint32_t nClumps, uint32_t firstAnal, uint32_t nAnalOwners,
                      const uint32_t* ownerSphStart, const uint32_t* rowStart, const uint32_t* bRowStart,
                      const float* aTerm, const float* bTerm, const uint8_t* aOn, const uint8_t* bOn,
                      const float* analAcc, const deme_family_rule* rules,
                      const OwnerRec* own, OwnerRec* ownOut, float* acc, int doIntegrate, const double* clock,
                      hipStream_t st, SpecGuard guard = SpecGuard()) const;
};
std::string jit_owner_source(const std::vector<std::pair<std::string, std::string>>& values);

// Inspection kernels (query.hip; regions through hipRTC). inspect_blocks(n) partials, then one block folds them.
uint32_t inspect_blocks(uint32_t n);
hipError_t launch_inspect(const KParams& P, int quantity, int flavor, uint32_t n, const SphereRec* sph,
                          const OwnerRec* own, const float* vol, const float lbf[3], double* partial, hipStream_t st);
hipError_t launch_inspect_final(int flavor, uint32_t nPart, const double* partial, double* out, hipStream_t st);
struct JitInspect {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    void release();
    hipError_t load(const std::vector<char>& code);
    hipError_t launch(const KParams& P, int quantity, int flavor, uint32_t n, const SphereRec* sph,
                      const OwnerRec* own, const float* vol, const float lbf[3], double* partial,
                      hipStream_t st) const;
};
std::string jit_inspect_source(const std::string& region);
// ChangeClumpSizes: scale the offsets and radii of the spheres of clumps with mark[owner] != 0 by factor[owner]
hipError_t launch_scale_components(SphereRec* sph, uint32_t n, const uint8_t* mark, const float* factor,
                                   hipStream_t st);
bool jit_is_owner_key(const std::string& key);
// built-in model bodies: 0 frictional Hertz-Mindlin (the ahead-of-time kernel's), 1 frictionless Hertz
const char* jit_builtin_model(int which);
bool jit_force_source(const std::string& model, const std::vector<std::string>& wildcards, const deme_sim_params* p,
                      const std::vector<std::pair<std::string, std::string>>& overrides, std::string& src,
                      std::string& err);
int jit_compile(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& log);

// adj.hip (adjoint of the step for single-sphere owners; lam: 9 x nO SoA adjoints of X, v, w)
uint32_t adj_force_blocks(uint32_t nC);
hipError_t launch_adj_owner_pre(uint32_t nO, uint32_t nClumps, float h, const OwnerRec* ownSnap, const float* accSnap,
                                float* lam, float* lAcc, hipStream_t st);
hipError_t launch_adj_force(const KParams& P, uint32_t nC, uint32_t nClumps, const uint32_t* idA, const uint32_t* idB,
                            const uint8_t* ctype, const SphereRec* sph, const OwnerRec* ownSnap, const float4* moi,
                            const deme_anal_obj* anal, const PlanePose* pose, const deme_material* mats, uint32_t nMat,
                            const float4* histIn, const uint32_t* mirror, uint32_t nO, const float* lAcc,
                            const float4* lDtOut, float* adjA, float* adjB, float4* lDtIn, float* matPart,
                            int accumulate, hipStream_t st);
hipError_t launch_adj_owner_post(uint32_t nO, uint32_t nClumps, const uint32_t* ownerSphStart, const uint32_t* rowStart,
                                 const uint32_t* bRowStart, const float* adjA, const float* adjB, float* lam,
                                 float h, const OwnerRec* ownPrev, const float* accPrev, float* lAcc, hipStream_t st);
hipError_t launch_adj_mat_final(const float* matPart, uint32_t nBlocks, uint32_t n, float* grad, hipStream_t st);
hipError_t launch_adj_gather_hist(uint32_t nC, const uint32_t* map, uint32_t nPrev, const float4* prev, float4* out,
                                  hipStream_t st);
hipError_t launch_adj_scatter_hist(uint32_t nC, const uint32_t* map, uint32_t nPrev, const float4* lIn, float4* lPrev,
                                   hipStream_t st);
hipError_t launch_fill_f4(float4* p, uint64_t n, hipStream_t st);

}  // namespace deme_int
