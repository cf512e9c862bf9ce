// Contact detection for gfx950: sphere -> bin assignment, occupancy scan, per-bin sphere pair detection, canonical
// contact rows and the contact-history map.
//
// Reference: src/algorithms/DEMCubContactDetection.cu:39-397 and the kernels it launches
// (DEMBinSphereKernels.cu, DEMContactKernels_Blockwise.cu, DEMHistoryMappingKernels.cu).
//
// Layout choices (MI355X):
//  * Bins are a counting sort: one atomic increment per (sphere, bin) touch builds the occupancy histogram, an
//    exclusive scan turns it into bin starts, and a second pass scatters (sphere, bin) entries. The scatter consumes
//    the histogram back to zero, so no clearing pass is needed before the next detection.
//  * Pair detection runs one lane per bin entry over the bin-sorted entries. A 256-entry tile of per-sphere records
//    (double position, float CD radius, id) is staged in LDS; a lane tests its entry against every later entry of the
//    same bin. Hits are appended with one atomic per wave (ballot + popcount).
//  * The unordered hits become canonical rows per sphere A (plane contacts first, then partners by ascending id), so
//    the result is independent of atomic ordering.
#include "deme_internal.h"

namespace deme_int {
using namespace deme_dev;
namespace {

struct BinRange {
    uint32_t x0, x1, y0, y1, z0, z1;
};

// DEMBinSphereKernels.cu:52-64 (double position and radius span, truncating conversions)
__device__ __forceinline__ BinRange bin_range(const KParams& P, double X, double Y, double Z, double rCD) {
    const double bx = X / P.binSize, by = Y / P.binSize, bz = Z / P.binSize, span = rCD / P.binSize;
    BinRange r;
    r.x0 = sat_u32(bx - span);
    r.x1 = sat_u32(bx + span);
    r.y0 = sat_u32(by - span);
    r.y1 = sat_u32(by + span);
    r.z0 = sat_u32(bz - span);
    r.z1 = sat_u32(bz + span);
    return r;
}

// getPointBinID<binID_t> (DEMHelperKernels.cu:266-273)
__device__ __forceinline__ uint32_t point_bin(const KParams& P, double X, double Y, double Z) {
    const uint32_t ix = sat_u32(X / P.binSize), iy = sat_u32(Y / P.binSize), iz = sat_u32(Z / P.binSize);
    return ix + iy * P.nbX + iz * P.nbX * P.nbY;
}

// checkSphereEntityOverlap<double> without contact point (DEMHelperKernels.cu:356-391); rCD already holds r + beta.
__device__ __forceinline__ bool plane_cd(const KParams& P, const PlanePose& pp, double X, double Y, double Z,
                                         double rCD) {
    if (pp.type != DEME_ANAL_OBJ_TYPE_PLANE) return false;
    const double dx = X - pp.px, dy = Y - pp.py, dz = Z - pp.pz;
    const double dist = dx * (double)pp.nx + dy * (double)pp.ny + dz * (double)pp.nz;
    return !(dist > rCD + (double)P.beta);
}

__global__ __launch_bounds__(64) void k_plane_pose(KParams P, const deme_anal_obj* __restrict__ anal, uint32_t nA,
                                                   const OwnerRec* __restrict__ own, PlanePose* __restrict__ pose,
                                                   uint32_t* __restrict__ clear, uint32_t nClear) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    // the first workgroup also clears nClear words (a detection pass's counters: one launch instead of two)
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < nClear; k += blockDim.x) clear[k] = 0u;
    if (a >= nA) return;
    const deme_anal_obj ob = anal[a];
    const OwnerRec o = load_owner(own, ob.owner);
    double ox, oy, oz;
    owner_position(P, o.voxel, o.lx, o.ly, o.lz, ox, oy, oz);
    float px = ob.relPos[0], py = ob.relPos[1], pz = ob.relPos[2];
    float nx = ob.rot[0], ny = ob.rot[1], nz = ob.rot[2];
    rotate(px, py, pz, o.qw, o.qx, o.qy, o.qz);
    rotate(nx, ny, nz, o.qw, o.qx, o.qy, o.qz);
    PlanePose pp;
    pp.px = ox + (double)px;
    pp.py = oy + (double)py;
    pp.pz = oz + (double)pz;
    pp.nx = nx;
    pp.ny = ny;
    pp.nz = nz;
    pp.owner = ob.owner;
    pp.fam = o.fam;
    pp.type = ob.type;
    pose[a] = pp;
}

// Rows a6 + a10 (+ histogram of a8): getNumberOfBinsEachSphereTouches (DEMBinSphereKernels.cu:12-115).
__global__ __launch_bounds__(256) void k_sphere_bins(KParams P, const SphereRec* __restrict__ sph,
                                                     const OwnerRec* __restrict__ own, uint32_t nS,
                                                     const PlanePose* __restrict__ pose, uint32_t nA,
                                                     const uint8_t* __restrict__ mask, uint32_t* binCount,
                                                     uint16_t* __restrict__ numBins, uint32_t* __restrict__ numAnal,
                                                     CdRec* __restrict__ rec, DevCounters* ctr) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const SphereRec sr = load_sphere(sph, s);
    const OwnerRec o = load_owner(own, sr.owner);
    double X, Y, Z;
    sphere_position(P, o, sr, X, Y, Z);
    double rCD = (double)sr.r;
    rCD += (double)P.beta;
    const BinRange r = bin_range(P, X, Y, Z, rCD);
    const uint32_t nx = r.x1 - r.x0 + 1, ny = r.y1 - r.y0 + 1, nz = r.z1 - r.z0 + 1;
    const uint64_t full = (uint64_t)nx * ny * nz;
    numBins[s] = (uint16_t)((int)(uint16_t)nx * (int)(uint16_t)ny * (int)(uint16_t)nz);
    if (full > 65535u) {
        atomicOr(&ctr->err, ERRBIT_CAPACITY);
    } else {
        // Linear ids are formed in 32-bit arithmetic exactly as the reference does; an id past the grid means the
        // sphere left the simulation world, which is reported instead of indexing past the histogram.
        // loops count touches (<= 65535 in total) rather than compare against a bound that may be 2^32 - 1
        const uint32_t nbX = P.nbX, nbXY = P.nbX * P.nbY;
        for (uint32_t dk = 0; dk < nz; dk++)
            for (uint32_t dj = 0; dj < ny; dj++)
                for (uint32_t di = 0; di < nx; di++) {
                    const uint32_t bin = (r.x0 + di) + (r.y0 + dj) * nbX + (r.z0 + dk) * nbXY;
                    if (bin < P.nBins)
                        atomicAdd(&binCount[bin], 1u);
                    else
                        atomicOr(&ctr->err, ERRBIT_OUTSIDE);
                }
    }
    uint32_t cnt = 0;
    const uint32_t famS = o.fam;
    uint32_t a0, a1;
    scene_planes(P, o.scene, a0, a1);
    for (uint32_t a = a0; a < a1 && a < nA; a++) {
        const PlanePose pp = pose[a];
        if (mask[mask_index(famS, pp.fam)] != DEME_DONT_PREVENT_CONTACT) continue;
        if (plane_cd(P, pp, X, Y, Z, rCD)) cnt++;
    }
    numAnal[s] = cnt;
    float rf = sr.r;
    rf += P.beta;
    CdRec cr;
    cr.x = X;
    cr.y = Y;
    cr.z = Z;
    cr.r = rf;
    cr.id = s;
    rec[s] = cr;
}

// Number of bins touched by at least one sphere (B_a) and the maximum occupancy, from the histogram.
__global__ __launch_bounds__(256) void k_bin_stats(const uint32_t* __restrict__ cnt, uint64_t n, DevCounters* ctr) {
    uint32_t act = 0, mx = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = cnt[i];
        act += (c > 0);
        mx = c > mx ? c : mx;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        act += __shfl_down(act, d, 64);
        const uint32_t o = __shfl_down(mx, d, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        if (act) atomicAdd(&ctr->nActive, (unsigned long long)act);
        atomicMax(&ctr->maxOcc, mx);
    }
}

// Row a8 scatter: every (sphere, bin) touch goes to binStart[bin] + slot. Order within a bin is arbitrary; pair
// detection and row canonicalisation make the final contact list independent of it.
__global__ __launch_bounds__(256) void k_bin_scatter(KParams P, const CdRec* __restrict__ rec,
                                                     const SphereRec* __restrict__ sph, uint32_t nS,
                                                     const uint32_t* __restrict__ binStart, uint32_t* binCount,
                                                     uint2* __restrict__ binEnt, const DevCounters* ctr) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || ctr->err) return;
    const CdRec cr = load_cd(rec, s);
    double rCD = (double)sph[s].r;
    rCD += (double)P.beta;
    const BinRange r = bin_range(P, cr.x, cr.y, cr.z, rCD);
    const uint32_t nx = r.x1 - r.x0 + 1, ny = r.y1 - r.y0 + 1, nz = r.z1 - r.z0 + 1;
    const uint32_t nbX = P.nbX, nbXY = P.nbX * P.nbY;
    // k, j, i order (DEMBinSphereKernels.cu:170-183); the counting pass already bounded nx * ny * nz
    for (uint32_t dk = 0; dk < nz; dk++)
        for (uint32_t dj = 0; dj < ny; dj++)
            for (uint32_t di = 0; di < nx; di++) {
                const uint32_t bin = (r.x0 + di) + (r.y0 + dj) * nbX + (r.z0 + dk) * nbXY;
                if (bin >= P.nBins) continue;
                const uint32_t slot = atomicSub(&binCount[bin], 1u) - 1u;
                binEnt[binStart[bin] + slot] = make_uint2(s, bin);
            }
}

// checkSpheresOverlap<double> (DEMHelperKernels.cu:195-222) for the canonical (lower id, higher id) order, then
// the contact-point-in-this-bin rule (DEMContactKernels_Blockwise.cu:250-256).
__device__ __forceinline__ bool pair_in_bin(const KParams& P, const CdRec& A, const CdRec& B, uint32_t bin) {
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double ra = (double)A.r, rb = (double)B.r;
    if (d2 > (ra + rb) * (ra + rb)) return false;
    float nX = (float)dx, nY = (float)dy, nZ = (float)dz;
    const float mag = sqrtf(nX * nX + nY * nY + nZ * nZ);
    nX /= mag;
    nY /= mag;
    nZ /= mag;
    const double half = (ra + rb - sqrt(d2)) / 2.0;
    const double cx = B.x + (rb - half) * (double)nX;
    const double cy = B.y + (rb - half) * (double)nY;
    const double cz = B.z + (rb - half) * (double)nZ;
    return point_bin(P, cx, cy, cz) == bin;
}

// Row a9: per-bin pair detection (DEMContactKernels_Blockwise.cu:15-309). One lane per bin entry.
__global__ __launch_bounds__(256) void k_bin_pairs(KParams P, const uint2* __restrict__ binEnt, uint64_t nEnt,
                                                   const uint32_t* __restrict__ binStart, const CdRec* __restrict__ rec,
                                                   const SphereRec* __restrict__ sph, const OwnerRec* __restrict__ own,
                                                   const uint8_t* __restrict__ mask, uint2* __restrict__ pairs,
                                                   uint32_t pairCap, DevCounters* ctr, uint32_t* rowCount,
                                                   uint32_t* bCnt) {
    __shared__ CdRec tile[256];
    const uint64_t base = (uint64_t)blockIdx.x * 256u;
    const uint64_t e = base + threadIdx.x;
    const uint64_t tileEnd = (base + 256u < nEnt) ? base + 256u : nEnt;
    const int lane = threadIdx.x & 63;
    uint32_t myBin = 0, nIter = 0;
    CdRec me;
    me.x = me.y = me.z = 0.0;
    me.r = 0.f;
    me.id = 0;
    if (e < nEnt) {
        const uint2 be = binEnt[e];
        me = load_cd(rec, be.x);
        myBin = be.y;
        nIter = binStart[myBin + 1] - (uint32_t)e - 1u;
    }
    tile[threadIdx.x] = me;
    __syncthreads();
    for (uint32_t j = 1;; j++) {
        const bool active = j <= nIter;
        if (!__any(active)) break;
        bool hit = false;
        uint32_t ia = 0, ib = 0;
        if (active) {
            const uint64_t ep = e + j;
            const CdRec ot = (ep < tileEnd) ? tile[ep - base] : load_cd(rec, binEnt[ep].x);
            const bool meFirst = me.id < ot.id;
            const CdRec& A = meFirst ? me : ot;
            const CdRec& B = meFirst ? ot : me;
            if (pair_in_bin(P, A, B, myBin)) {
                const uint32_t oa = sph[A.id].owner, obb = sph[B.id].owner;
                if (oa != obb) {
                    const uint32_t fa = own[oa].fam, fb = own[obb].fam;
                    if (mask[mask_index(fa, fb)] == DEME_DONT_PREVENT_CONTACT) {
                        hit = true;
                        ia = A.id;
                        ib = B.id;
                    }
                }
            }
        }
        const uint64_t bal = __ballot(hit);
        if (bal) {
            const int leader = __ffsll((unsigned long long)bal) - 1;
            uint32_t first = 0;
            if (lane == leader) first = atomicAdd(&ctr->pairCount, (uint32_t)__popcll(bal));
            first = __shfl(first, leader, 64);
            if (hit) {
                const uint32_t idx = first + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                if (idx < pairCap) pairs[idx] = make_uint2(ia, ib);
                atomicAdd(&rowCount[ia], 1u);
                atomicAdd(&bCnt[ib], 1u);
            }
        }
    }
}

// Row a11: sphere-sphere partners into their rows (after plane contacts). Consumes rowCount back to zero.
__global__ __launch_bounds__(256) void k_row_fill(const uint2* __restrict__ pairs, uint32_t nPairs,
                                                  const uint32_t* __restrict__ rowStart,
                                                  const uint32_t* __restrict__ numAnal, uint32_t* rowCount,
                                                  uint32_t* __restrict__ idB, uint8_t* __restrict__ ctype) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPairs) return;
    const uint2 pr = pairs[p];
    const uint32_t slot = atomicSub(&rowCount[pr.x], 1u) - 1u;
    const uint32_t pos = rowStart[pr.x] + numAnal[pr.x] + slot;
    idB[pos] = pr.y;
    ctype[pos] = DEME_SPHERE_SPHERE_CONTACT;
}

// Row a10: sphere-plane contacts in ascending plane order at the head of each row
// (populateBinSphereTouchingPairs, DEMBinSphereKernels.cu:186-228).
__global__ __launch_bounds__(256) void k_geo_fill(KParams P, const CdRec* __restrict__ rec,
                                                  const SphereRec* __restrict__ sph, const OwnerRec* __restrict__ own,
                                                  uint32_t nS, const PlanePose* __restrict__ pose, uint32_t nA,
                                                  const uint8_t* __restrict__ mask, const uint32_t* __restrict__ numAnal,
                                                  const uint32_t* __restrict__ rowStart, uint32_t* __restrict__ idB,
                                                  uint8_t* __restrict__ ctype) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || numAnal[s] == 0) return;
    const CdRec cr = load_cd(rec, s);
    const SphereRec sr = load_sphere(sph, s);
    double rCD = (double)sr.r;
    rCD += (double)P.beta;
    const OwnerRec o = load_owner(own, sr.owner);
    const uint32_t famS = o.fam;
    uint32_t pos = rowStart[s];
    uint32_t a0, a1;
    scene_planes(P, o.scene, a0, a1);
    for (uint32_t a = a0; a < a1 && a < nA; a++) {
        const PlanePose pp = pose[a];
        if (mask[mask_index(famS, pp.fam)] != DEME_DONT_PREVENT_CONTACT) continue;
        if (plane_cd(P, pp, cr.x, cr.y, cr.z, rCD)) {
            idB[pos] = a;
            ctype[pos] = DEME_SPHERE_PLANE_CONTACT;
            pos++;
        }
    }
}

// Canonical order inside each row: sphere partners by ascending id; idA filled for the whole row.
__global__ __launch_bounds__(256) void k_row_finalize(uint32_t nS, const uint32_t* __restrict__ rowStart,
                                                      const uint32_t* __restrict__ numAnal, uint32_t* __restrict__ idA,
                                                      uint32_t* __restrict__ idB) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const uint32_t st = rowStart[s], en = rowStart[s + 1], sp = st + numAnal[s];
    for (uint32_t i = sp + 1; i < en; i++) {
        const uint32_t v = idB[i];
        uint32_t k = i;
        while (k > sp && idB[k - 1] > v) {
            idB[k] = idB[k - 1];
            k--;
        }
        idB[k] = v;
    }
    for (uint32_t i = st; i < en; i++) idA[i] = s;
}

// Row a12: persistent-contact map (buildPersistentMap, DEMHistoryMappingKernels.cu:16-60) as a merge join of the
// new and previous rows of the same sphere (both ordered by (type group, idB)). The same pass links every
// sphere-sphere contact (s, b) to its slot in b's mirror row (the contacts where b is the B body, ordered by A id):
// mirror[i] = bRowStart[b] + rank of s in that row (skipped when mirror is null: the map alone).
// Rows up to this long take the register path of the history map.
constexpr uint32_t HM_REG = 8;

__global__ __launch_bounds__(256) void k_history_map(uint32_t nS, const uint32_t* __restrict__ rowStart,
                                                     const uint32_t* __restrict__ idB,
                                                     const uint8_t* __restrict__ ctype,
                                                     const uint32_t* __restrict__ prevRowStart,
                                                     const uint32_t* __restrict__ prevIdB,
                                                     const uint8_t* __restrict__ prevType, uint32_t* __restrict__ map,
                                                     const uint32_t* __restrict__ bRowStart,
                                                     const uint32_t* __restrict__ bPart,
                                                     uint32_t* __restrict__ mirror, SpecGuard guard) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || spec_skip(guard)) return;
    const uint32_t ns = rowStart[s], ne = rowStart[s + 1];
    uint32_t j = prevRowStart ? prevRowStart[s] : 0;
    const uint32_t je = prevRowStart ? prevRowStart[s + 1] : 0;
    if (ne - ns <= HM_REG && je - j <= HM_REG) {
        // short rows (nearly all): both rows loaded at once and matched in registers. A row's keys (plane or sphere,
        // partner id) are distinct, so the entry of the previous row with the same partner and type is the merge's.
        const uint32_t nn = ne - ns, np = je - j;
        uint32_t nb[HM_REG], pb[HM_REG];
        uint8_t nt[HM_REG], pt[HM_REG];
#pragma unroll
        for (uint32_t k = 0; k < HM_REG; k++) {
            nb[k] = k < nn ? idB[ns + k] : 0u;
            nt[k] = k < nn ? ctype[ns + k] : (uint8_t)DEME_NOT_A_CONTACT;
            pb[k] = k < np ? prevIdB[j + k] : 0u;
            pt[k] = k < np ? prevType[j + k] : (uint8_t)0xFF;
        }
        // the partners' mirror-row bounds, all in flight
        uint32_t lo[HM_REG], hi[HM_REG];
#pragma unroll
        for (uint32_t k = 0; k < HM_REG; k++) {
            const bool ss = mirror && k < nn && nt[k] == DEME_SPHERE_SPHERE_CONTACT && nb[k] < nS;
            lo[k] = ss ? bRowStart[nb[k]] : 0u;
            hi[k] = ss ? bRowStart[nb[k] + 1] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < HM_REG; k++) {
            if (k >= nn) break;
            uint32_t m = DEME_NULL_MAPPING_PARTNER;
#pragma unroll
            for (uint32_t q = 0; q < HM_REG; q++)
                if (q < np && pb[q] == nb[k] && pt[q] == nt[k]) m = j + q;
            map[ns + k] = m;
            if (!mirror) continue;
            uint32_t link = DEME_NULL_MAPPING_PARTNER;
            for (uint32_t q = lo[k]; q < hi[k] && link == DEME_NULL_MAPPING_PARTNER; q += 4) {
                uint32_t v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = q + u < hi[k] ? bPart[q + u] : 0xFFFFFFFFu;
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (v[u] == s && link == DEME_NULL_MAPPING_PARTNER) link = q + (uint32_t)u;
                if (v[3] > s) break;  // past s in an ascending row
            }
            mirror[ns + k] = link;
        }
        return;
    }
    for (uint32_t i = ns; i < ne; i++) {
        const uint32_t t = ctype[i], b = idB[i];
        const uint32_t gi = t == DEME_SPHERE_SPHERE_CONTACT ? 1u : 0u;
        const uint64_t ki = ((uint64_t)gi << 32) | b;
        uint32_t m = DEME_NULL_MAPPING_PARTNER;
        while (j < je) {
            const uint32_t gj = prevType[j] == DEME_SPHERE_SPHERE_CONTACT ? 1u : 0u;
            const uint64_t kj = ((uint64_t)gj << 32) | prevIdB[j];
            if (kj < ki) {
                j++;
                continue;
            }
            if (kj == ki && prevType[j] == t) m = j;
            break;
        }
        map[i] = m;
        if (!mirror) continue;
        uint32_t link = DEME_NULL_MAPPING_PARTNER;
        if (t == DEME_SPHERE_SPHERE_CONTACT && b < nS) {
            // B's mirror row (ascending, distinct ids) is short: read it four entries per round, all four loads in
            // flight, instead of a binary search's chain of dependent loads
            const uint32_t lo = bRowStart[b], hi = bRowStart[b + 1];
            for (uint32_t q = lo; q < hi && link == DEME_NULL_MAPPING_PARTNER; q += 4) {
                uint32_t v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = q + u < hi ? bPart[q + u] : 0xFFFFFFFFu;
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (v[u] == s && link == DEME_NULL_MAPPING_PARTNER) link = q + (uint32_t)u;
                if (v[3] > s) break;  // past s in an ascending row
            }
        }
        mirror[i] = link;
    }
}

// Mirror rows for a contact list built without them (the staged path): per B sphere, the A ids of its sphere-sphere
// contacts. bCnt holds the per-sphere counts and is consumed (counted down) to hand out slots.
__global__ __launch_bounds__(256) void k_brow_fill(uint32_t nC, uint32_t nS, const uint32_t* __restrict__ idA,
                                                   const uint32_t* __restrict__ idB,
                                                   const uint8_t* __restrict__ ctype,
                                                   const uint32_t* __restrict__ bRowStart, uint32_t* bCnt,
                                                   uint32_t* __restrict__ bPart) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC || ctype[c] != DEME_SPHERE_SPHERE_CONTACT) return;
    const uint32_t b = idB[c];
    if (b >= nS) return;  // guard against a malformed list (never indexes past the row table)
    const uint32_t slot = atomicSub(&bCnt[b], 1u) - 1u;
    if (slot >= bRowStart[b + 1] - bRowStart[b]) return;
    bPart[bRowStart[b] + slot] = idA[c];
}

__global__ __launch_bounds__(256) void k_brow_sort(uint32_t nS, const uint32_t* __restrict__ bRowStart,
                                                   uint32_t* __restrict__ bPart) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const uint32_t st = bRowStart[s], en = bRowStart[s + 1];
    for (uint32_t i = st + 1; i < en; i++) {
        const uint32_t v = bPart[i];
        uint32_t k = i;
        while (k > st && bPart[k - 1] > v) {
            bPart[k] = bPart[k - 1];
            k--;
        }
        bPart[k] = v;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v, uint32_t* p2, uint32_t n2) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
    if (t < n2) p2[t] = v;
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st, uint32_t* p2, uint32_t n2) {
    if (n == 0 && n2 == 0) return hipSuccess;
    uint32_t nb = blocks(n > n2 ? n : n2, 256);
    if (nb > 65536) nb = 65536;
    k_fill_u32<<<nb, 256, 0, st>>>(p, n, v, p2, n2);
    return hipGetLastError();
}

hipError_t launch_plane_pose(const KParams& P, const deme_anal_obj* anal, uint32_t nA, const OwnerRec* own,
                             PlanePose* pose, hipStream_t st) {
    if (nA == 0) return hipSuccess;
    k_plane_pose<<<blocks(nA, 64), 64, 0, st>>>(P, anal, nA, own, pose, nullptr, 0);
    return hipGetLastError();
}

hipError_t launch_pass_start(const KParams& P, const deme_anal_obj* anal, uint32_t nA, const OwnerRec* own,
                             PlanePose* pose, uint32_t* clear, uint32_t nClear, hipStream_t st) {
    if (nA == 0 && nClear == 0) return hipSuccess;
    k_plane_pose<<<nA ? blocks(nA, 64) : 1u, 64, 0, st>>>(P, anal, nA, own, pose, clear, nClear);
    return hipGetLastError();
}

hipError_t launch_sphere_bins(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                              const PlanePose* pose, uint32_t nA, const uint8_t* mask, uint32_t* binCount,
                              uint16_t* numBins, uint32_t* numAnal, CdRec* rec, DevCounters* ctr, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_sphere_bins<<<blocks(nS, 256), 256, 0, st>>>(P, sph, own, nS, pose, nA, mask, binCount, numBins, numAnal, rec,
                                                   ctr);
    return hipGetLastError();
}

hipError_t launch_bin_stats(const uint32_t* binCount, uint64_t nBins, DevCounters* ctr, hipStream_t st) {
    uint32_t nb = blocks(nBins, 256);
    if (nb > 4096) nb = 4096;
    if (nb == 0) nb = 1;
    k_bin_stats<<<nb, 256, 0, st>>>(binCount, nBins, ctr);
    return hipGetLastError();
}

hipError_t launch_bin_scatter(const KParams& P, const CdRec* rec, const SphereRec* sph, uint32_t nS,
                              const uint32_t* binStart, uint32_t* binCount, uint2* binEnt, const DevCounters* ctr,
                              hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_bin_scatter<<<blocks(nS, 256), 256, 0, st>>>(P, rec, sph, nS, binStart, binCount, binEnt, ctr);
    return hipGetLastError();
}

hipError_t launch_bin_pairs(const KParams& P, const uint2* binEnt, uint64_t nEnt, const uint32_t* binStart,
                            const CdRec* rec, const SphereRec* sph, const OwnerRec* own, const uint8_t* mask,
                            uint2* pairs, uint32_t pairCap, DevCounters* ctr, uint32_t* rowCount, uint32_t* bCnt,
                            hipStream_t st) {
    if (nEnt == 0) return hipSuccess;
    k_bin_pairs<<<blocks(nEnt, 256), 256, 0, st>>>(P, binEnt, nEnt, binStart, rec, sph, own, mask, pairs, pairCap, ctr,
                                                   rowCount, bCnt);
    return hipGetLastError();
}

hipError_t launch_row_fill(const uint2* pairs, uint32_t nPairs, const uint32_t* rowStart, const uint32_t* numAnal,
                           uint32_t* rowCount, uint32_t* idB, uint8_t* ctype, hipStream_t st) {
    if (nPairs == 0) return hipSuccess;
    k_row_fill<<<blocks(nPairs, 256), 256, 0, st>>>(pairs, nPairs, rowStart, numAnal, rowCount, idB, ctype);
    return hipGetLastError();
}

hipError_t launch_geo_fill(const KParams& P, const CdRec* rec, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                           const PlanePose* pose, uint32_t nA, const uint8_t* mask, const uint32_t* numAnal,
                           const uint32_t* rowStart, uint32_t* idB, uint8_t* ctype, hipStream_t st) {
    if (nS == 0 || nA == 0) return hipSuccess;
    k_geo_fill<<<blocks(nS, 256), 256, 0, st>>>(P, rec, sph, own, nS, pose, nA, mask, numAnal, rowStart, idB, ctype);
    return hipGetLastError();
}

hipError_t launch_row_finalize(uint32_t nS, const uint32_t* rowStart, const uint32_t* numAnal, uint32_t* idA,
                               uint32_t* idB, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_row_finalize<<<blocks(nS, 256), 256, 0, st>>>(nS, rowStart, numAnal, idA, idB);
    return hipGetLastError();
}

hipError_t launch_history_map(uint32_t nS, const uint32_t* rowStart, const uint32_t* idB, const uint8_t* ctype,
                              const uint32_t* prevRowStart, const uint32_t* prevIdB, const uint8_t* prevType,
                              uint32_t* map, const uint32_t* bRowStart, const uint32_t* bPart, uint32_t* mirror,
                              hipStream_t st, SpecGuard guard) {
    if (nS == 0) return hipSuccess;
    k_history_map<<<blocks(nS, 256), 256, 0, st>>>(nS, rowStart, idB, ctype, prevRowStart, prevIdB, prevType, map,
                                                   bRowStart, bPart, mirror, guard);
    return hipGetLastError();
}

hipError_t launch_brow_fill(uint32_t nC, uint32_t nS, const uint32_t* idA, const uint32_t* idB, const uint8_t* ctype,
                            const uint32_t* bRowStart, uint32_t* bCnt, uint32_t* bPart, hipStream_t st) {
    if (nC == 0) return hipSuccess;
    k_brow_fill<<<blocks(nC, 256), 256, 0, st>>>(nC, nS, idA, idB, ctype, bRowStart, bCnt, bPart);
    return hipGetLastError();
}

hipError_t launch_brow_sort(uint32_t nS, const uint32_t* bRowStart, uint32_t* bPart, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_brow_sort<<<blocks(nS, 256), 256, 0, st>>>(nS, bRowStart, bPart);
    return hipGetLastError();
}

}  // namespace deme_int
