// Fast contact detection for gfx950: the reference's bin rule evaluated over a cell list.
//
// The reference (DEMBinSphereKernels.cu + DEMContactKernels_Blockwise.cu) detects a sphere pair (A < B) in bin b
// iff both expanded spheres' AABBs touch b, the expanded spheres overlap and the contact point falls in b. The pair
// is therefore detected exactly when the overlap holds and the contact point's bin lies in the bin ranges of both
// spheres. This path evaluates that predicate directly:
//   1. every sphere goes to one cell (cell edge >= the largest interaction distance 2 * max(r + beta));
//   2. cells are counting-sorted (one atomic per sphere, returning the slot) and scanned;
//   3. each sphere scans its 27 neighbour cells, keeps partners with a larger id that satisfy the bin rule, and
//      writes them into its own slab row -- no shared counters;
//   4. rows become the canonical contact list (plane contacts first, partners by ascending id).
// The staged per-bin path (cd.hip) materialises the bin table for parity of the intermediate arrays; both paths
// produce the same contact list.
#include "deme_internal.h"

namespace deme_int {
using namespace deme_dev;
namespace {

struct BinRange {
    uint32_t x0, x1, y0, y1, z0, z1;
};

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

__device__ __forceinline__ bool plane_cd(const KParams& P, const PlanePose& pp, double X, double Y, double Z,
                                         double rCD) {
    if (pp.type != DEME_ANAL_OBJ_TYPE_PLANE) return false;
    const double dx = X - pp.px, dy = Y - pp.py, dz = Z - pp.pz;
    const double dist = dx * (double)pp.nx + dy * (double)pp.ny + dz * (double)pp.nz;
    return !(dist > rCD + (double)P.beta);
}

__device__ __forceinline__ bool range_in_grid(const KParams& P, const BinRange& r) {
    return r.x1 < P.nbX && r.y1 < P.nbY && r.z1 < P.nbZ;
}

// Membership of a linear bin id in the linear ids of a range (the reference's bins may alias when a range crosses
// the grid's upper faces; this enumerates exactly the ids the reference would form).
// (Out of line: the rare path. Arguments by value, so the caller's ranges stay in registers.)
__device__ __noinline__ bool range_has_linear(uint32_t nbX, uint32_t nbY, BinRange r, uint32_t lin) {
    const uint32_t nbXY = nbX * nbY;
    for (uint32_t k = 0; k <= r.z1 - r.z0; k++)
        for (uint32_t j = 0; j <= r.y1 - r.y0; j++)
            for (uint32_t i = 0; i <= r.x1 - r.x0; i++)
                if ((r.x0 + i) + (r.y0 + j) * nbX + (r.z0 + k) * nbXY == lin) return true;
    return false;
}

// Cell of a position, clamped to the cell grid (positions outside the world collapse onto edge cells, which only
// adds candidates).
__device__ __forceinline__ void cell_of(const CellGrid& G, double X, double Y, double Z, int& cx, int& cy, int& cz) {
    double fx = floor(X * G.inv), fy = floor(Y * G.inv), fz = floor(Z * G.inv);
    fx = fx < 0 ? 0 : (fx > G.nx - 1 ? G.nx - 1 : fx);
    fy = fy < 0 ? 0 : (fy > G.ny - 1 ? G.ny - 1 : fy);
    fz = fz < 0 ? 0 : (fz > G.nz - 1 ? G.nz - 1 : fz);
    cx = (int)fx;
    cy = (int)fy;
    cz = (int)fz;
}

// Slot of this lane in counter[key] (atomic fetch-add of 1), with one atomic per run of equal keys among adjacent
// active lanes: the spheres of a clump are consecutive and usually share a cell, and same-address atomics of one
// wave serialize. Lanes of a run get consecutive slots in lane order.
__device__ __forceinline__ uint32_t run_slot(uint32_t* counter, uint32_t key) {
    const int lane = threadIdx.x & 63;
    const unsigned long long active = __ballot(1);
    const unsigned long long upTo = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const uint32_t prevKey = __shfl_up(key, 1, 64);
    const bool prevActive = lane > 0 && ((active >> (lane - 1)) & 1ull);
    const bool starts = !prevActive || prevKey != key;
    const unsigned long long starters = __ballot(starts);
    // run: from the highest starter at or below this lane up to (not including) the next starter or inactive lane
    const int start = 63 - __clzll(starters & upTo);
    const unsigned long long fromStart = ~((1ull << start) - 1ull);  // lanes start..63
    const unsigned long long stops = (starters | ~active) & ~upTo;  // first of these above this lane ends the run
    const unsigned long long runMask = fromStart & (stops ? ((1ull << (__ffsll(stops) - 1)) - 1ull) : ~0ull);
    uint32_t base = 0;
    if (lane == start) base = atomicAdd(&counter[key], (uint32_t)__popcll(runMask & active));
    base = __shfl(base, start, 64);
    return base + (uint32_t)__popcll(runMask & active & (upTo >> 1));
}

// Per sphere, shared by the cell assignment and the per-step position pass: world position (a4, a5), bin counts and
// range checks (a6), plane contacts (a10).
__device__ __forceinline__ void sphere_detect(const KParams& P, const SphereRec& sr, const OwnerRec& o, uint32_t s,
                                              const PlanePose* __restrict__ pose, uint32_t nA,
                                              const uint8_t* __restrict__ mask, uint16_t* __restrict__ numBins,
                                              uint32_t* __restrict__ numAnal, unsigned long long* __restrict__ planeBits,
                                              DevCounters* ctr, double& X, double& Y, double& Z) {
    sphere_position(P, o, sr, X, Y, Z);
    double rCD = (double)sr.r;
    rCD += (double)P.beta;
    const BinRange r = bin_range(P, X, Y, Z, rCD);
    const uint32_t nx = r.x1 - r.x0 + 1, ny = r.y1 - r.y0 + 1, nz = r.z1 - r.z0 + 1;
    if (numBins) numBins[s] = (uint16_t)((int)(uint16_t)nx * (int)(uint16_t)ny * (int)(uint16_t)nz);
    if ((uint64_t)nx * ny * nz > 65535u) {
        atomicOr(&ctr->err, ERRBIT_CAPACITY);
    } else {
        // The linear id grows with each of i, j, k, so without 32-bit wrap the range's largest id is its last corner.
        const uint32_t nbXY = P.nbX * P.nbY;
        const uint64_t last = (uint64_t)r.x1 + (uint64_t)r.y1 * P.nbX + (uint64_t)r.z1 * nbXY;
        bool outside = last >= P.nBins;
        if (last > 0xFFFFFFFFull) {
            // ids wrap: enumerate them as the reference forms them
            outside = false;
            for (uint32_t dk = 0; dk < nz; dk++)
                for (uint32_t dj = 0; dj < ny; dj++)
                    for (uint32_t di = 0; di < nx; di++)
                        outside |= ((r.x0 + di) + (r.y0 + dj) * P.nbX + (r.z0 + dk) * nbXY) >= P.nBins;
        }
        if (outside) atomicOr(&ctr->err, ERRBIT_OUTSIDE);
    }
    uint32_t cnt = 0;
    unsigned long long bits = 0;  // bit k: component a0 + k of the sphere's scene
    uint32_t a0, a1;
    scene_planes(P, o.scene, a0, a1);
    for (uint32_t a = a0; a < a1 && a < nA && a - a0 < 64; a++) {
        const PlanePose pp = pose[a];
        if (mask && mask[mask_index(o.fam, pp.fam)] != DEME_DONT_PREVENT_CONTACT) continue;
        if (plane_cd(P, pp, X, Y, Z, rCD)) {
            cnt++;
            bits |= 1ull << (a - a0);
        }
    }
    numAnal[s] = cnt;
    planeBits[s] = bits;
}

// Per sphere: the shared detection work, then the cell assignment with its slot.
__global__ __launch_bounds__(256) void k_cell_assign(KParams P, CellGrid G, const SphereRec* __restrict__ sph,
                                                     const OwnerRec* __restrict__ own, uint32_t nS,
                                                     const PlanePose* __restrict__ pose, uint32_t nA,
                                                     const uint8_t* __restrict__ mask, uint32_t* cellCount,
                                                     uint32_t* __restrict__ cellOf, uint32_t* __restrict__ slotOf,
                                                     uint16_t* __restrict__ numBins, uint32_t* __restrict__ numAnal,
                                                     unsigned long long* __restrict__ planeBits,
                                                     DevCounters* ctr) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const SphereRec sr = load_sphere(sph, s);
    const OwnerRec o = load_owner(own, sr.owner);
    double X, Y, Z;
    sphere_detect(P, sr, o, s, pose, nA, mask, numBins, numAnal, planeBits, ctr, X, Y, Z);
    int cx, cy, cz;
    cell_of(G, X, Y, Z, cx, cy, cz);
    const uint32_t cell = (uint32_t)cx + (uint32_t)cy * G.nx + (uint32_t)cz * G.nx * G.ny;
    cellOf[s] = cell;
    slotOf[s] = run_slot(cellCount, cell);
}

// Per sphere, each step of the candidate-list path: the shared detection work, the detection record in id order and
// (owner, family), and the largest squared displacement since the candidate lists were built (one atomic per wave).
__global__ __launch_bounds__(256) void k_sphere_pos(KParams P, const SphereRec* __restrict__ sph,
                                                    const OwnerRec* __restrict__ own, uint32_t nS,
                                                    const PlanePose* __restrict__ pose, uint32_t nA,
                                                    const uint8_t* __restrict__ mask, CdRec* __restrict__ pos,
                                                    const BuildPos* __restrict__ pos0, double unit,
                                                    uint2* __restrict__ meta,
                                                    uint16_t* __restrict__ numBins, uint32_t* __restrict__ numAnal,
                                                    unsigned long long* __restrict__ planeBits, DevCounters* ctr,
                                                    const double* __restrict__ uIn) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    // displacements since the build are measured against a common translation u (the midpoint of the displacement
    // extremes of a sample -- each workgroup's first sphere -- at the previous pass; any fixed vector keeps the lists
    // exact), so spheres that move together do not age the lists; this pass's sample extremes are folded for the next
    const double ux = uIn[0], uy = uIn[1], uz = uIn[2];
    double d2 = 0.0;
    double m3[3] = {0.0, 0.0, 0.0};
    if (s < nS) {
        const SphereRec sr = load_sphere(sph, s);
        const OwnerRec o = load_owner(own, sr.owner);
        CdRec cr;
        sphere_detect(P, sr, o, s, pose, nA, mask, numBins, numAnal, planeBits, ctr, cr.x, cr.y, cr.z);
        cr.r = sr.r;
        cr.id = s;
        uint4* d4 = reinterpret_cast<uint4*>(pos + s);
        const uint4* s4 = reinterpret_cast<const uint4*>(&cr);
        d4[0] = s4[0];
        d4[1] = s4[1];
        if (meta) meta[s] = make_uint2(sr.owner, o.fam);
        // the build position in fixed point (int -> double and the power-of-two scaling are exact); one that did
        // not fit counts as an infinite displacement, so the lists are found stale
        const BuildPos b = pos0[s];
        if (b.x == INT32_MIN || b.y == INT32_MIN || b.z == INT32_MIN) {
            d2 = INFINITY;
        } else {
            const double mx = cr.x - (double)b.x * unit, my = cr.y - (double)b.y * unit,
                         mz = cr.z - (double)b.z * unit;
            m3[0] = mx;
            m3[1] = my;
            m3[2] = mz;
            const double dx = mx - ux, dy = my - uy, dz = mz - uz;
            d2 = dx * dx + dy * dy + dz * dz;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off, 64));
    __shared__ double waveMax[4];
    if ((threadIdx.x & 63) == 0) waveMax[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) {
        // (lane 0 of every workgroup holds a sphere: the grid covers nS exactly)
        const uint32_t slot = blockIdx.x % DISP_SLOTS;
        const double m = fmax(fmax(waveMax[0], waveMax[1]), fmax(waveMax[2], waveMax[3]));
        if (m > 0.0) atomicMax(&ctr->disp2[slot], (unsigned long long)__double_as_longlong(m));
        for (int a = 0; a < 3; a++) {
            atomicMax(&ctr->dispLo[a][slot], ~ordered_bits(m3[a]));
            atomicMax(&ctr->dispHi[a][slot], ordered_bits(m3[a]));
        }
    }
}

// The sorted CD record of each sphere, its position recomputed from the sphere and owner records exactly as the
// assignment pass computed it (same function, same inputs): writing the records there and reading them back here
// would move 240 bytes per sphere more than the recomputation's gathers.
__global__ __launch_bounds__(256) void k_cell_scatter(KParams P, const SphereRec* __restrict__ sph,
                                                      const OwnerRec* __restrict__ own, uint32_t nS,
                                                      const uint32_t* __restrict__ cellOf,
                                                      const uint32_t* __restrict__ slotOf,
                                                      const uint32_t* __restrict__ cellStart,
                                                      CdRec* __restrict__ sorted, uint2* __restrict__ sortedMeta,
                                                      uint32_t* __restrict__ posOf, uint32_t* __restrict__ aCnt,
                                                      uint32_t* __restrict__ bCnt, float beta,
                                                      float4* __restrict__ sortedF) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const SphereRec sr = load_sphere(sph, s);
    const OwnerRec o = load_owner(own, sr.owner);
    CdRec cr;
    sphere_position(P, o, sr, cr.x, cr.y, cr.z);
    cr.r = sr.r;  // raw radius: the bin rule needs both (double)r + beta and float(r + beta)
    cr.id = s;
    const uint32_t pos = cellStart[cellOf[s]] + slotOf[s];
    posOf[s] = pos;
    sortedMeta[pos] = make_uint2(sr.owner, o.fam);
    // the pair pass accumulates its per-sphere counts with atomics
    aCnt[pos] = 0u;
    bCnt[pos] = 0u;
    CdRec* dst = sorted + pos;
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    const uint4* s4 = reinterpret_cast<const uint4*>(&cr);
    d4[0] = s4[0];
    d4[1] = s4[1];
    // the pair sweep's compact copy, when it uses one: position rounded to float, expanded radius float(r + beta)
    if (sortedF) {
        float rf = cr.r;
        rf += beta;
        sortedF[pos] = make_float4((float)cr.x, (float)cr.y, (float)cr.z, rf);
    }
}

// Expanded radius as the reference's detection kernel forms it: float(r + beta), widened to double.
__device__ __forceinline__ double expanded_radius(const KParams& P, float r) {
    float rf = r;
    rf += P.beta;
    return (double)rf;
}

// First half of the reference's sphere-pair test (checkSpheresOverlap<double>, DEMHelperKernels.cu:195-205): the
// expanded spheres overlap.
__device__ __forceinline__ bool expanded_overlap(double ra, const CdRec& A, double rb, const CdRec& B) {
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    return !(d2 > (ra + rb) * (ra + rb));
}

// Second half, for a pair that overlaps: the contact point (DEMHelperKernels.cu:206-222) computed with A the lower id,
// and its bin inside both spheres' AABB bin ranges.
struct RuleBins {
    uint32_t ix, iy, iz;
    BinRange ga, gb;
};

__device__ __forceinline__ void contact_point(const KParams& P, const CdRec& A, const CdRec& B, double& cx, double& cy,
                                              double& cz) {
    const double ra = expanded_radius(P, A.r), rb = expanded_radius(P, B.r);
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    float nX = (float)dx, nY = (float)dy, nZ = (float)dz;
    const float mag = sqrtf(nX * nX + nY * nY + nZ * nZ);
    nX /= mag;
    nY /= mag;
    nZ /= mag;
    const double half = (ra + rb - sqrt(d2)) / 2.0;
    cx = B.x + (rb - half) * (double)nX;
    cy = B.y + (rb - half) * (double)nY;
    cz = B.z + (rb - half) * (double)nZ;
}

__device__ __forceinline__ double cd_radius(const KParams& P, float r) {
    double v = (double)r;
    v += (double)P.beta;
    return v;
}

// The reference's bin indices, with its divisions by binSize.
__device__ __forceinline__ void rule_bins_exact(const KParams& P, const CdRec& A, const CdRec& B, double cx, double cy,
                                                double cz, RuleBins& R) {
    R.ix = sat_u32(cx / P.binSize);
    R.iy = sat_u32(cy / P.binSize);
    R.iz = sat_u32(cz / P.binSize);
    R.ga = bin_range(P, A.x, A.y, A.z, cd_radius(P, A.r));
    R.gb = bin_range(P, B.x, B.y, B.z, cd_radius(P, B.r));
}

// The same bin indices from multiplications by the reciprocal of the bin size, used when every floor input lies
// farther from an integer than the product and the quotient can differ (a few ulps; the margin is 2^-40 relative):
// then each floor, and so the rule, equals the division's. Returns false when a value is too close to call.
__device__ __forceinline__ bool far_from_int(double t, double scale) {
    const double f = t - floor(t);
    const double eps = 0x1p-40 * (scale + 1.0);
    return f > eps && f < 1.0 - eps;
}

__device__ __forceinline__ bool range_fast(const KParams& P, double inv, double X, double Y, double Z, double rCD,
                                           BinRange& r) {
    const double bx = X * inv, by = Y * inv, bz = Z * inv, span = rCD * inv;
    const double ax = fabs(bx) + span, ay = fabs(by) + span, az = fabs(bz) + span;
    const double x0 = bx - span, x1 = bx + span, y0 = by - span, y1 = by + span, z0 = bz - span, z1 = bz + span;
    if (!(far_from_int(x0, ax) && far_from_int(x1, ax) && far_from_int(y0, ay) && far_from_int(y1, ay) &&
          far_from_int(z0, az) && far_from_int(z1, az)))
        return false;
    r.x0 = sat_u32(x0);
    r.x1 = sat_u32(x1);
    r.y0 = sat_u32(y0);
    r.y1 = sat_u32(y1);
    r.z0 = sat_u32(z0);
    r.z1 = sat_u32(z1);
    return true;
}

__device__ __forceinline__ bool rule_bins_fast(const KParams& P, double inv, const CdRec& A, const CdRec& B, double cx,
                                               double cy, double cz, RuleBins& R) {
    const double qx = cx * inv, qy = cy * inv, qz = cz * inv;
    if (!(far_from_int(qx, fabs(qx)) && far_from_int(qy, fabs(qy)) && far_from_int(qz, fabs(qz)))) return false;
    R.ix = sat_u32(qx);
    R.iy = sat_u32(qy);
    R.iz = sat_u32(qz);
    return range_fast(P, inv, A.x, A.y, A.z, cd_radius(P, A.r), R.ga) &&
           range_fast(P, inv, B.x, B.y, B.z, cd_radius(P, B.r), R.gb);
}

__device__ __forceinline__ bool rule_member(const KParams& P, const RuleBins& R) {
    if (range_in_grid(P, R.ga) && range_in_grid(P, R.gb)) {
        // no aliasing: a linear id is one (i, j, k) -- unless the point itself lies past the grid, where the
        // reference's linear id aliases; fall through to the exact enumeration then
        if (R.ix < P.nbX && R.iy < P.nbY && R.iz < P.nbZ)
            return R.ix >= R.ga.x0 && R.ix <= R.ga.x1 && R.iy >= R.ga.y0 && R.iy <= R.ga.y1 && R.iz >= R.ga.z0 &&
                   R.iz <= R.ga.z1 && R.ix >= R.gb.x0 && R.ix <= R.gb.x1 && R.iy >= R.gb.y0 && R.iy <= R.gb.y1 &&
                   R.iz >= R.gb.z0 && R.iz <= R.gb.z1;
    }
    const uint32_t lin = R.ix + R.iy * P.nbX + R.iz * P.nbX * P.nbY;
    return range_has_linear(P.nbX, P.nbY, R.ga, lin) && range_has_linear(P.nbX, P.nbY, R.gb, lin);
}

// A: the lower id of an overlapping pair. inv: 1 / binSize.
__device__ __forceinline__ bool bin_rule_point(const KParams& P, double inv, const CdRec& A, const CdRec& B) {
    double cx, cy, cz;
    contact_point(P, A, B, cx, cy, cz);
    RuleBins R;
    if (!rule_bins_fast(P, inv, A, B, cx, cy, cz, R)) rule_bins_exact(P, A, B, cx, cy, cz, R);
    return rule_member(P, R);
}

// One lane per sphere in cell order, half-shell: each lane pairs its sphere with the later spheres of its own cell
// and with the spheres of 13 of its 26 neighbour cells (the cell after it in x; the three cells of the next row in y;
// the nine cells of the next layer in z), so every candidate pair is met exactly once. Cells are numbered x fastest,
// so each neighbour row of three cells is one contiguous range of the cell-sorted records.
//
// Two phases keep the block's lanes busy. The sweep runs the reference's double-precision expanded-overlap test and
// the same-owner check and queues every surviving pair in LDS. The block then shares the queue out evenly: the family
// mask and the contact-point bin rule (evaluated with the lower id as A, whichever lane met the pair), and each
// accepted pair is appended to both spheres' columns: the higher id to the lower-id sphere's contact row (rowSlab,
// count aCnt), the lower id to the higher-id sphere's mirror row (mirSlab, count bCnt; a candidate build keeps the
// rows alone). Slabs and counters are indexed
// by sorted position (entry k of sorted sphere i at slab[k * nS + i]), so the appends of a block hit neighbouring
// addresses. Pairs beyond the queue's capacity are parked in the lane's column of `surv` and finished by that lane.
// A column that overflows a slab is reported through maxOcc and the pass is redone with wider columns.
constexpr uint32_t PAIR_QUEUE = 1024;
static_assert(PAIR_QUEUE < (1u << 23), "queue codes hold a kind bit, an 8-bit lane and a 23-bit slot");
// Queue code of an accepted pair in the second phase: kind (0: the finder's contact row, 1: its mirror row) in bit
// 31, the finder's lane in bits 23-30, its slot among the block's entries for that lane below.
__host__ __device__ constexpr uint32_t queue_code(uint32_t kind, uint32_t lane, uint32_t slot) {
    return (kind << 31) | (lane << 23) | slot;
}
__host__ __device__ constexpr uint32_t code_kind(uint32_t c) { return c >> 31; }
__host__ __device__ constexpr uint32_t code_lane(uint32_t c) { return (c >> 23) & 0xFFu; }
__host__ __device__ constexpr uint32_t code_slot(uint32_t c) { return c & 0x7FFFFFu; }
static_assert(code_kind(queue_code(1, 255, PAIR_QUEUE - 1)) == 1 && code_lane(queue_code(1, 255, PAIR_QUEUE - 1)) == 255 &&
                  code_slot(queue_code(1, 255, PAIR_QUEUE - 1)) == PAIR_QUEUE - 1 &&
                  code_lane(queue_code(0, 1, 0)) == 1 && code_slot(queue_code(0, 1, 0)) == 0 &&
                  queue_code(1, 255, PAIR_QUEUE - 1) != 0xFFFFFFFFu,
              "queue codes round-trip and never equal the rejection mark");
constexpr int PAIR_SEGS = 5;

// segment k of the half shell: (dy, dz) = (0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)
__device__ __forceinline__ int seg_dy(int k) { return k == 0 ? 0 : (k == 1 ? 1 : k - 3); }
__device__ __forceinline__ int seg_dz(int k) { return k < 2 ? 0 : 1; }

__device__ __forceinline__ void append_pair(uint32_t* __restrict__ cnt, uint32_t* __restrict__ slab, uint32_t idx,
                                            uint32_t val, uint32_t nS, uint32_t slabCap, DevCounters* ctr) {
    const uint32_t slot = atomicAdd(&cnt[idx], 1u);
    if (slot < slabCap) slab[(uint64_t)slot * nS + idx] = val;
    else atomicMax(&ctr->maxOcc, slot + 1);
}

// Candidate-list predicate: the pair lies within the skin of touching, d <= r'A + r'B + skin (double precision).
__device__ __forceinline__ bool within_skin(const KParams& P, const CdRec& A, const CdRec& B, double skin) {
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double lim = expanded_radius(P, A.r) + expanded_radius(P, B.r) + skin;
    return !(dx * dx + dy * dy + dz * dz > lim * lim);
}

// What the shared second phase still has to check of a queued pair (sorted positions i, j): CHK_NONE, the sweep
// checked owners and the exact overlap; CHK_FLOAT, the pair passed only the float sweep, so the same-owner rule and the
// exact expanded-overlap test come first; CHK_CAND, a candidate-list build: the same-owner rule and the skin test only.
enum : int { CHK_NONE = 0, CHK_FLOAT = 1, CHK_CAND = 2 };

// Second phase, test half: on acceptance the ids of i and j.
template <int CHECK>
__device__ __forceinline__ bool accept_pair(const KParams& P, double inv, const CdRec* __restrict__ sorted,
                                            const uint2* __restrict__ sortedMeta, const uint8_t* __restrict__ mask,
                                            uint32_t i, uint32_t j, double skin, uint32_t& idI, uint32_t& idJ) {
    const bool useMask = CHECK != CHK_CAND && mask;
    if (CHECK != CHK_NONE || useMask) {
        const uint2 mi = sortedMeta[i], mj = sortedMeta[j];
        if (CHECK != CHK_NONE && mi.x == mj.x) return false;
        if (useMask && mask[mask_index(mi.y, mj.y)] != DEME_DONT_PREVENT_CONTACT) return false;
    }
    const CdRec a = load_cd(sorted, i), b = load_cd(sorted, j);
    idI = a.id;
    idJ = b.id;
    if (CHECK == CHK_CAND) return within_skin(P, a, b, skin);
    if (CHECK == CHK_FLOAT && !expanded_overlap(expanded_radius(P, a.r), a, expanded_radius(P, b.r), b)) return false;
    return a.id < b.id ? bin_rule_point(P, inv, a, b) : bin_rule_point(P, inv, b, a);
}

// Second phase for one queued pair (sorted positions i, j) that did not fit the queue; CHECK as above.
template <int CHECK>
__device__ __forceinline__ void finish_pair(const KParams& P, double inv, uint32_t nS, const CdRec* __restrict__ sorted,
                                            const uint2* __restrict__ sortedMeta, const uint8_t* __restrict__ mask,
                                            uint32_t i, uint32_t j, double skin, uint32_t slabCap, uint32_t* rowSlab,
                                            uint32_t* mirSlab, uint32_t* aCnt, uint32_t* bCnt, DevCounters* ctr) {
    uint32_t idI, idJ;
    if (!accept_pair<CHECK>(P, inv, sorted, sortedMeta, mask, i, j, skin, idI, idJ)) return;
    const bool iLow = idI < idJ;
    const uint32_t loIdx = iLow ? i : j, hiIdx = iLow ? j : i;
    const uint32_t loId = iLow ? idI : idJ, hiId = iLow ? idJ : idI;
    append_pair(aCnt, rowSlab, loIdx, hiId, nS, slabCap, ctr);
    if (CHECK != CHK_CAND) append_pair(bCnt, mirSlab, hiIdx, loId, nS, slabCap, ctr);  // (candidates: rows only)
}

// Sweep modes: the exact test on every candidate (PAIR_EXACT); a float pre-reject with the exact test on the
// survivors (PAIR_PRE); the float test alone, with the same-owner rule and the exact test in the shared second phase
// (PAIR_FLOAT).
// PAIR_CAND builds the candidate lists of the candidate-list path: the float sweep with the skin added to its margin,
// then in the shared phase the same-owner rule and the exact test d <= r'A + r'B + skin (no family mask, no bin rule:
// both are applied each step to the pairs found close enough).
enum : int { PAIR_EXACT = 0, PAIR_PRE = 1, PAIR_FLOAT = 2, PAIR_CAND = 3 };

template <int MODE, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void k_cell_pairs(KParams P, CellGrid G, uint32_t nS, uint32_t nBlocks,
                                                           const CdRec* __restrict__ sorted,
                                                           const float4* __restrict__ sortedF, float margin,
                                                           double skin, const uint2* __restrict__ sortedMeta,
                                                           const uint32_t* __restrict__ cellStart,
                                                           const uint8_t* __restrict__ mask, uint32_t slabCap,
                                                           uint32_t* __restrict__ surv, uint32_t* __restrict__ rowSlab,
                                                           uint32_t* __restrict__ mirSlab, uint32_t* aCnt,
                                                           uint32_t* bCnt, DevCounters* ctr) {
    __shared__ uint32_t qI[PAIR_QUEUE], qJ[PAIR_QUEUE];
    __shared__ uint32_t qLen;
    static_assert(BLOCK <= 256, "queue codes hold an 8-bit lane");
    __shared__ uint32_t selfCnt[2][BLOCK], selfBase[2][BLOCK];
    // XCD-aware order: workgroups are dealt to the 8 XCDs round-robin, so XCD k gets a contiguous share of the
    // sphere range and the rows its blocks share stay in its L2
    const uint32_t per = (nBlocks + 7) / 8;
    const uint32_t lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
    // a flagged assignment error (oversized or out-of-world ranges) aborts the pass; both tests are block-uniform
    if (lb >= nBlocks || ctr->err) return;
    if (threadIdx.x == 0) qLen = 0;
    __syncthreads();
    const uint32_t i = lb * BLOCK + threadIdx.x;
    uint32_t nc = 0;  // pairs parked in surv
    if (i < nS) {
        const CdRec me = load_cd(sorted, i);
        constexpr bool PRE = MODE != PAIR_EXACT;
        const float4 mf = PRE ? sortedF[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t myOwner = sortedMeta[i].x;
        int cx, cy, cz;
        cell_of(G, me.x, me.y, me.z, cx, cy, cz);
        const uint32_t nxny = G.nx * G.ny;
        const uint32_t myCell = (uint32_t)cx + (uint32_t)cy * G.nx + (uint32_t)cz * nxny;
        const double ra = expanded_radius(P, me.r);
        const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < (int)G.nx - 1 ? cx + 1 : cx;
        uint32_t* col = surv + i;
        // every segment's bounds are loaded before the first candidate, so their latencies overlap
        uint32_t segB[PAIR_SEGS], segE[PAIR_SEGS];
#pragma unroll
        for (int sg = 0; sg < PAIR_SEGS; sg++) {
            const int y = cy + seg_dy(sg), z = cz + seg_dz(sg);
            segB[sg] = segE[sg] = 0;
            if (y < 0 || y >= (int)G.ny || z >= (int)G.nz) continue;
            if (sg == 0) {
                // own cell after this sphere, then the next cell in x
                segB[sg] = i + 1;
                segE[sg] = cellStart[myCell + (uint32_t)(x1 - cx) + 1];
            } else {
                const uint32_t c0 = (uint32_t)x0 + (uint32_t)y * G.nx + (uint32_t)z * nxny;
                segB[sg] = cellStart[c0];
                segE[sg] = cellStart[c0 + (uint32_t)(x1 - x0) + 1];
            }
        }
        for (int sg = 0; sg < PAIR_SEGS; sg++) {
            const uint32_t b = segB[sg], e = segE[sg];
            // two candidates per iteration: both loads are in flight before either test. With PRE, the float copies
            // reject the pairs that are farther apart than their radii by more than the rounding margin, and only the
            // survivors load the full records for the reference's double-precision test (worth it where most
            // candidates are far: drift margins, wide radius spreads).
            // the candidate's owner is loaded with its record, so an overlapping pair costs no second round trip
            auto push = [&](uint32_t j) {
                const uint32_t q = atomicAdd(&qLen, 1u);
                if (q < PAIR_QUEUE) {
                    qI[q] = i;
                    qJ[q] = j;
                } else {
                    if (nc < slabCap) col[(uint64_t)nc * nS] = j;
                    nc++;
                }
            };
            auto test = [&](const CdRec& ot, uint32_t otOwner, uint32_t j) {
                if (!expanded_overlap(ra, me, expanded_radius(P, ot.r), ot)) return;
                if (otOwner == myOwner) return;  // same clump: never a contact
                const uint32_t q = atomicAdd(&qLen, 1u);
                if (q < PAIR_QUEUE) {
                    qI[q] = i;
                    qJ[q] = j;
                } else {
                    if (nc < slabCap) col[(uint64_t)nc * nS] = j;
                    nc++;
                }
            };
            uint32_t j = b;
            auto pre = [&](const float4& of) {
                const float dx = mf.x - of.x, dy = mf.y - of.y, dz = mf.z - of.z;
                const float lim = mf.w + of.w + margin;
                return !(dx * dx + dy * dy + dz * dz > lim * lim);
            };
            if (MODE == PAIR_FLOAT || MODE == PAIR_CAND) {
                // float copies only: 16 bytes per candidate; the few survivors get their exact tests in the queue
                for (; j + 3 < e; j += 4) {
                    const float4 o0 = sortedF[j], o1 = sortedF[j + 1], o2 = sortedF[j + 2], o3 = sortedF[j + 3];
                    if (pre(o0)) push(j);
                    if (pre(o1)) push(j + 1);
                    if (pre(o2)) push(j + 2);
                    if (pre(o3)) push(j + 3);
                }
                for (; j < e; j++)
                    if (pre(sortedF[j])) push(j);
            } else if (PRE) {
                for (; j + 1 < e; j += 2) {
                    const float4 o0 = sortedF[j], o1 = sortedF[j + 1];
                    if (pre(o0)) test(load_cd(sorted, j), sortedMeta[j].x, j);
                    if (pre(o1)) test(load_cd(sorted, j + 1), sortedMeta[j + 1].x, j + 1);
                }
                if (j < e && pre(sortedF[j])) test(load_cd(sorted, j), sortedMeta[j].x, j);
            } else {
                for (; j + 1 < e; j += 2) {
                    // one address per pair of candidates; the second record at a constant offset
                    const CdRec* pj = sorted + j;
                    const uint2* mj = sortedMeta + j;
                    const CdRec o0 = load_cd(pj, 0), o1 = load_cd(pj, 1);
                    const uint32_t w0 = mj[0].x, w1 = mj[1].x;
                    test(o0, w0, j);
                    test(o1, w1, j + 1);
                }
                if (j < e) test(load_cd(sorted, j), sortedMeta[j].x, j);
            }
        }
    }
    __syncthreads();
    const double inv = 1.0 / P.binSize;
    const uint32_t nq = min(qLen, PAIR_QUEUE);
    const uint32_t blockBase = lb * BLOCK;
    selfCnt[0][threadIdx.x] = selfCnt[1][threadIdx.x] = 0;
    __syncthreads();
    // Every accepted pair makes one entry in each of the two spheres' columns. The partner's entry is appended at
    // once (global atomic); the finder's entry -- always a sphere of this block -- is counted in LDS, so each of the
    // block's spheres takes all its own entries with a single global atomic and they are written together.
    constexpr uint32_t REJECTED = 0xFFFFFFFFu;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const uint32_t si = qI[q], sj = qJ[q];
        uint32_t idI, idJ;
        constexpr int CHECK = MODE == PAIR_CAND ? CHK_CAND : (MODE == PAIR_FLOAT ? CHK_FLOAT : CHK_NONE);
        if (!accept_pair<CHECK>(P, inv, sorted, sortedMeta, mask, si, sj, skin, idI, idJ)) {
            qI[q] = REJECTED;
            continue;
        }
        // i lower id: i's contact row gets j and j's mirror row gets i; else the other way round. A candidate build
        // keeps rows only (its transpose is counted from the rows): no mirror entries.
        const uint32_t kind = idI < idJ ? 0u : 1u;
        if (kind == 0) {
            if (MODE != PAIR_CAND) append_pair(bCnt, mirSlab, sj, idI, nS, slabCap, ctr);
        } else {
            append_pair(aCnt, rowSlab, sj, idI, nS, slabCap, ctr);
            if (MODE == PAIR_CAND) {
                qI[q] = REJECTED;  // no mirror entry of its own
                continue;
            }
        }
        const uint32_t local = si - blockBase;
        const uint32_t ls = atomicAdd(&selfCnt[kind][local], 1u);
        qI[q] = queue_code(kind, local, ls);  // ls < PAIR_QUEUE
        qJ[q] = idJ;
    }
    __syncthreads();
    {
        const uint32_t t = threadIdx.x;
        if (selfCnt[0][t]) selfBase[0][t] = atomicAdd(&aCnt[blockBase + t], selfCnt[0][t]);
        if (selfCnt[1][t]) selfBase[1][t] = atomicAdd(&bCnt[blockBase + t], selfCnt[1][t]);
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const uint32_t code = qI[q];
        if (code == REJECTED) continue;
        const uint32_t kind = code_kind(code), local = code_lane(code);
        const uint32_t slot = selfBase[kind][local] + code_slot(code);
        if (slot < slabCap) (kind == 0 ? rowSlab : mirSlab)[(uint64_t)slot * nS + blockBase + local] = qJ[q];
        else atomicMax(&ctr->maxOcc, slot + 1);
    }
    if (nc > slabCap) {
        atomicMax(&ctr->maxOcc, nc);
        return;
    }
    for (uint32_t k = 0; k < nc; k++)
        finish_pair<MODE == PAIR_CAND ? CHK_CAND : (MODE == PAIR_FLOAT ? CHK_FLOAT : CHK_NONE)>(
            P, inv, nS, sorted, sortedMeta, mask, i, surv[(uint64_t)k * nS + i], skin, slabCap, rowSlab, mirSlab, aCnt,
            bCnt, ctr);
}

// Canonical rows: plane contacts (ascending plane index) then partners (ascending id); idA for the row. Mirror rows:
// the sphere's lower-id partners, ascending. Both come from the sphere's slab columns (sorted position).
// A column of up to SORT_REG entries is loaded at once (all loads in flight), sorted in registers by an unrolled
// insertion network and stored; longer columns (rare) are sorted in place.
constexpr uint32_t SORT_REG = 8;
__device__ __forceinline__ void sorted_copy(const uint32_t* __restrict__ src, uint32_t n, uint32_t nS,
                                            uint32_t* __restrict__ dst) {
    if (n <= SORT_REG) {
        uint32_t v[SORT_REG];
#pragma unroll
        for (uint32_t k = 0; k < SORT_REG; k++) v[k] = k < n ? src[(uint64_t)k * nS] : 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t i = 1; i < SORT_REG; i++)
#pragma unroll
            for (uint32_t j = i; j > 0; j--) {
                const uint32_t lo = min(v[j - 1], v[j]), hi = max(v[j - 1], v[j]);
                v[j - 1] = lo;
                v[j] = hi;
            }
#pragma unroll
        for (uint32_t k = 0; k < SORT_REG; k++)
            if (k < n) dst[k] = v[k];
        return;
    }
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t v = src[(uint64_t)k * nS];
        uint32_t q = k;
        while (q > 0 && dst[q - 1] > v) {
            dst[q] = dst[q - 1];
            q--;
        }
        dst[q] = v;
    }
}

__global__ __launch_bounds__(256) void k_cell_rows(KParams P, const SphereRec* __restrict__ sph,
                                                   const OwnerRec* __restrict__ own, uint32_t nS,
                                                   const uint32_t* __restrict__ rowSlab,
                                                   const uint32_t* __restrict__ mirSlab,
                                                   const uint32_t* __restrict__ posOf,
                                                   const uint32_t* __restrict__ rowStart,
                                                   const unsigned long long* __restrict__ planeBits,
                                                   const uint32_t* __restrict__ bRowStart, uint32_t* __restrict__ idA,
                                                   uint32_t* __restrict__ idB, uint8_t* __restrict__ ctype,
                                                   uint32_t* __restrict__ bPart, SpecGuard guard) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || spec_skip(guard)) return;
    const uint32_t st = rowStart[s], en = rowStart[s + 1];
    uint32_t pos = st;
    unsigned long long bits = planeBits[s];
    uint32_t a0 = 0, a1 = 0;
    if (bits) scene_planes(P, own[sph[s].owner].scene, a0, a1);
    while (bits && pos < en) {
        const int a = __ffsll(bits) - 1;
        bits &= bits - 1;
        idB[pos] = a0 + (uint32_t)a;
        ctype[pos] = DEME_SPHERE_PLANE_CONTACT;
        pos++;
    }
    const uint32_t sp = pos;
    const uint32_t p = posOf ? posOf[s] : s;  // slab columns by sorted position (or by id when posOf is null)
    sorted_copy(rowSlab + p, en - sp, nS, idB + sp);
    const uint32_t bs = bRowStart[s];
    sorted_copy(mirSlab + p, bRowStart[s + 1] - bs, nS, bPart + bs);
    for (uint32_t q = sp; q < en; q++) ctype[q] = DEME_SPHERE_SPHERE_CONTACT;
    for (uint32_t q = st; q < en; q++) idA[q] = s;
}

// A build position coordinate in fixed point (scale = 1 / unit): rounded to the nearest unit, INT32_MIN when it does
// not fit in (-2^31, 2^31) units (a NaN fails both comparisons).
__device__ __forceinline__ int32_t build_fix(double x, double scale) {
    const double v = rint(x * scale);
    return (v > -2147483648.0 && v < 2147483648.0) ? (int32_t)v : INT32_MIN;
}

// Candidate lists from a candidate build: sphere s's row of higher-id partners within the skin (ascending), copied
// from its slab column at its sorted position; the sphere's position at the build is kept (in fixed point) for the
// displacement test.
__global__ __launch_bounds__(256) void k_cand_rows(uint32_t nS, const uint32_t* __restrict__ rowSlab,
                                                   const uint32_t* __restrict__ candStart, uint32_t* __restrict__ candA,
                                                   uint32_t* __restrict__ candB, const CdRec* __restrict__ sorted,
                                                   BuildPos* __restrict__ pos0, double scale) {
    // one lane per sorted position: the lanes of a wave read adjacent slab columns (coalesced) and write their
    // spheres' rows wherever the ids put them
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nS) return;
    const CdRec r = load_cd(sorted, p);
    const uint32_t s = r.id;
    const uint32_t st = candStart[s], en = candStart[s + 1];
    sorted_copy(rowSlab + p, en - st, nS, candB + st);
    for (uint32_t c = st; c < en; c++) candA[c] = s;
    pos0[s] = BuildPos{build_fix(r.x, scale), build_fix(r.y, scale), build_fix(r.z, scale)};
}

// The candidates of one build are indexed once (candidate c: sphere candA[c], partner candB[c] of higher id; the
// rows candStart[A]..candStart[A + 1] ascending in partner) and transposed once (tC[tStart[B]..tStart[B + 1]]: the
// candidates whose partner is B, ascending A). Every pass between builds marks which candidates touch (acc[c]) and
// derives the canonical rows, the mirror rows and the persistent-contact map from those indices alone: a contact of
// this pass that was a contact of the previous pass is the same candidate, so the map is the previous pass's contact
// index of c, and no list is sorted, merged or searched.

// Per candidate: the reference's predicate (the exact expanded overlap, the family mask, the contact-point bin rule)
// on the pair, which holds every partner that can touch while no sphere has moved more than half the skin since the
// build. If a sphere moved further the pass is marked stale and does nothing; the host rebuilds the lists and
// repeats it.
__global__ __launch_bounds__(256) void k_cand_flags(KParams P, uint32_t nK, const CdRec* __restrict__ pos,
                                                    const uint2* __restrict__ meta,
                                                    const uint32_t* __restrict__ candA,
                                                    const uint32_t* __restrict__ candB,
                                                    const uint8_t* __restrict__ mask, double dispLimit2,
                                                    uint8_t* __restrict__ acc, DevCounters* ctr,
                                                    double* __restrict__ uOut) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctr->err) return;
    if (c == 0) {
        // the next pass's common translation: the midpoint of this pass's displacement extremes, per axis
        for (int a = 0; a < 3; a++) {
            unsigned long long l = 0, h = 0;
            for (uint32_t k = 0; k < DISP_SLOTS; k++) {
                l = max(l, ctr->dispLo[a][k]);
                h = max(h, ctr->dispHi[a][k]);
            }
            const double u = h ? 0.5 * (from_ordered_bits(~l) + from_ordered_bits(h)) : 0.0;  // (no spheres: none)
            uOut[a] = isfinite(u) ? u : 0.0;  // (a non-finite position must not hide a displacement)
        }
    }
    // (the grid has at least one lane, so a build without candidates is still found stale). The largest displacement
    // is folded across the wave, one slot per lane: a loop over the slots would run on the scalar unit, which the
    // CU's waves share.
    static_assert(DISP_SLOTS == 64, "one displacement slot per lane of a wave");
    unsigned long long dm = ctr->disp2[threadIdx.x & 63u];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(dm, off, 64);
        dm = o > dm ? o : dm;
    }
    double dmax;
    __builtin_memcpy(&dmax, &dm, sizeof(dmax));
    if (dmax > dispLimit2) {
        if (c == 0) atomicOr(&ctr->err, ERRBIT_STALE);
        return;
    }
    if (c >= nK) return;
    const uint32_t ia = candA[c], ib = candB[c];
    const CdRec a = load_cd(pos, ia), b = load_cd(pos, ib);
    uint8_t f = 0;
    if (expanded_overlap(expanded_radius(P, a.r), a, expanded_radius(P, b.r), b) &&
        (!mask || mask[mask_index(meta[ia].y, meta[ib].y)] == DEME_DONT_PREVENT_CONTACT))
        f = bin_rule_point(P, 1.0 / P.binSize, a, b) ? 1 : 0;  // a has the lower id
    acc[c] = f;
}

// Per sphere: its row's touching candidates (aCnt) and those with it as partner (bCnt, its mirror row's length), with
// each touching transposed slot's rank in the mirror row (tRank).
__global__ __launch_bounds__(256) void k_cand_count(uint32_t nS, const uint32_t* __restrict__ candStart,
                                                    const uint32_t* __restrict__ tStart,
                                                    const uint32_t* __restrict__ tC,
                                                    const uint8_t* __restrict__ acc, uint32_t* __restrict__ aCnt,
                                                    uint32_t* __restrict__ bCnt, uint32_t* __restrict__ tRank,
                                                    const DevCounters* ctr) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || ctr->err) return;
    uint32_t na = 0, nb = 0;
    const uint32_t c1 = candStart[s + 1], t1 = tStart[s + 1];
    for (uint32_t c = candStart[s]; c < c1; c++) na += acc[c];
    uint32_t t = tStart[s];
    for (; t + 3 < t1; t += 4) {
        const uint32_t q0 = tC[t], q1 = tC[t + 1], q2 = tC[t + 2], q3 = tC[t + 3];
        const uint32_t f0 = acc[q0], f1 = acc[q1], f2 = acc[q2], f3 = acc[q3];
        if (f0) tRank[t] = nb++;
        if (f1) tRank[t + 1] = nb++;
        if (f2) tRank[t + 2] = nb++;
        if (f3) tRank[t + 3] = nb++;
    }
    for (; t < t1; t++)
        if (acc[tC[t]]) tRank[t] = nb++;
    aCnt[s] = na;
    bCnt[s] = nb;
}

// Per touching candidate: its contact in A's canonical row (after A's plane contacts, in candidate order), idA, its
// mirror link (B's mirror row start plus the rank the row scan gave c's transposed slot), cidx[c] and, when prevAcc
// is given (the previous pass used the same candidates), its persistent-contact map entry:
// the previous pass's contact index of c, or none if c did not touch then. A wave covers 64 consecutive candidates:
// c's rank among the touching candidates of its row is a population count of the wave's ballot, plus, for a row
// that began before the wave, the touching candidates of the row ahead of it (counted once by lane 0).
__global__ __launch_bounds__(256) void k_cand_fill(uint32_t nK, const uint32_t* __restrict__ candA,
                                                   const uint32_t* __restrict__ candB,
                                                   const uint32_t* __restrict__ candStart,
                                                   const uint8_t* __restrict__ acc,
                                                   const uint32_t* __restrict__ rowStart,
                                                   const uint32_t* __restrict__ numAnal,
                                                   const uint32_t* __restrict__ tPos,
                                                   const uint32_t* __restrict__ tRank,
                                                   const uint32_t* __restrict__ bRowStart,
                                                   const uint8_t* __restrict__ prevAcc,
                                                   const uint32_t* __restrict__ prevCidx, uint32_t* __restrict__ cidx,
                                                   uint32_t* __restrict__ idA, uint32_t* __restrict__ idB,
                                                   uint8_t* __restrict__ ctype, uint32_t* __restrict__ map,
                                                   uint32_t* __restrict__ mirror, SpecGuard guard) {
    if (spec_skip(guard)) return;  // uniform: the whole grid leaves
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w0 = c - lane;  // the wave's first candidate
    if (w0 >= nK) return;          // whole wave past the end
    const bool in = c < nK;
    const uint32_t a = in ? candA[c] : 0u;
    const bool t = in && acc[c];
    const unsigned long long ball = __ballot(t);
    const uint32_t r0 = in ? candStart[a] : 0u;
    // everything a touching lane reads, requested before the lanes that do not touch leave and before the row-ahead
    // count, so the loads overlap
    const uint32_t rs = in ? rowStart[a] : 0u, re = in ? rowStart[a + 1] : 0u, na = in ? numAnal[a] : 0u;
    const uint32_t b = in ? candB[c] : 0u;
    const bool pa = in && prevAcc && prevAcc[c];
    const uint32_t pc = in && prevAcc ? prevCidx[c] : DEME_NULL_MAPPING_PARTNER;
    // (a lane that does not touch reads a rank this pass did not write, in bounds, and leaves before using it)
    const uint32_t mir = in ? bRowStart[b] + tRank[tPos[c]] : 0u;
    // touching candidates of the wave's first row ahead of the wave (lane 0's row is the only one that can begin
    // before it)
    uint32_t ahead = 0;
    if (lane == 0 && in) {
        const uint32_t rr = candStart[a];
        for (uint32_t q = rr; q < c; q++) ahead += acc[q];
    }
    ahead = __shfl(ahead, 0, 64);
    if (!t) return;
    const uint32_t lo = r0 > w0 ? r0 - w0 : 0u;  // first lane of c's row in this wave
    const unsigned long long below = (lane ? (~0ull >> (64u - lane)) : 0ull) & ~(lo ? (~0ull >> (64u - lo)) : 0ull);
    const uint32_t k = (uint32_t)__popcll(ball & below) + (r0 < w0 ? ahead : 0u);
    const uint32_t i = rs + na + k;
    if (i >= re) {  // counts that disagree with the scan (not reached)
        cidx[c] = DEME_NULL_MAPPING_PARTNER;
        return;
    }
    idA[i] = a;
    idB[i] = b;
    ctype[i] = DEME_SPHERE_SPHERE_CONTACT;
    mirror[i] = mir;
    cidx[c] = i;
    if (prevAcc) map[i] = pa ? pc : DEME_NULL_MAPPING_PARTNER;
}

// Per sphere: its plane contacts at the head of its canonical row (ascending plane index; the map entry, when
// mapPlanes, is the same plane's entry in its previous row; no mirror slot). Spheres touching no plane do nothing.
__global__ __launch_bounds__(256) void k_cand_planes(KParams P, const SphereRec* __restrict__ sph,
                                                     const OwnerRec* __restrict__ own, uint32_t nS,
                                                     const uint32_t* __restrict__ rowStart,
                                                     const unsigned long long* __restrict__ planeBits,
                                                     const uint32_t* __restrict__ prevRowStart,
                                                     const uint32_t* __restrict__ prevIdB,
                                                     const uint8_t* __restrict__ prevType, bool mapPlanes,
                                                     uint32_t* __restrict__ idA, uint32_t* __restrict__ idB,
                                                     uint8_t* __restrict__ ctype, uint32_t* __restrict__ map,
                                                     uint32_t* __restrict__ mirror, SpecGuard guard) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS || spec_skip(guard)) return;
    unsigned long long bits = planeBits[s];
    if (!bits) return;
    uint32_t pos = rowStart[s];
    const uint32_t en = rowStart[s + 1];
    uint32_t a0 = 0, a1 = 0;
    scene_planes(P, own[sph[s].owner].scene, a0, a1);
    // the previous row starts with its plane contacts, ascending: a merge over them
    uint32_t j = 0, je = 0;
    if (mapPlanes && prevRowStart) {
        j = prevRowStart[s];
        je = prevRowStart[s + 1];
    }
    while (bits && pos < en) {
        const uint32_t pid = a0 + (uint32_t)(__ffsll(bits) - 1);
        bits &= bits - 1;
        idA[pos] = s;
        idB[pos] = pid;
        ctype[pos] = DEME_SPHERE_PLANE_CONTACT;
        mirror[pos] = DEME_NULL_MAPPING_PARTNER;
        if (mapPlanes) {
            while (j < je && prevType[j] == DEME_SPHERE_PLANE_CONTACT && prevIdB[j] < pid) j++;
            map[pos] = (j < je && prevType[j] == DEME_SPHERE_PLANE_CONTACT && prevIdB[j] == pid)
                           ? j
                           : DEME_NULL_MAPPING_PARTNER;
        }
        pos++;
    }
}

// The transpose of a candidate build, counted from its rows: per candidate, its place among the candidates with the
// same partner (an atomic count per partner, in no particular order) ...
__global__ __launch_bounds__(256) void k_cand_tcount(uint32_t nK, const uint32_t* __restrict__ candB,
                                                     uint32_t* __restrict__ tCnt, uint32_t* __restrict__ slot) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nK) return;
    slot[c] = atomicAdd(&tCnt[candB[c]], 1u);
}

// ... each candidate stored at that place of its partner's transposed row (tStart from a scan of the counts) ...
__global__ __launch_bounds__(256) void k_cand_tscatter(uint32_t nK, const uint32_t* __restrict__ candB,
                                                       const uint32_t* __restrict__ slot,
                                                       const uint32_t* __restrict__ tStart,
                                                       uint32_t* __restrict__ tC) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nK) return;
    tC[tStart[candB[c]] + slot[c]] = c;
}

// ... and each transposed row sorted ascending (candidate order is row order, so ascending A), in place: rows of up to
// SORT_REG entries in registers, longer ones by insertion; then each candidate's slot recorded, tPos[tC[t]] = t.
__global__ __launch_bounds__(256) void k_cand_tsort(uint32_t nS, const uint32_t* __restrict__ tStart,
                                                    uint32_t* __restrict__ tC, uint32_t* __restrict__ tPos) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const uint32_t t0 = tStart[s], n = tStart[s + 1] - t0;
    uint32_t* row = tC + t0;
    if (n < 2) {
        if (n == 1) tPos[row[0]] = t0;
        return;
    }
    if (n <= SORT_REG) {
        uint32_t v[SORT_REG];
#pragma unroll
        for (uint32_t k = 0; k < SORT_REG; k++) v[k] = k < n ? row[k] : 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t i = 1; i < SORT_REG; i++)
#pragma unroll
            for (uint32_t j = i; j > 0; j--) {
                const uint32_t lo = min(v[j - 1], v[j]), hi = max(v[j - 1], v[j]);
                v[j - 1] = lo;
                v[j] = hi;
            }
#pragma unroll
        for (uint32_t k = 0; k < SORT_REG; k++)
            if (k < n) {
                row[k] = v[k];
                tPos[v[k]] = t0 + k;
            }
        return;
    }
    for (uint32_t k = 1; k < n; k++) {
        const uint32_t v = row[k];
        uint32_t q = k;
        while (q > 0 && row[q - 1] > v) {
            row[q] = row[q - 1];
            q--;
        }
        row[q] = v;
    }
    for (uint32_t k = 0; k < n; k++) tPos[row[k]] = t0 + k;
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_cell_assign(const KParams& P, const CellGrid& G, const SphereRec* sph, const OwnerRec* own,
                              uint32_t nS, const PlanePose* pose, uint32_t nA, const uint8_t* mask,
                              uint32_t* cellCount, uint32_t* cellOf, uint32_t* slotOf, uint16_t* numBins,
                              uint32_t* numAnal, unsigned long long* planeBits, DevCounters* ctr, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_assign<<<blocks(nS, 256), 256, 0, st>>>(P, G, sph, own, nS, pose, nA, mask, cellCount, cellOf, slotOf,
                                                   numBins, numAnal, planeBits, ctr);
    return hipGetLastError();
}

hipError_t launch_cell_scatter(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                               const uint32_t* cellOf, const uint32_t* slotOf, const uint32_t* cellStart,
                               CdRec* sorted, uint2* sortedMeta, uint32_t* posOf, uint32_t* aCnt, uint32_t* bCnt,
                               float beta, float4* sortedF, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_scatter<<<blocks(nS, 256), 256, 0, st>>>(P, sph, own, nS, cellOf, slotOf, cellStart, sorted, sortedMeta,
                                                    posOf, aCnt, bCnt, beta, sortedF);
    return hipGetLastError();
}

template <uint32_t BLOCK>
static void launch_pairs_block(int mode, uint32_t grid, hipStream_t st, const KParams& P, const CellGrid& G,
                               uint32_t nS, uint32_t nb, const CdRec* sorted, const float4* sortedF, float margin,
                               double skin, const uint2* sortedMeta, const uint32_t* cellStart, const uint8_t* mask,
                               uint32_t slabCap, uint32_t* surv, uint32_t* rowSlab, uint32_t* mirSlab, uint32_t* aCnt,
                               uint32_t* bCnt, DevCounters* ctr) {
    if (mode == PAIR_CAND)
        k_cell_pairs<PAIR_CAND, BLOCK><<<grid, BLOCK, 0, st>>>(P, G, nS, nb, sorted, sortedF, margin, skin, sortedMeta,
                                                               cellStart, nullptr, slabCap, surv, rowSlab, mirSlab,
                                                               aCnt, bCnt, ctr);
    else if (mode == PAIR_FLOAT)
        k_cell_pairs<PAIR_FLOAT, BLOCK><<<grid, BLOCK, 0, st>>>(P, G, nS, nb, sorted, sortedF, margin, 0.0, sortedMeta,
                                                                cellStart, mask, slabCap, surv, rowSlab, mirSlab,
                                                                aCnt, bCnt, ctr);
    else if (mode == PAIR_PRE)
        k_cell_pairs<PAIR_PRE, BLOCK><<<grid, BLOCK, 0, st>>>(P, G, nS, nb, sorted, sortedF, margin, 0.0, sortedMeta,
                                                              cellStart, mask, slabCap, surv, rowSlab, mirSlab, aCnt,
                                                              bCnt, ctr);
    else
        k_cell_pairs<PAIR_EXACT, BLOCK><<<grid, BLOCK, 0, st>>>(P, G, nS, nb, sorted, nullptr, 0.f, 0.0, sortedMeta,
                                                                cellStart, mask, slabCap, surv, rowSlab, mirSlab,
                                                                aCnt, bCnt, ctr);
}

hipError_t launch_cell_pairs(const KParams& P, const CellGrid& G, uint32_t nS, const CdRec* sorted,
                             const float4* sortedF, float margin, int mode, uint32_t block,
                             const uint2* sortedMeta, const uint32_t* cellStart, const uint8_t* mask, uint32_t slabCap,
                             uint32_t* surv, uint32_t* rowSlab, uint32_t* mirSlab, uint32_t* aCnt, uint32_t* bCnt,
                             DevCounters* ctr, hipStream_t st, double skin) {
    if (nS == 0) return hipSuccess;
    if (mode < PAIR_EXACT || mode > PAIR_CAND) return hipErrorInvalidValue;
    if (mode != PAIR_EXACT && !sortedF) return hipErrorInvalidValue;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    // the candidate build's float test covers the skin as well
    const float m = mode == PAIR_CAND ? margin + (float)skin : margin;
    const uint32_t nb = blocks(nS, block), grid = 8 * ((nb + 7) / 8);
    if (block == 64)
        launch_pairs_block<64>(mode, grid, st, P, G, nS, nb, sorted, sortedF, m, skin, sortedMeta, cellStart, mask,
                               slabCap, surv, rowSlab, mirSlab, aCnt, bCnt, ctr);
    else if (block == 256)
        launch_pairs_block<256>(mode, grid, st, P, G, nS, nb, sorted, sortedF, m, skin, sortedMeta, cellStart, mask,
                                slabCap, surv, rowSlab, mirSlab, aCnt, bCnt, ctr);
    else
        launch_pairs_block<128>(mode, grid, st, P, G, nS, nb, sorted, sortedF, m, skin, sortedMeta, cellStart, mask,
                                slabCap, surv, rowSlab, mirSlab, aCnt, bCnt, ctr);
    return hipGetLastError();
}

hipError_t launch_sphere_pos(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                             const PlanePose* pose, uint32_t nA, const uint8_t* mask, CdRec* pos,
                             const BuildPos* pos0, double unit, uint2* meta, uint16_t* numBins, uint32_t* numAnal,
                             unsigned long long* planeBits, DevCounters* ctr, const double* uIn, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_sphere_pos<<<blocks(nS, 256), 256, 0, st>>>(P, sph, own, nS, pose, nA, mask, pos, pos0, unit, meta, numBins,
                                                  numAnal, planeBits, ctr, uIn);
    return hipGetLastError();
}

hipError_t launch_cand_rows(uint32_t nS, const uint32_t* rowSlab, const uint32_t* candStart, uint32_t* candA,
                            uint32_t* candB, const CdRec* sorted, BuildPos* pos0, double scale, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cand_rows<<<blocks(nS, 256), 256, 0, st>>>(nS, rowSlab, candStart, candA, candB, sorted, pos0, scale);
    return hipGetLastError();
}

hipError_t launch_cand_flags(const KParams& P, uint32_t nK, const CdRec* pos, const uint2* meta, const uint32_t* candA,
                             const uint32_t* candB, const uint8_t* mask, double dispLimit2, uint8_t* acc,
                             DevCounters* ctr, double* uOut, hipStream_t st) {
    k_cand_flags<<<blocks(nK > 0 ? nK : 1u, 256), 256, 0, st>>>(P, nK, pos, meta, candA, candB, mask, dispLimit2, acc,
                                                                ctr, uOut);
    return hipGetLastError();
}

hipError_t launch_cand_count(uint32_t nS, const uint32_t* candStart, const uint32_t* tStart, const uint32_t* tC,
                             const uint8_t* acc, uint32_t* aCnt, uint32_t* bCnt, uint32_t* tRank,
                             const DevCounters* ctr, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cand_count<<<blocks(nS, 256), 256, 0, st>>>(nS, candStart, tStart, tC, acc, aCnt, bCnt, tRank, ctr);
    return hipGetLastError();
}

hipError_t launch_cand_fill(uint32_t nK, const uint32_t* candA, const uint32_t* candB, const uint32_t* candStart,
                            const uint8_t* acc, const uint32_t* rowStart, const uint32_t* numAnal,
                            const uint32_t* tPos, const uint32_t* tRank, const uint32_t* bRowStart,
                            const uint8_t* prevAcc, const uint32_t* prevCidx, uint32_t* cidx, uint32_t* idA,
                            uint32_t* idB, uint8_t* ctype, uint32_t* map, uint32_t* mirror, hipStream_t st,
                            SpecGuard guard) {
    if (nK == 0) return hipSuccess;
    k_cand_fill<<<blocks(nK, 256), 256, 0, st>>>(nK, candA, candB, candStart, acc, rowStart, numAnal, tPos, tRank,
                                                 bRowStart, prevAcc, prevCidx, cidx, idA, idB, ctype, map, mirror,
                                                 guard);
    return hipGetLastError();
}

hipError_t launch_cand_planes(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                              const uint32_t* rowStart, const unsigned long long* planeBits,
                              const uint32_t* prevRowStart, const uint32_t* prevIdB, const uint8_t* prevType,
                              bool mapPlanes, uint32_t* idA, uint32_t* idB, uint8_t* ctype, uint32_t* map,
                              uint32_t* mirror, hipStream_t st, SpecGuard guard) {
    if (nS == 0) return hipSuccess;
    k_cand_planes<<<blocks(nS, 256), 256, 0, st>>>(P, sph, own, nS, rowStart, planeBits, prevRowStart, prevIdB,
                                                   prevType, mapPlanes, idA, idB, ctype, map, mirror, guard);
    return hipGetLastError();
}

hipError_t launch_cand_tcount(uint32_t nK, const uint32_t* candB, uint32_t* tCnt, uint32_t* slot, hipStream_t st) {
    if (nK == 0) return hipSuccess;
    k_cand_tcount<<<blocks(nK, 256), 256, 0, st>>>(nK, candB, tCnt, slot);
    return hipGetLastError();
}

hipError_t launch_cand_tscatter(uint32_t nK, const uint32_t* candB, const uint32_t* slot, const uint32_t* tStart,
                                uint32_t* tC, hipStream_t st) {
    if (nK == 0) return hipSuccess;
    k_cand_tscatter<<<blocks(nK, 256), 256, 0, st>>>(nK, candB, slot, tStart, tC);
    return hipGetLastError();
}

hipError_t launch_cand_tsort(uint32_t nS, const uint32_t* tStart, uint32_t* tC, uint32_t* tPos, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cand_tsort<<<blocks(nS, 256), 256, 0, st>>>(nS, tStart, tC, tPos);
    return hipGetLastError();
}

hipError_t launch_cell_rows(const KParams& P, const SphereRec* sph, const OwnerRec* own, uint32_t nS,
                            const uint32_t* rowSlab, const uint32_t* mirSlab, const uint32_t* posOf,
                            const uint32_t* rowStart, const unsigned long long* planeBits, const uint32_t* bRowStart,
                            uint32_t* idA, uint32_t* idB, uint8_t* ctype, uint32_t* bPart, hipStream_t st,
                            SpecGuard guard) {
    if (nS == 0) return hipSuccess;
    k_cell_rows<<<blocks(nS, 256), 256, 0, st>>>(P, sph, own, nS, rowSlab, mirSlab, posOf, rowStart, planeBits,
                                                 bRowStart, idA, idB, ctype, bPart, guard);
    return hipGetLastError();
}

}  // namespace deme_int
