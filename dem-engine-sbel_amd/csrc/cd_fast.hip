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
__device__ __noinline__ bool range_has_linear(const KParams& P, const BinRange& r, uint32_t lin) {
    const uint32_t nbXY = P.nbX * P.nbY;
    for (uint32_t k = 0; k <= r.z1 - r.z0; k++)
        for (uint32_t j = 0; j <= r.y1 - r.y0; j++)
            for (uint32_t i = 0; i <= r.x1 - r.x0; i++)
                if ((r.x0 + i) + (r.y0 + j) * P.nbX + (r.z0 + k) * nbXY == lin) return true;
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

// Per sphere: world position, plane contacts (a10), bin counts (a6), cell assignment with its slot.
__global__ __launch_bounds__(256) void k_cell_assign(KParams P, CellGrid G, const SphereRec* __restrict__ sph,
                                                     const OwnerRec* __restrict__ own, uint32_t nS,
                                                     const PlanePose* __restrict__ pose, uint32_t nA,
                                                     const uint8_t* __restrict__ mask, uint32_t* cellCount,
                                                     uint32_t* __restrict__ cellOf, uint32_t* __restrict__ slotOf,
                                                     uint16_t* __restrict__ numBins, uint32_t* __restrict__ numAnal,
                                                     unsigned long long* __restrict__ planeBits,
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
    numBins[s] = (uint16_t)((int)(uint16_t)nx * (int)(uint16_t)ny * (int)(uint16_t)nz);
    if ((uint64_t)nx * ny * nz > 65535u) {
        atomicOr(&ctr->err, ERRBIT_CAPACITY);
    } else {
        const uint32_t nbXY = P.nbX * P.nbY;
        bool outside = false;
        for (uint32_t dk = 0; dk < nz; dk++)
            for (uint32_t dj = 0; dj < ny; dj++)
                for (uint32_t di = 0; di < nx; di++)
                    outside |= ((r.x0 + di) + (r.y0 + dj) * P.nbX + (r.z0 + dk) * nbXY) >= P.nBins;
        if (outside) atomicOr(&ctr->err, ERRBIT_OUTSIDE);
    }
    uint32_t cnt = 0;
    unsigned long long bits = 0;
    for (uint32_t a = 0; a < nA; a++) {
        const PlanePose pp = pose[a];
        if (mask[mask_index(o.fam, pp.fam)] != DEME_DONT_PREVENT_CONTACT) continue;
        if (plane_cd(P, pp, X, Y, Z, rCD)) {
            cnt++;
            bits |= 1ull << a;
        }
    }
    numAnal[s] = cnt;
    planeBits[s] = bits;
    int cx, cy, cz;
    cell_of(G, X, Y, Z, cx, cy, cz);
    const uint32_t cell = (uint32_t)cx + (uint32_t)cy * G.nx + (uint32_t)cz * G.nx * G.ny;
    cellOf[s] = cell;
    slotOf[s] = atomicAdd(&cellCount[cell], 1u);
    CdRec cr;
    cr.x = X;
    cr.y = Y;
    cr.z = Z;
    cr.r = sr.r;  // raw radius: the bin rule needs both (double)r + beta and float(r + beta)
    cr.id = s;
    rec[s] = cr;
}

__global__ __launch_bounds__(256) void k_cell_scatter(uint32_t nS, const uint32_t* __restrict__ cellOf,
                                                      const uint32_t* __restrict__ slotOf,
                                                      const uint32_t* __restrict__ cellStart,
                                                      const CdRec* __restrict__ rec, CdRec* __restrict__ sorted) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const CdRec cr = load_cd(rec, s);
    CdRec* dst = sorted + cellStart[cellOf[s]] + slotOf[s];
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    const uint4* s4 = reinterpret_cast<const uint4*>(&cr);
    d4[0] = s4[0];
    d4[1] = s4[1];
}

// The reference's bin rule for A (lower id) and B: expanded overlap (float radii r + beta), contact point
// (checkSpheresOverlap<double>, DEMHelperKernels.cu:195-222), and the contact point's bin inside both AABB ranges.
__device__ __forceinline__ bool bin_rule(const KParams& P, const CdRec& A, const CdRec& B) {
    float rfa = A.r;
    rfa += P.beta;
    float rfb = B.r;
    rfb += P.beta;
    const double ra = (double)rfa, rb = (double)rfb;
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
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
    const uint32_t ix = sat_u32(cx / P.binSize), iy = sat_u32(cy / P.binSize), iz = sat_u32(cz / P.binSize);
    double rcdA = (double)A.r;
    rcdA += (double)P.beta;
    double rcdB = (double)B.r;
    rcdB += (double)P.beta;
    const BinRange ga = bin_range(P, A.x, A.y, A.z, rcdA);
    const BinRange gb = bin_range(P, B.x, B.y, B.z, rcdB);
    if (range_in_grid(P, ga) && range_in_grid(P, gb)) {
        // no aliasing: a linear id is one (i, j, k) -- unless the point itself lies past the grid, where the
        // reference's linear id aliases; fall through to the exact enumeration then
        if (ix < P.nbX && iy < P.nbY && iz < P.nbZ)
            return ix >= ga.x0 && ix <= ga.x1 && iy >= ga.y0 && iy <= ga.y1 && iz >= ga.z0 && iz <= ga.z1 &&
                   ix >= gb.x0 && ix <= gb.x1 && iy >= gb.y0 && iy <= gb.y1 && iz >= gb.z0 && iz <= gb.z1;
    }
    const uint32_t lin = ix + iy * P.nbX + iz * P.nbX * P.nbY;
    return range_has_linear(P, ga, lin) && range_has_linear(P, gb, lin);
}

// One lane per sphere in cell order: scan 27 neighbour cells, keep partners with larger ids.
__global__ __launch_bounds__(256) void k_cell_pairs(KParams P, CellGrid G, uint32_t nS,
                                                    const CdRec* __restrict__ sorted,
                                                    const uint32_t* __restrict__ cellStart,
                                                    const SphereRec* __restrict__ sph,
                                                    const OwnerRec* __restrict__ own,
                                                    const uint8_t* __restrict__ mask, uint32_t slabCap,
                                                    uint32_t* __restrict__ slab, uint32_t* __restrict__ rowCount,
                                                    uint32_t* bCount, DevCounters* ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // a flagged assignment error (oversized or out-of-world ranges) aborts the pass before any range is walked
    if (i >= nS || ctr->err) return;
    const CdRec me = load_cd(sorted, i);
    int cx, cy, cz;
    cell_of(G, me.x, me.y, me.z, cx, cy, cz);
    const uint32_t myOwner = sph[me.id].owner;
    const uint32_t myFam = own[myOwner].fam;
    uint32_t n = 0;
    uint32_t* row = slab + (uint64_t)me.id * slabCap;
    const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < (int)G.nz - 1 ? cz + 1 : cz;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < (int)G.ny - 1 ? cy + 1 : cy;
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < (int)G.nx - 1 ? cx + 1 : cx;
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++) {
            // cells x0..x1 of one (y, z) row are contiguous in the sorted array
            const uint32_t c0 = (uint32_t)x0 + (uint32_t)y * G.nx + (uint32_t)z * G.nx * G.ny;
            const uint32_t b = cellStart[c0], e = cellStart[c0 + (uint32_t)(x1 - x0) + 1];
            for (uint32_t j = b; j < e; j++) {
                const CdRec ot = load_cd(sorted, j);
                if (ot.id <= me.id) continue;
                if (!bin_rule(P, me, ot)) continue;
                const uint32_t ob = sph[ot.id].owner;
                if (ob == myOwner) continue;
                if (mask[mask_index(myFam, own[ob].fam)] != DEME_DONT_PREVENT_CONTACT) continue;
                if (n < slabCap) row[n] = ot.id;
                n++;
                atomicAdd(&bCount[ob], 1u);
            }
        }
    rowCount[me.id] = n;
    if (n > slabCap) atomicMax(&ctr->maxOcc, n);
}

// Canonical rows: plane contacts (ascending plane index) then partners (ascending id); idA for the row.
__global__ __launch_bounds__(256) void k_cell_rows(uint32_t nS, uint32_t slabCap, const uint32_t* __restrict__ slab,
                                                   const uint32_t* __restrict__ rowStart,
                                                   const uint32_t* __restrict__ numAnal,
                                                   const unsigned long long* __restrict__ planeBits,
                                                   uint32_t* __restrict__ idA, uint32_t* __restrict__ idB,
                                                   uint8_t* __restrict__ ctype) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nS) return;
    const uint32_t st = rowStart[s], en = rowStart[s + 1];
    uint32_t pos = st;
    unsigned long long bits = planeBits[s];
    while (bits) {
        const int a = __ffsll(bits) - 1;
        bits &= bits - 1;
        idB[pos] = (uint32_t)a;
        ctype[pos] = DEME_SPHERE_PLANE_CONTACT;
        pos++;
    }
    const uint32_t sp = pos;
    const uint32_t* src = slab + (uint64_t)s * slabCap;
    for (uint32_t k = 0; sp + k < en; k++) {
        const uint32_t v = src[k];
        uint32_t q = sp + k;
        while (q > sp && idB[q - 1] > v) {
            idB[q] = idB[q - 1];
            q--;
        }
        idB[q] = v;
    }
    for (uint32_t q = sp; q < en; q++) ctype[q] = DEME_SPHERE_SPHERE_CONTACT;
    for (uint32_t q = st; q < en; q++) idA[q] = s;
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_cell_assign(const KParams& P, const CellGrid& G, const SphereRec* sph, const OwnerRec* own,
                              uint32_t nS, const PlanePose* pose, uint32_t nA, const uint8_t* mask,
                              uint32_t* cellCount, uint32_t* cellOf, uint32_t* slotOf, uint16_t* numBins,
                              uint32_t* numAnal, unsigned long long* planeBits, CdRec* rec, DevCounters* ctr,
                              hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_assign<<<blocks(nS, 256), 256, 0, st>>>(P, G, sph, own, nS, pose, nA, mask, cellCount, cellOf, slotOf,
                                                   numBins, numAnal, planeBits, rec, ctr);
    return hipGetLastError();
}

hipError_t launch_cell_scatter(uint32_t nS, const uint32_t* cellOf, const uint32_t* slotOf, const uint32_t* cellStart,
                               const CdRec* rec, CdRec* sorted, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_scatter<<<blocks(nS, 256), 256, 0, st>>>(nS, cellOf, slotOf, cellStart, rec, sorted);
    return hipGetLastError();
}

hipError_t launch_cell_pairs(const KParams& P, const CellGrid& G, uint32_t nS, const CdRec* sorted,
                             const uint32_t* cellStart, const SphereRec* sph, const OwnerRec* own, const uint8_t* mask,
                             uint32_t slabCap, uint32_t* slab, uint32_t* rowCount, uint32_t* bCount, DevCounters* ctr,
                             hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_pairs<<<blocks(nS, 256), 256, 0, st>>>(P, G, nS, sorted, cellStart, sph, own, mask, slabCap, slab, rowCount,
                                                  bCount, ctr);
    return hipGetLastError();
}

hipError_t launch_cell_rows(uint32_t nS, uint32_t slabCap, const uint32_t* slab, const uint32_t* rowStart,
                            const uint32_t* numAnal, const unsigned long long* planeBits, uint32_t* idA, uint32_t* idB,
                            uint8_t* ctype, hipStream_t st) {
    if (nS == 0) return hipSuccess;
    k_cell_rows<<<blocks(nS, 256), 256, 0, st>>>(nS, slabCap, slab, rowStart, numAnal, planeBits, idA, idB, ctype);
    return hipGetLastError();
}

}  // namespace deme_int
