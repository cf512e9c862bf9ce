// Dynamics for gfx950: frictional Hertz-Mindlin contact force with tangential history, owner-keyed force/torque
// reduction without global atomics, and clump integration.
//
// Reference: src/DEM/dT.cpp:1349-1432 and the kernels it launches (DEMPrepForceKernels.cu, DEMCalcForceKernels.cu +
// DEMCustomizablePolicies/FullHertzianForceModel.cu, DEMCollectForceKernels.cu + DEMCubForceCollection.cu,
// DEMIntegrationKernels.cu + IntegrationVelPassOn*.cu).
//
// Every float/double promotion of the reference's expressions is kept (e.g. `2. * E_cnt * sqrt_Rd` is evaluated in
// double and rounded to float), and the device code is compiled without contraction.
#include <algorithm>

#include "deme_internal.h"

// Rows a14-a15, the contact force kernel: generated at build time from jit/force_kernel.tpl with the default
// frictional Hertz-Mindlin model (jit/hertz_frictional.model), the same source the runtime hipRTC path specialises.
#include "force_aot.hip"
// Rows a16 (clump owners) + a17, the owner update: generated at build time from jit/owner_kernel.tpl with no motion
// strategies (numeric prescriptions from the family rule table).
#include "owner_aot.hip"

namespace deme_int {
using namespace deme_dev;
namespace {

// Row a16 for analytical owners (planes): every sphere-plane contact contributes its B-side term (stored by the
// force kernel) to the plane's owner. Deterministic two-level reduction over tiles of ANAL_TILE contacts: lane t of a
// tile sums contacts t, t + 256, ... (ANAL_ITEMS = 4 of them) in order, lanes combine by wave butterflies and the four
// waves in order, tiles combine in a fixed tree in k_anal_final. The contact types of a lane's contacts are loaded
// together (coalesced across the wave), so the gathers of the few plane contacts are the only dependent loads.

__global__ __launch_bounds__(256) void k_anal_collect(uint32_t nC, const uint32_t* __restrict__ idB,
                                                      const uint8_t* __restrict__ ctype,
                                                      const float* __restrict__ gTerm,
                                                      const deme_anal_obj* __restrict__ anal, uint32_t nA,
                                                      uint32_t firstAnal, uint32_t nAnalOwners,
                                                      float* __restrict__ partials, SpecGuard guard) {
    __shared__ float wsum[4][6];
    if (guard.ctr) {
        // captured with its detection pass: the grid covers the capacity; tiles past the pass's count write nothing
        // (the final pass reads only the tiles of the count)
        if (spec_skip(guard)) return;
        nC = (uint32_t)guard.ctr->total[1];
        if ((uint64_t)blockIdx.x * ANAL_TILE >= nC && blockIdx.x > 0) return;
    }
    float acc[MAX_ANAL_OWNERS][6];
#pragma unroll
    for (uint32_t k = 0; k < MAX_ANAL_OWNERS; k++)
#pragma unroll
        for (int c = 0; c < 6; c++) acc[k][c] = 0.f;
    const uint64_t c0 = (uint64_t)blockIdx.x * ANAL_TILE + threadIdx.x;
    bool any = false;
    uint8_t ty[ANAL_ITEMS];
#pragma unroll
    for (uint32_t it = 0; it < ANAL_ITEMS; it++) {
        const uint64_t c = c0 + (uint64_t)it * 256;
        ty[it] = c < nC ? ctype[c] : (uint8_t)DEME_NOT_A_CONTACT;
    }
    // the dependent gathers are issued a stage at a time for all the lane's contacts (component id, its owner, the
    // term), so their latencies overlap instead of adding up per contact
    uint32_t own[ANAL_ITEMS];
#pragma unroll
    for (uint32_t it = 0; it < ANAL_ITEMS; it++)
        own[it] = ty[it] == DEME_SPHERE_PLANE_CONTACT ? idB[c0 + (uint64_t)it * 256] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t it = 0; it < ANAL_ITEMS; it++) own[it] = own[it] < nA ? anal[own[it]].owner - firstAnal : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t it = 0; it < ANAL_ITEMS; it++) {
        const uint64_t c = c0 + (uint64_t)it * 256;
        if (own[it] == 0xFFFFFFFFu) continue;
        const uint32_t k = own[it];
        const float2* g = reinterpret_cast<const float2*>(gTerm + c * 6);
        const float2 g0 = g[0], g1 = g[1], g2 = g[2];
        any = true;
#pragma unroll
        for (uint32_t kk = 0; kk < MAX_ANAL_OWNERS; kk++) {
            if (kk == k) {
                acc[kk][0] += g0.x;
                acc[kk][1] += g0.y;
                acc[kk][2] += g1.x;
                acc[kk][3] += g1.y;
                acc[kk][4] += g2.x;
                acc[kk][5] += g2.y;
            }
        }
    }
    if (!__syncthreads_or(any)) {
        // a tile without plane contacts contributes exact zeros; skip the tree
        if (threadIdx.x < nAnalOwners * 6) partials[(uint64_t)blockIdx.x * nAnalOwners * 6 + threadIdx.x] = 0.f;
        return;
    }
    // per owner: 64-lane butterflies, then the four waves in order (a fixed order: deterministic)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t k = 0; k < MAX_ANAL_OWNERS; k++) {
        if (k >= nAnalOwners) break;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            float v = acc[k][c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) wsum[wave][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const uint32_t c = threadIdx.x;
            partials[((uint64_t)blockIdx.x * nAnalOwners + k) * 6 + c] =
                ((wsum[0][c] + wsum[1][c]) + wsum[2][c]) + wsum[3][c];
        }
        __syncthreads();
    }
}

// One workgroup per (owner, component): lane t sums tiles t, t + 256, ... in order, then a fixed LDS tree.
__global__ __launch_bounds__(256) void k_anal_final(const float* __restrict__ partials, uint32_t nBlocks,
                                                    uint32_t nAnalOwners, float* __restrict__ analAcc,
                                                    SpecGuard guard) {
    __shared__ float red[256];
    if (guard.ctr) {
        if (spec_skip(guard)) return;
        const uint64_t nb = (guard.ctr->total[1] + ANAL_TILE - 1) / ANAL_TILE;
        nBlocks = nb > 0 ? (uint32_t)nb : 1u;  // an empty list still has its one (zero) tile
    }
    const uint32_t v = blockIdx.x, t = threadIdx.x, stride = nAnalOwners * 6;
    float s = 0.f;
    for (uint32_t b = t; b < nBlocks; b += 256) s += partials[(uint64_t)b * stride + v];
    red[t] = s;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) analAcc[v] = red[0];
}

// The simulation time read by the owner update's motion rules (a kernel, not a copy: it may sit in a captured graph).
__global__ void k_set_clock(double* clock, double t) {
    if (threadIdx.x == 0) clock[0] = t;
}

__global__ void k_advance_clock(double* clock, float h) {
    if (threadIdx.x == 0) clock[0] += (double)h;
}

// ChangeFamily (APIPublic.cpp:184-207): every owner of family `from` moves to family `to`.
__global__ __launch_bounds__(256) void k_change_family(OwnerRec* own, uint32_t nO, uint8_t from, uint8_t to) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o < nO && own[o].fam == from) own[o].fam = to;
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_force(const KParams& P, uint32_t nC, const uint32_t* idA, const uint32_t* idB, const uint8_t* ctype,
                        const SphereRec* sph, const OwnerRec* own, const float4* moi, const deme_anal_obj* anal,
                        const PlanePose* pose, const deme_material* mats, const float4* histIn, const uint32_t* map,
                        float4* histOut, const uint32_t* mirror, float* aTerm, float* bTerm, float* gTerm,
                        uint8_t* aOn, uint8_t* bOn, ForceRec* out, hipStream_t st, SpecGuard guard) {
    if (nC == 0) return hipSuccess;
    if (out)
        deme_force_full<<<blocks(nC, 256), 256, 0, st>>>(P, nC, idA, idB, ctype, sph, own, moi, anal, pose, mats, histIn,
                                                       map, histOut, mirror, aTerm, bTerm, gTerm, aOn, bOn, out,
                                                       guard);
    else
        deme_force_terms<<<blocks(nC, 256), 256, 0, st>>>(P, nC, idA, idB, ctype, sph, own, moi, anal, pose, mats,
                                                        histIn, map, histOut, mirror, aTerm, bTerm, gTerm, aOn, bOn,
                                                        nullptr, guard);
    return hipGetLastError();
}

hipError_t launch_anal_collect(uint32_t nC, const uint32_t* idB, const uint8_t* ctype, const float* gTerm,
                               const deme_anal_obj* anal, uint32_t nA, uint32_t firstAnalOwner, uint32_t nAnalOwners,
                               float* blockPartials, uint64_t partialCap, uint32_t* nBlocksOut, hipStream_t st,
                               SpecGuard guard) {
    // with a guard the list may come out empty: one tile still writes its zero partials for the final pass
    const uint32_t nb = guard.ctr ? std::max(anal_blocks(nC), 1u) : anal_blocks(nC);
    *nBlocksOut = nb;
    if (nb == 0 || nAnalOwners == 0) return hipSuccess;
    if (nAnalOwners > MAX_ANAL_OWNERS) return hipErrorInvalidValue;
    // every workgroup writes its tile's nAnalOwners x 6 partials, plane contacts or not
    if ((uint64_t)nb * nAnalOwners * 6 > partialCap) return hipErrorInvalidValue;
    k_anal_collect<<<nb, 256, 0, st>>>(nC, idB, ctype, gTerm, anal, nA, firstAnalOwner, nAnalOwners, blockPartials,
                                       guard);
    return hipGetLastError();
}

hipError_t launch_anal_final(const float* blockPartials, uint32_t nBlocks, uint32_t nAnalOwners, float* analAcc,
                             hipStream_t st, SpecGuard guard) {
    if (nAnalOwners == 0) return hipSuccess;
    if (nAnalOwners > MAX_ANAL_OWNERS) return hipErrorInvalidValue;
    k_anal_final<<<nAnalOwners * 6, 256, 0, st>>>(blockPartials, nBlocks, nAnalOwners, analAcc, guard);
    return hipGetLastError();
}

hipError_t launch_owner_update(const KParams& P, uint32_t nO, uint32_t nClumps, uint32_t firstAnalOwner,
                               uint32_t nAnalOwners, const uint32_t* ownerSphStart, const uint32_t* rowStart,
                               const uint32_t* bRowStart, const float* aTerm, const float* bTerm,
                               const uint8_t* aOn, const uint8_t* bOn, const float* analAcc, const deme_family_rule* rules, OwnerRec* own, float* acc,
                               int doIntegrate, const double* clock, hipStream_t st, OwnerRec* ownOut,
                               SpecGuard guard) {
    if (nO == 0) return hipSuccess;
    deme_owner_update<<<blocks(nO, 256), 256, 0, st>>>(P, nO, nClumps, firstAnalOwner, nAnalOwners, ownerSphStart,
                                                    rowStart, bRowStart, aTerm, bTerm, aOn, bOn, analAcc, rules,
                                                    own,
                                                    ownOut ? ownOut : own, acc, doIntegrate, clock, guard);
    return hipGetLastError();
}

hipError_t launch_set_clock(double* clock, double t, hipStream_t st) {
    k_set_clock<<<1, 64, 0, st>>>(clock, t);
    return hipGetLastError();
}

hipError_t launch_advance_clock(double* clock, float h, hipStream_t st) {
    k_advance_clock<<<1, 64, 0, st>>>(clock, h);
    return hipGetLastError();
}

hipError_t launch_change_family(OwnerRec* own, uint32_t nO, uint8_t from, uint8_t to, hipStream_t st) {
    if (nO == 0) return hipSuccess;
    k_change_family<<<blocks(nO, 256), 256, 0, st>>>(own, nO, from, to);
    return hipGetLastError();
}

}  // namespace deme_int
