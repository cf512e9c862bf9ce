// Dynamics for gfx950: frictional Hertz-Mindlin contact force with tangential history, owner-keyed force/torque
// reduction without global atomics, and clump integration.
//
// Reference: src/DEM/dT.cpp:1349-1432 and the kernels it launches (DEMPrepForceKernels.cu, DEMCalcForceKernels.cu +
// DEMCustomizablePolicies/FullHertzianForceModel.cu, DEMCollectForceKernels.cu + DEMCubForceCollection.cu,
// DEMIntegrationKernels.cu + IntegrationVelPassOn*.cu).
//
// Every float/double promotion of the reference's expressions is kept (e.g. `2. * E_cnt * sqrt_Rd` is evaluated in
// double and rounded to float), and the device code is compiled without contraction.
#include "deme_internal.h"

// Rows a14-a15, the contact force kernel: generated at build time from jit/force_kernel.tpl with the default
// frictional Hertz-Mindlin model (jit/hertz_frictional.model), the same source the runtime hipRTC path specialises.
#include "force_aot.hip"

namespace deme_int {
using namespace deme_dev;
namespace {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
__device__ __forceinline__ f3 operator*(float b, f3 a) { return mk(b * a.x, b * a.y, b * a.z); }
__device__ __forceinline__ f3 operator/(f3 a, float b) { return mk(a.x / b, a.y / b, a.z / b); }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ void operator+=(f3& a, f3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}
__device__ __forceinline__ void operator-=(f3& a, f3 b) {
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
}
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ void rotq(f3& v, float w, float x, float y, float z) { rotate(v.x, v.y, v.z, w, x, y, z); }

// Row a16 for analytical owners (planes): every sphere-plane contact contributes its B-side term (stored by the
// force kernel) to the plane's owner. Deterministic two-level reduction: lane t of a tile sums contacts t, t + 256, ...
// (16 of them) in order, lanes combine in a fixed tree, tiles combine in a fixed tree in k_anal_final. The contact
// types of a lane's 16 contacts are loaded together (coalesced across the wave), so the gathers of the few plane
// contacts are the only dependent loads.
constexpr uint32_t ANAL_ITEMS = 16;
constexpr uint32_t ANAL_TILE = 256 * ANAL_ITEMS;

__global__ __launch_bounds__(256) void k_anal_collect(uint32_t nC, const uint32_t* __restrict__ idB,
                                                      const uint8_t* __restrict__ ctype,
                                                      const float* __restrict__ gTerm,
                                                      const deme_anal_obj* __restrict__ anal, uint32_t nA,
                                                      uint32_t firstAnal, uint32_t nAnalOwners,
                                                      float* __restrict__ partials) {
    __shared__ float red[256][6];
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
#pragma unroll
    for (uint32_t it = 0; it < ANAL_ITEMS; it++) {
        const uint64_t c = c0 + (uint64_t)it * 256;
        if (ty[it] != DEME_SPHERE_PLANE_CONTACT) continue;
        const uint32_t b = idB[c];
        if (b >= nA) continue;
        const uint32_t k = anal[b].owner - firstAnal;
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
#pragma unroll
    for (uint32_t k = 0; k < MAX_ANAL_OWNERS; k++) {
        if (k >= nAnalOwners) break;
#pragma unroll
        for (int c = 0; c < 6; c++) red[threadIdx.x][c] = acc[k][c];
        __syncthreads();
        for (uint32_t st = 128; st > 0; st >>= 1) {
            if (threadIdx.x < st)
#pragma unroll
                for (int c = 0; c < 6; c++) red[threadIdx.x][c] += red[threadIdx.x + st][c];
            __syncthreads();
        }
        if (threadIdx.x < 6) partials[((uint64_t)blockIdx.x * nAnalOwners + k) * 6 + threadIdx.x] = red[0][threadIdx.x];
        __syncthreads();
    }
}

// One workgroup per (owner, component): lane t sums tiles t, t + 256, ... in order, then a fixed LDS tree.
__global__ __launch_bounds__(256) void k_anal_final(const float* __restrict__ partials, uint32_t nBlocks,
                                                    uint32_t nAnalOwners, float* __restrict__ analAcc) {
    __shared__ float red[256];
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

__device__ __forceinline__ void add_term(f3& a, f3& al, const float* __restrict__ src) {
    const float2* g = reinterpret_cast<const float2*>(src);
    const float2 g0 = g[0], g1 = g[1], g2 = g[2];
    a += mk(g0.x, g0.y, g1.x);
    al += mk(g1.y, g2.x, g2.y);
}

// Rows a16 (clump owners) + a17: owner reduction fused with integrateOwners (DEMIntegrationKernels.cu:170-179).
// A clump's A-side terms are the contiguous rows of its spheres; its B-side terms are the contiguous mirror rows of
// its spheres (ordered by B sphere, then A sphere). Terms are summed sequentially in that order (A side, then B side)
// in float.
__global__ __launch_bounds__(256) void k_owner_update(
    KParams P, uint32_t nO, uint32_t nClumps, uint32_t firstAnal, uint32_t nAnalOwners,
    const uint32_t* __restrict__ ownerSphStart, const uint32_t* __restrict__ rowStart,
    const uint32_t* __restrict__ bRowStart, const float* __restrict__ aTerm, const float* __restrict__ bTerm,
    const float* __restrict__ analAcc, const deme_family_rule* __restrict__ rules, const OwnerRec* own,
    OwnerRec* ownOut, float* __restrict__ accOut, int doIntegrate, float t) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nO) return;
    OwnerRec r = load_owner(own, o);
    f3 a = mk(0, 0, 0), al = mk(0, 0, 0);
    if (o < nClumps) {
        const uint32_t s0 = ownerSphStart[o], s1 = ownerSphStart[o + 1];
        const uint32_t c1 = rowStart[s1];
        for (uint32_t c = rowStart[s0]; c < c1; c++) add_term(a, al, aTerm + (uint64_t)c * 6);
        const uint32_t m1 = bRowStart[s1];
        for (uint32_t m = bRowStart[s0]; m < m1; m++) add_term(a, al, bTerm + (uint64_t)m * 6);
        a = mk(0.f + a.x, 0.f + a.y, 0.f + a.z);
        al = mk(0.f + al.x, 0.f + al.y, 0.f + al.z);
    } else if (o - firstAnal < nAnalOwners) {
        const float* s = analAcc + (uint64_t)(o - firstAnal) * 6;
        a = mk(s[0], s[1], s[2]);
        al = mk(s[3], s[4], s[5]);
    }
    accOut[o] = a.x;
    accOut[(uint64_t)nO + o] = a.y;
    accOut[2ull * nO + o] = a.z;
    accOut[3ull * nO + o] = al.x;
    accOut[4ull * nO + o] = al.y;
    accOut[5ull * nO + o] = al.z;
    if (!doIntegrate) return;

    // ---- integrateVel (DEMIntegrationKernels.cu:179-221) ----
    const deme_family_rule R = rules[r.fam];
    bool LinPrescribed = false, RotPrescribed = false;
    const f3 old_v = mk(r.vx, r.vy, r.vz), old_w = mk(r.wx, r.wy, r.wz);
    if (R.used) {
        if (R.velSet & 1) r.vx = R.vel[0];
        if (R.velSet & 2) r.vy = R.vel[1];
        if (R.velSet & 4) r.vz = R.vel[2];
        if (R.velSet & 8) r.wx = R.vel[3];
        if (R.velSet & 16) r.wy = R.vel[4];
        if (R.velSet & 32) r.wz = R.vel[5];
        LinPrescribed = R.linVelPrescribed;
        RotPrescribed = R.rotVelPrescribed;
    }
    f3 v_update = mk(0, 0, 0), w_update = mk(0, 0, 0);
    if (!LinPrescribed) {
        v_update.x = (a.x + P.Gx) * P.h;
        v_update.y = (a.y + P.Gy) * P.h;
        v_update.z = (a.z + P.Gz) * P.h;
        r.vx += v_update.x;
        r.vy += v_update.y;
        r.vz += v_update.z;
    }
    if (!RotPrescribed) {
        w_update.x = al.x * P.h;
        w_update.y = al.y * P.h;
        w_update.z = al.z * P.h;
        r.wx += w_update.x;
        r.wy += w_update.y;
        r.wz += w_update.z;
    }
    f3 v, w;
    if (P.integrator == DEME_INTEGRATOR_FORWARD_EULER) {
        v = old_v;
        w = old_w;
    } else if (P.integrator == DEME_INTEGRATOR_CENTERED_DIFFERENCE) {
        v = old_v + v_update;
        w = old_w + w_update;
    } else {
        v = old_v + v_update * 0.5f;
        w = old_w + w_update * 0.5f;
    }
    // ---- integratePos (DEMIntegrationKernels.cu:241-301) ----
    double X, Y, Z;
    owner_position(P, r.voxel, r.lx, r.ly, r.lz, X, Y, Z);
    LinPrescribed = false;
    RotPrescribed = false;
    if (R.used) {
        if (R.posSet & 1) X = R.pos[0];
        if (R.posSet & 2) Y = R.pos[1];
        if (R.posSet & 4) Z = R.pos[2];
        LinPrescribed = R.linPosPrescribed;
        RotPrescribed = R.rotPosPrescribed;
    }
    if (!LinPrescribed) {
        X += (double)v.x * P.h;
        Y += (double)v.y * P.h;
        Z += (double)v.z * P.h;
    }
    position_to_voxel(P, X, Y, Z, r.voxel, r.lx, r.ly, r.lz);
    if (!RotPrescribed) {
        double dq0 = 1, dq1 = w.x, dq2 = w.y, dq3 = w.z;
        const double len = sqrt(dq1 * dq1 + dq2 * dq2 + dq3 * dq3);
        const double theta = 0.5 * P.h * len;
        if (len > 0) {
            dq0 = cos(theta);
            const double s = sin(theta) / len;
            dq1 *= s;
            dq2 *= s;
            dq3 *= s;
        }
        // HamiltonProduct<float> (DEMHelperKernels.cu:169-186)
        const float a1 = r.qw, b1 = r.qx, c1 = r.qy, d1 = r.qz;
        const float a2 = (float)dq0, b2 = (float)dq1, c2 = (float)dq2, d2 = (float)dq3;
        r.qw = a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2;
        r.qx = a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2;
        r.qy = a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2;
        r.qz = a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2;
    }
    (void)t;  // numeric prescriptions do not depend on time
    uint4* dst = reinterpret_cast<uint4*>(ownOut + o);
    const uint4* src = reinterpret_cast<const uint4*>(&r);
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
    dst[3] = src[3];
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t launch_force(const KParams& P, uint32_t nC, const uint32_t* idA, const uint32_t* idB, const uint8_t* ctype,
                        const SphereRec* sph, const OwnerRec* own, const float4* moi, const deme_anal_obj* anal,
                        const PlanePose* pose, const deme_material* mats, const float4* histIn, const uint32_t* map,
                        float4* histOut, const uint32_t* mirror, float* aTerm, float* bTerm, float* gTerm,
                        ForceRec* out, hipStream_t st) {
    if (nC == 0) return hipSuccess;
    if (out)
        deme_force_full<<<blocks(nC, 256), 256, 0, st>>>(P, nC, idA, idB, ctype, sph, own, moi, anal, pose, mats, histIn,
                                                       map, histOut, mirror, aTerm, bTerm, gTerm, out);
    else
        deme_force_terms<<<blocks(nC, 256), 256, 0, st>>>(P, nC, idA, idB, ctype, sph, own, moi, anal, pose, mats,
                                                        histIn, map, histOut, mirror, aTerm, bTerm, gTerm, nullptr);
    return hipGetLastError();
}

hipError_t launch_anal_collect(uint32_t nC, const uint32_t* idB, const uint8_t* ctype, const float* gTerm,
                               const deme_anal_obj* anal, uint32_t nA, uint32_t firstAnalOwner, uint32_t nAnalOwners,
                               float* blockPartials, uint32_t* nBlocksOut, hipStream_t st) {
    const uint32_t nb = nC == 0 ? 0 : blocks(nC, ANAL_TILE);
    *nBlocksOut = nb;
    if (nb == 0 || nAnalOwners == 0) return hipSuccess;
    if (nAnalOwners > MAX_ANAL_OWNERS) return hipErrorInvalidValue;
    k_anal_collect<<<nb, 256, 0, st>>>(nC, idB, ctype, gTerm, anal, nA, firstAnalOwner, nAnalOwners, blockPartials);
    return hipGetLastError();
}

hipError_t launch_anal_final(const float* blockPartials, uint32_t nBlocks, uint32_t nAnalOwners, float* analAcc,
                             hipStream_t st) {
    if (nAnalOwners == 0) return hipSuccess;
    if (nAnalOwners > MAX_ANAL_OWNERS) return hipErrorInvalidValue;
    k_anal_final<<<nAnalOwners * 6, 256, 0, st>>>(blockPartials, nBlocks, nAnalOwners, analAcc);
    return hipGetLastError();
}

hipError_t launch_owner_update(const KParams& P, uint32_t nO, uint32_t nClumps, uint32_t firstAnalOwner,
                               uint32_t nAnalOwners, const uint32_t* ownerSphStart, const uint32_t* rowStart,
                               const uint32_t* bRowStart, const float* aTerm, const float* bTerm,
                               const float* analAcc, const deme_family_rule* rules, OwnerRec* own, float* acc,
                               int doIntegrate, float t, hipStream_t st, OwnerRec* ownOut) {
    if (nO == 0) return hipSuccess;
    k_owner_update<<<blocks(nO, 256), 256, 0, st>>>(P, nO, nClumps, firstAnalOwner, nAnalOwners, ownerSphStart,
                                                    rowStart, bRowStart, aTerm, bTerm, analAcc, rules, own,
                                                    ownOut ? ownOut : own, acc, doIntegrate, t);
    return hipGetLastError();
}

}  // namespace deme_int
