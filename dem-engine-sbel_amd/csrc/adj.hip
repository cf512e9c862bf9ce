// Adjoint (reverse-mode) kernels of the DEM step for systems of sphere owners and fixed analytical walls: gradients of
// a loss on the final state with respect to the initial velocities and the per-material E, nu, CoR, mu, with the
// contact set frozen between detection passes. The math lives in adjoint.h; these kernels run it over owners and
// contacts with the same contiguous-row layout as the forward owner reduction (no atomics, fixed summation order).
//
// Per step, backwards:
//   k_adj_owner_pre   owner adjoints (X, v, w at step n+1) -> adjoints of the step's accelerations, and the parts of
//                     the step-n adjoints that come through the integration itself (the last step of a window only);
//   k_adj_force       per contact: re-evaluate the force with its tape and push the acceleration adjoints back to both
//                     bodies' positions, velocities and angular velocities, the history in, and the materials;
//   k_adj_owner_post  per owner: add the contact contributions (its rows, then its mirror rows), then run the
//                     integration of the step before backwards, as k_adj_owner_pre does.
#include "adjoint.h"
#include "deme_internal.h"

namespace deme_int {
using namespace deme_dev;
using deme_adj::V3;
using deme_adj::v3;
namespace {

__device__ __forceinline__ V3<float> ld3(const float* p, uint64_t i, uint64_t n) {
    return v3<float>(p[i], p[n + i], p[2 * n + i]);
}
__device__ __forceinline__ void st3(float* p, uint64_t i, uint64_t n, V3<float> v) {
    p[i] = v.x;
    p[n + i] = v.y;
    p[2 * n + i] = v.z;
}
__device__ __forceinline__ V3<float> rot_q(V3<float> v, const OwnerRec& o) {
    float x = v.x, y = v.y, z = v.z;
    rotate(x, y, z, o.qw, o.qx, o.qy, o.qz);
    return v3(x, y, z);
}

// One owner through the integration of a step, backwards: s (the adjoints of X, v, w after the step) becomes the
// adjoints before it, through the integration alone, and lAcc receives the adjoints of the step's world linear and
// angular accelerations. accSnap: 6 x nO SoA (a, alpha in the body frame) of the step; ownSnap: records before it.
__device__ __forceinline__ void owner_pre_one(uint32_t o, uint32_t nO, float h, const OwnerRec* __restrict__ ownSnap,
                                              const float* __restrict__ accSnap, float s[9],
                                              float* __restrict__ lAcc) {
    const OwnerRec r = load_owner(ownSnap, o);
    const V3<float> lXn = v3(s[0], s[1], s[2]);
    const V3<float> lv = v3(s[3], s[4], s[5]);
    const V3<float> lw = v3(s[6], s[7], s[8]);
    // X' = X + h v + h^2/2 a, v' = v + h a (a includes gravity, a constant)
    const V3<float> la = (0.5f * h * h) * lXn + h * lv;
    const V3<float> lvn = lv + h * lXn;
    // w' = Rot(h (w + h/2 alpha)) (w + h alpha), world frame
    const V3<float> w = rot_q(v3(r.wx, r.wy, r.wz), r);
    const V3<float> al = rot_q(v3(accSnap[3ull * nO + o], accSnap[4ull * nO + o], accSnap[5ull * nO + o]), r);
    const V3<float> u = w + h * al;
    const V3<float> th = h * (w + (0.5f * h) * al);
    V3<float> lu, lth;
    deme_adj::rot_reverse(th, u, lw, lu, lth);
    const V3<float> lwn = lu + h * lth;
    const V3<float> lal = h * lu + (0.5f * h * h) * lth;
    s[3] = lvn.x;
    s[4] = lvn.y;
    s[5] = lvn.z;
    s[6] = lwn.x;
    s[7] = lwn.y;
    s[8] = lwn.z;
    st3(lAcc, o, nO, la);
    lAcc[3ull * nO + o] = lal.x;
    lAcc[4ull * nO + o] = lal.y;
    lAcc[5ull * nO + o] = lal.z;
}

__device__ __forceinline__ void zero_owner(uint32_t o, uint32_t nO, float* __restrict__ lam,
                                           float* __restrict__ lAcc) {
    // fixed walls: nothing flows through them
    for (int k = 0; k < 9; k++) lam[(uint64_t)k * nO + o] = 0.f;
    for (int k = 0; k < 6; k++) lAcc[(uint64_t)k * nO + o] = 0.f;
}

// lam: 9 x nO SoA (X, v, w) after the step -> before it, through the integration; lAcc as in owner_pre_one.
__global__ __launch_bounds__(256) void k_adj_owner_pre(uint32_t nO, uint32_t nClumps, float h,
                                                       const OwnerRec* __restrict__ ownSnap,
                                                       const float* __restrict__ accSnap, float* __restrict__ lam,
                                                       float* __restrict__ lAcc) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nO) return;
    if (o >= nClumps) {
        zero_owner(o, nO, lam, lAcc);
        return;
    }
    float s[9];
    for (int k = 0; k < 9; k++) s[k] = lam[(uint64_t)k * nO + o];
    owner_pre_one(o, nO, h, ownSnap, accSnap, s, lAcc);
    for (int k = 0; k < 9; k++) lam[(uint64_t)k * nO + o] = s[k];
}

constexpr uint32_t ADJ_BLOCK = 256;

// Per contact; adjA (9 floats per contact) and adjB (9 floats per mirror slot) receive the contributions to the two
// bodies, lDtIn the history-in adjoint; matPart (per block, nMat x 4) the material gradients summed in lane order.
__global__ __launch_bounds__(ADJ_BLOCK) void k_adj_force(
    KParams P, uint32_t nC, uint32_t nClumps, const uint32_t* __restrict__ idA, const uint32_t* __restrict__ idB,
    const uint8_t* __restrict__ ctype, const SphereRec* __restrict__ sph, const OwnerRec* __restrict__ ownSnap,
    const float4* __restrict__ moi, const deme_anal_obj* __restrict__ anal, const PlanePose* __restrict__ pose,
    const deme_material* __restrict__ mats, uint32_t nMat, const float4* __restrict__ histIn,
    const uint32_t* __restrict__ mirror, uint32_t nO, const float* __restrict__ lAcc,
    const float4* __restrict__ lDtOut, float* __restrict__ adjA, float* __restrict__ adjB, float4* __restrict__ lDtIn,
    float* __restrict__ matPart, int accumulate) {
    __shared__ float wsum[ADJ_BLOCK / 64][256];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t mA = 0, mB = 0;
    if (c < nC) {
        const uint32_t a = idA[c], b = idB[c], t = ctype[c];
        const bool ok = a < P.nS && ((t == DEME_SPHERE_SPHERE_CONTACT && b < P.nS) ||
                                     (t == DEME_SPHERE_PLANE_CONTACT && b < P.nA));
        float* dA = adjA + (uint64_t)c * 9;
        for (int k = 0; k < 9; k++) dA[k] = 0.f;
        lDtIn[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) {
            deme_adj::ContactIn<float> in;
            const SphereRec sa = load_sphere(sph, a);
            const OwnerRec oa = load_owner(ownSnap, sa.owner);
            double X, Y, Z;
            owner_position(P, oa.voxel, oa.lx, oa.ly, oa.lz, X, Y, Z);
            in.XA[0] = X;
            in.XA[1] = Y;
            in.XA[2] = Z;
            in.vA = v3(oa.vx, oa.vy, oa.vz);
            in.wA = rot_q(v3(oa.wx, oa.wy, oa.wz), oa);
            in.rA = sa.r;
            in.mA = oa.mass;
            mA = sa.mat;
            const deme_material& MA = mats[sa.mat];
            in.mA_ = deme_adj::Mat{MA.E, MA.nu, MA.CoR, MA.mu};
            in.h = P.h;
            const float4 w = histIn[c];
            in.dtIn = v3(w.x, w.y, w.z);
            uint32_t ownerB = 0;
            if (t == DEME_SPHERE_SPHERE_CONTACT) {
                const SphereRec sb = load_sphere(sph, b);
                const OwnerRec ob = load_owner(ownSnap, sb.owner);
                owner_position(P, ob.voxel, ob.lx, ob.ly, ob.lz, X, Y, Z);
                in.XB[0] = X;
                in.XB[1] = Y;
                in.XB[2] = Z;
                in.vB = v3(ob.vx, ob.vy, ob.vz);
                in.wB = rot_q(v3(ob.wx, ob.wy, ob.wz), ob);
                in.rB = sb.r;
                in.mB = ob.mass;
                mB = sb.mat;
                ownerB = sb.owner;
                in.plane = false;
                in.nP = v3<float>(0, 0, 0);
            } else {
                const PlanePose pp = pose[b];
                in.XB[0] = pp.px;
                in.XB[1] = pp.py;
                in.XB[2] = pp.pz;
                in.vB = in.wB = v3<float>(0, 0, 0);
                in.rB = 1e15f;
                in.mB = anal[b].mass;
                mB = anal[b].material;
                in.plane = true;
                in.nP = v3(pp.nx, pp.ny, pp.nz);
                ownerB = anal[b].owner;
            }
            const deme_material& MB = mats[mB];
            in.mB_ = deme_adj::Mat{MB.E, MB.nu, MB.CoR, MB.mu};
            deme_adj::ContactTape<float> tape;
            deme_adj::ContactOut<float> out;
            deme_adj::contact_forward(in, tape, out);
            // adjoints of this contact's force and arms from the owners' acceleration adjoints
            const uint32_t oA = sa.owner;
            const V3<float> laA = ld3(lAcc, oA, nO), lalA = ld3(lAcc + 3ull * nO, oA, nO);
            const float IA = moi[oA].x;
            const V3<float> ltA = (1.f / IA) * lalA;
            V3<float> lF = (1.f / oa.mass) * laA;
            deme_adj::acc(lF, deme_adj::cross(ltA, out.armA));
            const V3<float> lArmA = deme_adj::cross(out.F, ltA);
            V3<float> lArmB = v3<float>(0, 0, 0);
            if (!in.plane && ownerB < nClumps) {
                const V3<float> laB = ld3(lAcc, ownerB, nO), lalB = ld3(lAcc + 3ull * nO, ownerB, nO);
                const float IB = moi[ownerB].x;
                const V3<float> ltB = (1.f / IB) * lalB;
                deme_adj::acc(lF, (-1.f / in.mB) * laB);
                deme_adj::acc(lF, -deme_adj::cross(ltB, out.armB));
                lArmB = -deme_adj::cross(out.F, ltB);
            }
            const float4 ld = lDtOut[c];
            deme_adj::ContactAdj<float> adj;
            deme_adj::contact_reverse(in, tape, lF, lArmA, lArmB, v3(ld.x, ld.y, ld.z), adj);
            dA[0] = adj.XA.x;
            dA[1] = adj.XA.y;
            dA[2] = adj.XA.z;
            dA[3] = adj.vA.x;
            dA[4] = adj.vA.y;
            dA[5] = adj.vA.z;
            dA[6] = adj.wA.x;
            dA[7] = adj.wA.y;
            dA[8] = adj.wA.z;
            if (!in.plane) {
                const uint32_t m = mirror[c];
                if (m != DEME_NULL_MAPPING_PARTNER) {
                    float* dB = adjB + (uint64_t)m * 9;
                    dB[0] = adj.XB.x;
                    dB[1] = adj.XB.y;
                    dB[2] = adj.XB.z;
                    dB[3] = adj.vB.x;
                    dB[4] = adj.vB.y;
                    dB[5] = adj.vB.z;
                    dB[6] = adj.wB.x;
                    dB[7] = adj.wB.y;
                    dB[8] = adj.wB.z;
                }
            }
            lDtIn[c] = make_float4(adj.dtIn.x, adj.dtIn.y, adj.dtIn.z, 0.f);
            g[0] = adj.gA.E;
            g[1] = adj.gA.nu;
            g[2] = adj.gA.CoR;
            g[3] = adj.gA.mu;
            g[4] = adj.gB.E;
            g[5] = adj.gB.nu;
            g[6] = adj.gB.CoR;
            g[7] = adj.gB.mu;
        }
    }
    // per-block material sums: for each (material, property) a butterfly sum over each wave's lanes (each lane's A
    // side plus its B side), then the block's waves in order -- a fixed order, so the result is deterministic. A block
    // keeps its slot over the steps of a window, so the slots accumulate in step order.
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t m = 0; m < nMat; m++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float v = (mA == m ? g[q] : 0.f) + (mB == m ? g[4 + q] : 0.f);
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
            if (lane == 0) wsum[wave][m * 4 + q] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < nMat * 4) {
        float s = 0.f;
#pragma unroll
        for (uint32_t w = 0; w < ADJ_BLOCK / 64; w++) s += wsum[w][threadIdx.x];
        float* slot = matPart + (uint64_t)blockIdx.x * nMat * 4 + threadIdx.x;
        *slot = accumulate ? *slot + s : s;
    }
}

// lam += the contact contributions of the owner's rows (A side) and mirror rows (B side), in that order; then, when
// ownPrev is given, the integration of the previous step backwards (owner_pre_one with ownPrev / accPrev).
__global__ __launch_bounds__(256) void k_adj_owner_post(uint32_t nO, uint32_t nClumps,
                                                        const uint32_t* __restrict__ ownerSphStart,
                                                        const uint32_t* __restrict__ rowStart,
                                                        const uint32_t* __restrict__ bRowStart,
                                                        const float* __restrict__ adjA, const float* __restrict__ adjB,
                                                        float* __restrict__ lam, float h,
                                                        const OwnerRec* __restrict__ ownPrev,
                                                        const float* __restrict__ accPrev, float* __restrict__ lAcc) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nO) return;
    if (o >= nClumps) {
        if (ownPrev) zero_owner(o, nO, lam, lAcc);
        return;
    }
    float s[9];
    for (int k = 0; k < 9; k++) s[k] = lam[(uint64_t)k * nO + o];
    const uint32_t s0 = ownerSphStart[o], s1 = ownerSphStart[o + 1];
    for (uint32_t c = rowStart[s0]; c < rowStart[s1]; c++)
        for (int k = 0; k < 9; k++) s[k] += adjA[(uint64_t)c * 9 + k];
    for (uint32_t m = bRowStart[s0]; m < bRowStart[s1]; m++)
        for (int k = 0; k < 9; k++) s[k] += adjB[(uint64_t)m * 9 + k];
    if (ownPrev) owner_pre_one(o, nO, h, ownPrev, accPrev, s, lAcc);
    for (int k = 0; k < 9; k++) lam[(uint64_t)k * nO + o] = s[k];
}

// Material gradients: per (material, property), the block partials summed in block order.
__global__ __launch_bounds__(256) void k_adj_mat_final(const float* __restrict__ matPart, uint32_t nBlocks,
                                                       uint32_t n, float* __restrict__ grad) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float s = 0.f;
    for (uint32_t b = 0; b < nBlocks; b++) s += matPart[(uint64_t)b * n + v];
    grad[v] += s;
}

// History in of a window's first step, gathered from the previous list through the persistent-contact map.
__global__ __launch_bounds__(256) void k_adj_gather_hist(uint32_t nC, const uint32_t* __restrict__ map, uint32_t nPrev,
                                                         const float4* __restrict__ prev, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC) return;
    const uint32_t m = map ? map[i] : DEME_NULL_MAPPING_PARTNER;
    out[i] = (m == DEME_NULL_MAPPING_PARTNER || m >= nPrev) ? make_float4(0.f, 0.f, 0.f, 0.f) : prev[m];
}

// The reverse of that gather: the previous list's history adjoint (cleared first, map is injective).
__global__ __launch_bounds__(256) void k_adj_scatter_hist(uint32_t nC, const uint32_t* __restrict__ map,
                                                          uint32_t nPrev, const float4* __restrict__ lIn,
                                                          float4* __restrict__ lPrev) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC || !map) return;
    const uint32_t m = map[i];
    if (m != DEME_NULL_MAPPING_PARTNER && m < nPrev) lPrev[m] = lIn[i];
}

__global__ void k_fill_f4(float4* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

uint32_t adj_force_blocks(uint32_t nC) { return blocks(nC, ADJ_BLOCK); }

hipError_t launch_adj_owner_pre(uint32_t nO, uint32_t nClumps, float h, const OwnerRec* ownSnap, const float* accSnap,
                                float* lam, float* lAcc, hipStream_t st) {
    if (nO == 0) return hipSuccess;
    k_adj_owner_pre<<<blocks(nO, 256), 256, 0, st>>>(nO, nClumps, h, ownSnap, accSnap, lam, lAcc);
    return hipGetLastError();
}

hipError_t launch_adj_force(const KParams& P, uint32_t nC, uint32_t nClumps, const uint32_t* idA, const uint32_t* idB,
                            const uint8_t* ctype, const SphereRec* sph, const OwnerRec* ownSnap, const float4* moi,
                            const deme_anal_obj* anal, const PlanePose* pose, const deme_material* mats, uint32_t nMat,
                            const float4* histIn, const uint32_t* mirror, uint32_t nO, const float* lAcc,
                            const float4* lDtOut, float* adjA, float* adjB, float4* lDtIn, float* matPart,
                            int accumulate, hipStream_t st) {
    if (nC == 0) return hipSuccess;
    if (nMat * 4 > ADJ_BLOCK) return hipErrorInvalidValue;
    k_adj_force<<<blocks(nC, ADJ_BLOCK), ADJ_BLOCK, 0, st>>>(P, nC, nClumps, idA, idB, ctype, sph, ownSnap, moi, anal,
                                                            pose, mats, nMat, histIn, mirror, nO, lAcc, lDtOut, adjA,
                                                            adjB, lDtIn, matPart, accumulate);
    return hipGetLastError();
}

hipError_t launch_adj_owner_post(uint32_t nO, uint32_t nClumps, const uint32_t* ownerSphStart, const uint32_t* rowStart,
                                 const uint32_t* bRowStart, const float* adjA, const float* adjB, float* lam,
                                 float h, const OwnerRec* ownPrev, const float* accPrev, float* lAcc, hipStream_t st) {
    if (nO == 0) return hipSuccess;
    k_adj_owner_post<<<blocks(nO, 256), 256, 0, st>>>(nO, nClumps, ownerSphStart, rowStart, bRowStart, adjA, adjB,
                                                      lam, h, ownPrev, accPrev, lAcc);
    return hipGetLastError();
}

hipError_t launch_adj_mat_final(const float* matPart, uint32_t nBlocks, uint32_t n, float* grad, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_adj_mat_final<<<blocks(n, 256), 256, 0, st>>>(matPart, nBlocks, n, grad);
    return hipGetLastError();
}

hipError_t launch_adj_gather_hist(uint32_t nC, const uint32_t* map, uint32_t nPrev, const float4* prev, float4* out,
                                  hipStream_t st) {
    if (nC == 0) return hipSuccess;
    k_adj_gather_hist<<<blocks(nC, 256), 256, 0, st>>>(nC, map, nPrev, prev, out);
    return hipGetLastError();
}

hipError_t launch_adj_scatter_hist(uint32_t nC, const uint32_t* map, uint32_t nPrev, const float4* lIn, float4* lPrev,
                                   hipStream_t st) {
    if (nPrev) {
        uint32_t nb = blocks(nPrev, 256);
        if (nb > 65536) nb = 65536;
        k_fill_f4<<<nb, 256, 0, st>>>(lPrev, nPrev);
    }
    if (nC == 0 || !map) return hipGetLastError();
    k_adj_scatter_hist<<<blocks(nC, 256), 256, 0, st>>>(nC, map, nPrev, lIn, lPrev);
    return hipGetLastError();
}

hipError_t launch_fill_f4(float4* p, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint32_t nb = blocks(n, 256);
    if (nb > 65536) nb = 65536;
    k_fill_f4<<<nb, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace deme_int
