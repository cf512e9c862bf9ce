// Device-side records and arithmetic helpers for the DEM hot path on gfx950.
//
// All device code is compiled with -ffp-contract=off and IEEE division/sqrt, so every double/float operation rounds
// exactly like the reference's arithmetic (no fused multiply-adds are introduced).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "../../include/deme/deme_types.h"

namespace deme_dev {

// --------------------------------------------------------------------------------------------------------------
// HBM-resident records. Owner state is kept as one 64-byte record per owner so that the force kernel's random
// owner gathers touch a single 64-byte segment instead of a dozen SoA cache lines.
// --------------------------------------------------------------------------------------------------------------
struct __attribute__((aligned(16))) OwnerRec {
    uint64_t voxel;       // voxelID_t
    uint16_t lx, ly, lz;  // subVoxelPos_t
    uint8_t fam;          // family_t
    uint8_t pad0;
    float qw, qx, qy, qz;  // oriQ
    float vx, vy, vz;      // linear velocity
    float wx, wy, wz;      // body-frame angular velocity (omgBar)
    float mass;
    uint32_t scene;  // independent scene of a batch (0 for a single scene)
};
static_assert(sizeof(OwnerRec) == 64, "OwnerRec must be 64 bytes");

// Sphere component template, flattened per sphere (ClumpCompAcqStratAllFlatten.cu).
struct __attribute__((aligned(16))) SphereRec {
    uint32_t owner;
    float r;
    float rx, ry, rz;
    uint32_t mat;
    uint32_t pad0, pad1;
};
static_assert(sizeof(SphereRec) == 32, "SphereRec must be 32 bytes");

// Per-sphere contact-detection record written once per detection pass: world position (relative to the world's
// left-bottom-front corner) in double, the sphere radius in float, and the sphere id.
struct __attribute__((aligned(16))) CdRec {
    double x, y, z;
    float r;
    uint32_t id;
};
static_assert(sizeof(CdRec) == 32, "CdRec must be 32 bytes");

// A sphere's position at a candidate build, in fixed point: coordinate = q * unit, unit a power of two chosen so that
// +-2^31 units cover twice the bin grid's extent (rounding error at most unit / 2 per axis). INT32_MIN on an axis marks
// a position that did not fit (or was not a number).
struct BuildPos {
    int32_t x, y, z;
};
static_assert(sizeof(BuildPos) == 12, "BuildPos must be 12 bytes");

// Per-contact outputs of the force kernel (contactForces, contactTorque_convToForce, contactPointGeometryA/B).
struct __attribute__((aligned(16))) ForceRec {
    float fx, fy, fz;
    float tx, ty, tz;
    float ax, ay, az;
    float bx, by, bz;
};
static_assert(sizeof(ForceRec) == 48, "ForceRec must be 48 bytes");

// World pose of one analytical component, recomputed whenever owner states change.
struct __attribute__((aligned(16))) PlanePose {
    double px, py, pz;
    float nx, ny, nz;
    uint32_t owner;
    uint32_t fam;
    uint32_t type;
};
static_assert(sizeof(PlanePose) == 48, "PlanePose must be 48 bytes");

// Inspector quantities (DEMInspector::switch_quantity_type, AuxClasses.cpp:58-103) and reductions; the tile of the
// inspection kernels (256 threads, 8 coalesced strides each).
enum { INSPECT_CLUMP_MAX_Z = 0, INSPECT_CLUMP_MIN_Z = 1, INSPECT_CLUMP_MAX_ABSV = 2, INSPECT_CLUMP_VOLUME = 3 };
enum { INSPECT_MAX = 0, INSPECT_MIN = 1, INSPECT_SUM = 2 };
constexpr int INSPECT_BLOCK = 256, INSPECT_ITEMS = 8;

// Kernel-visible simulation constants.
struct KParams {
    double l, voxelSize, binSize;
    uint32_t nbX, nbY, nbZ;
    uint32_t nBins;  // nbX * nbY * nbZ (< 2^32, checked on the host)
    uint32_t nvX, nvY;
    float beta, h;
    float Gx, Gy, Gz;
    uint32_t integrator;
    uint32_t nS, nA;  // sphere and analytical-component counts (bounds of data-dependent gathers)
    // Analytical components are grouped by scene: scene k owns components planeStart[k] .. planeStart[k + 1] - 1
    // (at most 64). A single scene has planeStart = {0, nA}.
    const uint32_t* planeStart;
    uint32_t nScenes;
};

// Range of analytical components a body of the given scene can touch.
__device__ __forceinline__ void scene_planes(const KParams& P, uint32_t scene, uint32_t& a0, uint32_t& a1) {
    if (P.nScenes == 1) {
        // one scene: every component (no table lookup on the dependent-load chain)
        a0 = 0;
        a1 = scene == 0 ? P.nA : 0;
    } else if (scene < P.nScenes) {
        a0 = P.planeStart[scene];
        a1 = P.planeStart[scene + 1];
    } else {
        a0 = a1 = 0;
    }
}

// --------------------------------------------------------------------------------------------------------------
// Saturating float->integer conversions (the reference's device conversions saturate; NaN -> 0).
// --------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sat_u32(double x) {
    if (!(x > 0.0)) return 0u;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
__device__ __forceinline__ uint16_t sat_u16(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65535.0) return 65535;
    return (uint16_t)(uint32_t)x;
}
__device__ __forceinline__ uint64_t sat_u64(double x) {
    if (!(x > 0.0)) return 0ull;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)x;
}

// voxelIDToPosition<double, voxelID_t, subVoxelPos_t> (DEMHelperKernels.cu:93-110)
__device__ __forceinline__ void owner_position(const KParams& P, uint64_t id, uint16_t lx, uint16_t ly, uint16_t lz,
                                               double& X, double& Y, double& Z) {
    const uint64_t vx = id & (((uint64_t)1 << P.nvX) - 1);
    const uint64_t vy = (id >> P.nvX) & (((uint64_t)1 << P.nvY) - 1);
    const uint64_t vz = id >> (P.nvX + P.nvY);
    X = (double)vx * P.voxelSize + (double)lx * P.l;
    Y = (double)vy * P.voxelSize + (double)ly * P.l;
    Z = (double)vz * P.voxelSize + (double)lz * P.l;
}

// positionToVoxelID<voxelID_t, subVoxelPos_t, double> (DEMHelperKernels.cu:113-135)
__device__ __forceinline__ void position_to_voxel(const KParams& P, double X, double Y, double Z, uint64_t& id,
                                                  uint16_t& lx, uint16_t& ly, uint16_t& lz) {
    const uint64_t nx = sat_u64(X / P.voxelSize), ny = sat_u64(Y / P.voxelSize), nz = sat_u64(Z / P.voxelSize);
    lx = sat_u16((X - (double)nx * P.voxelSize) / P.l);
    ly = sat_u16((Y - (double)ny * P.voxelSize) / P.l);
    lz = sat_u16((Z - (double)nz * P.voxelSize) / P.l);
    id = nx + (ny << P.nvX) + (nz << (P.nvX + P.nvY));
}

// applyOriQToVector3<float, float> (DEMHelperKernels.cu:137-148)
__device__ __forceinline__ void rotate(float& X, float& Y, float& Z, float Qw, float Qx, float Qy, float Qz) {
    const float oX = X, oY = Y, oZ = Z;
    X = (2.0f * (Qw * Qw + Qx * Qx) - 1.0f) * oX + (2.0f * (Qx * Qy - Qw * Qz)) * oY + (2.0f * (Qx * Qz + Qw * Qy)) * oZ;
    Y = (2.0f * (Qx * Qy + Qw * Qz)) * oX + (2.0f * (Qw * Qw + Qy * Qy) - 1.0f) * oY + (2.0f * (Qy * Qz - Qw * Qx)) * oZ;
    Z = (2.0f * (Qx * Qz - Qw * Qy)) * oX + (2.0f * (Qy * Qz + Qw * Qx)) * oY + (2.0f * (Qw * Qw + Qz * Qz) - 1.0f) * oZ;
}

// locateMaskPair (DEMHelperKernels.cu:51-56)
__device__ __forceinline__ uint32_t mask_index(uint32_t i, uint32_t j) {
    if (i > j) {
        const uint32_t t = i;
        i = j;
        j = t;
    }
    return (1 + j) * j / 2 + i;
}

// Sphere world position as every reference kernel derives it: owner position in double plus the float-rotated
// component offset (DEMBinSphereKernels.cu:39-51).
__device__ __forceinline__ void sphere_position(const KParams& P, const OwnerRec& o, const SphereRec& s, double& X,
                                                double& Y, double& Z) {
    double ox, oy, oz;
    owner_position(P, o.voxel, o.lx, o.ly, o.lz, ox, oy, oz);
    float rx = s.rx, ry = s.ry, rz = s.rz;
    rotate(rx, ry, rz, o.qw, o.qx, o.qy, o.qz);
    X = ox + (double)rx;
    Y = oy + (double)ry;
    Z = oz + (double)rz;
}

__device__ __forceinline__ OwnerRec load_owner(const OwnerRec* __restrict__ p, uint32_t i) {
    // four 16-byte loads
    const uint4* q = reinterpret_cast<const uint4*>(p + i);
    uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    OwnerRec r;
    uint4* w = reinterpret_cast<uint4*>(&r);
    w[0] = a;
    w[1] = b;
    w[2] = c;
    w[3] = d;
    return r;
}

__device__ __forceinline__ SphereRec load_sphere(const SphereRec* __restrict__ p, uint32_t i) {
    const uint4* q = reinterpret_cast<const uint4*>(p + i);
    uint4 a = q[0], b = q[1];
    SphereRec r;
    uint4* w = reinterpret_cast<uint4*>(&r);
    w[0] = a;
    w[1] = b;
    return r;
}

__device__ __forceinline__ CdRec load_cd(const CdRec* __restrict__ p, uint32_t i) {
    const uint4* q = reinterpret_cast<const uint4*>(p + i);
    uint4 a = q[0], b = q[1];
    CdRec r;
    uint4* w = reinterpret_cast<uint4*>(&r);
    w[0] = a;
    w[1] = b;
    return r;
}

// Device-side counters read back by the host after a detection pass.
struct DevCounters {
    uint32_t pairCount;
    uint32_t err;
    uint32_t maxOcc;
    uint32_t pad;
    unsigned long long nActive;
    unsigned long long total[4];
    // largest squared displacement (bits of a non-negative double) of a sphere since the candidate lists were built,
    // spread over DISP_SLOTS words (workgroup b folds into word b % DISP_SLOTS: one address per workgroup would
    // serialize the atomics); the largest word is the maximum
    unsigned long long disp2[64];
    // per axis, the extremes of the spheres' displacements since the build (same slots; order-preserving integer
    // images of the doubles, the minima complemented, so every word folds with atomicMax from a cleared zero)
    unsigned long long dispLo[3][64], dispHi[3][64];
};
constexpr uint32_t DISP_SLOTS = 64;
// The host reads the counters up to the displacement extremes (the device alone uses those).
constexpr size_t DEV_COUNTERS_HOST_BYTES = __builtin_offsetof(DevCounters, dispLo);

// Order-preserving image of a double as an unsigned 64-bit integer (and back).
__host__ __device__ inline unsigned long long ordered_bits(double x) {
    unsigned long long b;
    __builtin_memcpy(&b, &x, sizeof(b));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(unsigned long long b) {
    b = (b >> 63) ? (b & 0x7FFFFFFFFFFFFFFFull) : ~b;
    double x;
    __builtin_memcpy(&x, &b, sizeof(x));
    return x;
}

__host__ __device__ inline double max_disp2(const DevCounters& c) {
    unsigned long long m = 0;
    for (uint32_t k = 0; k < DISP_SLOTS; k++) m = c.disp2[k] > m ? c.disp2[k] : m;
    double d;
    __builtin_memcpy(&d, &m, sizeof(d));
    return d;
}

// ERRBIT_STALE: a sphere moved too far since the candidate lists were built; they must be rebuilt and the pass redone
enum : uint32_t { ERRBIT_CAPACITY = 1u, ERRBIT_PAIR_OVERFLOW = 2u, ERRBIT_OUTSIDE = 4u, ERRBIT_STALE = 8u };

// Kernels enqueued before the host has read a detection pass's counters (so the GPU does not wait for the host, or
// the whole step sits in a captured graph) exit at once, writing nothing, when the pass failed, a slab row
// overflowed, or the contact / mirror-row totals exceed the buffers they were given; the host then redoes them.
// Kernels over contacts then take the contact count from ctr->total[1]. ctr == nullptr: no guard.
struct SpecGuard {
    const DevCounters* ctr = nullptr;
    uint32_t slabCap = 0;
    unsigned long long capC = 0, capP = 0;
};
__device__ __forceinline__ bool spec_skip(const SpecGuard& g) {
    return g.ctr && (g.ctr->err || g.ctr->maxOcc > g.slabCap || g.ctr->total[1] > g.capC || g.ctr->total[2] > g.capP);
}

}  // namespace deme_dev
