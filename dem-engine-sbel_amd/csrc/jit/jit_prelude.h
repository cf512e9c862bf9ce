// Names a contact force model body can use, with the meaning the reference's force-model strings give them
// (DEMCalcForceKernels.cu:14-179, Models.h:197-252, DEMHelperKernels.cu:137-148, 316-339, Defines.h:24-31):
// float3 arithmetic (HIP's vector types), dot / cross / length / normalize, applyOriQToVector3,
// matProxy2ContactParam, the deme:: constants and DEME_TINY_FLOAT / DEME_HUGE_FLOAT. Included by the force kernel
// template both when it is compiled ahead of time and when hipRTC specialises it at run time.
#pragma once

#include "../deme_device.h"

#define DEME_TINY_FLOAT 1e-12
#define DEME_HUGE_FLOAT 1e15

namespace deme {
typedef float oriQ_t;
typedef uint8_t family_t;
typedef uint8_t contact_t;
typedef uint8_t materialsOffset_t;
typedef uint32_t bodyID_t;
constexpr double PI = 3.1415926535897932385;
constexpr double PI_SQUARED = 9.869604401089358;
constexpr double TWO_OVER_THREE = 0.666666666666667;
constexpr double FOUR_OVER_THREE = 1.333333333333333;
constexpr double FIVE_OVER_THREE = 1.666666666666667;
constexpr double TWO_TIMES_SQRT_FIVE_OVER_SIX = 1.825741858350554;
constexpr uint8_t NOT_A_CONTACT = DEME_NOT_A_CONTACT;
constexpr uint8_t SPHERE_SPHERE_CONTACT = DEME_SPHERE_SPHERE_CONTACT;
constexpr uint8_t SPHERE_PLANE_CONTACT = DEME_SPHERE_PLANE_CONTACT;
}  // namespace deme

__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(float3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ float3 normalize(float3 v) {
    const float m = length(v);
    return make_float3(v.x / m, v.y / m, v.z / m);
}

// Rotation of (X, Y, Z) by the unit quaternion (Qw, Qx, Qy, Qz), rotation-matrix form.
template <typename T1, typename T2>
__device__ __forceinline__ void applyOriQToVector3(T1& X, T1& Y, T1& Z, T2 Qw, T2 Qx, T2 Qy, T2 Qz) {
    deme_dev::rotate(X, Y, Z, Qw, Qx, Qy, Qz);
}

// Effective contact parameters of a material pair: E* = 1 / sum (1 - nu^2) / E, G* = 1 / sum 2 (2 - nu)(1 + nu) / E,
// CoR the smaller, mu and Crr the larger (the reference's promotions: the sums in double, rounded once).
template <typename T>
__device__ __forceinline__ void matProxy2ContactParam(T& E_eff, T& G_eff, T& CoR_eff, T& mu_eff, T& Crr_eff, T E_A,
                                                      T nu_A, T CoR_A, T mu_A, T Crr_A, T E_B, T nu_B, T CoR_B,
                                                      T mu_B, T Crr_B) {
    const T invE = (1. - nu_A * nu_A) / E_A + (1. - nu_B * nu_B) / E_B;
    E_eff = 1. / invE;
    const T invG = 2. * (2. - nu_A) * (1. + nu_A) / E_A + 2. * (2. - nu_B) * (1. + nu_B) / E_B;
    G_eff = 1. / invG;
    CoR_eff = fminf(CoR_A, CoR_B);
    mu_eff = fmaxf(mu_A, mu_B);
    Crr_eff = fmaxf(Crr_A, Crr_B);
}

// The two-output form (frictionless models): E* and CoR only. Model strings also call it mat2ProxyContactParam.
template <typename T>
__device__ __forceinline__ void matProxy2ContactParam(T& E_eff, T& CoR_eff, T E_A, T nu_A, T CoR_A, T E_B, T nu_B,
                                                      T CoR_B) {
    const T invE = (1. - nu_A * nu_A) / E_A + (1. - nu_B * nu_B) / E_B;
    E_eff = 1. / invE;
    CoR_eff = fminf(CoR_A, CoR_B);
}
template <typename T>
__device__ __forceinline__ void mat2ProxyContactParam(T& E_eff, T& CoR_eff, T E_A, T nu_A, T CoR_A, T E_B, T nu_B,
                                                      T CoR_B) {
    matProxy2ContactParam<T>(E_eff, CoR_eff, E_A, nu_A, CoR_A, E_B, nu_B, CoR_B);
}

namespace deme_jit {
// A material property read as an array indexed by material (E[bodyAMatType]) from the uploaded material table.
struct MatProp {
    const deme_material* m;
    int k;
    __device__ __forceinline__ float operator[](uint32_t i) const { return reinterpret_cast<const float*>(m + i)[k]; }
};
}  // namespace deme_jit
