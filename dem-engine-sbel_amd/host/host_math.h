// Float helpers of the DEMSolver facade's host code, rounding as the reference's host helpers do.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace deme {
namespace hostmath {

// normalize on the host (thirdparty helper_math: v * (1 / sqrtf(dot(v, v))))
inline float3 normalize3(float3 v) {
    const float invLen = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return make_float3(v.x * invLen, v.y * invLen, v.z * invLen);
}

// hostApplyOriQToVector3<float, float> (HostSideHelpers.hpp:383-394): rotate (X, Y, Z) by the unit quaternion Q
inline void host_rotate(float& X, float& Y, float& Z, float Qw, float Qx, float Qy, float Qz) {
    const float oX = X, oY = Y, oZ = Z;
    X = (2.0f * (Qw * Qw + Qx * Qx) - 1.0f) * oX + (2.0f * (Qx * Qy - Qw * Qz)) * oY + (2.0f * (Qx * Qz + Qw * Qy)) * oZ;
    Y = (2.0f * (Qx * Qy + Qw * Qz)) * oX + (2.0f * (Qw * Qw + Qy * Qy) - 1.0f) * oY + (2.0f * (Qy * Qz - Qw * Qx)) * oZ;
    Z = (2.0f * (Qx * Qz - Qw * Qy)) * oX + (2.0f * (Qy * Qz + Qw * Qx)) * oY + (2.0f * (Qw * Qw + Qz * Qz) - 1.0f) * oZ;
}

}  // namespace hostmath
}  // namespace deme
