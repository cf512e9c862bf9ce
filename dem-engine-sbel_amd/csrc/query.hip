// Inspectors: reductions of a per-sphere or per-clump quantity over the whole domain or a region (DEMInspector,
// AuxClasses.cpp:17-160; dT::inspectCall, dT.cpp:1660-1720). The kernels are generated at build time from
// jit/inspect_kernel.tpl with a region that counts everything; regions are compiled through hipRTC (jit_rt.cpp).
#include "deme_internal.h"
#include "inspect_aot.hip"

namespace deme_int {
using namespace deme_dev;

uint32_t inspect_blocks(uint32_t n) {
    const uint32_t per = (uint32_t)(INSPECT_BLOCK * INSPECT_ITEMS);
    const uint32_t nb = (uint32_t)(((uint64_t)n + per - 1) / per);
    return nb ? nb : 1;
}

hipError_t launch_inspect(const KParams& P, int quantity, int flavor, uint32_t n, const SphereRec* sph,
                          const OwnerRec* own, const float* vol, const float lbf[3], double* partial, hipStream_t st) {
    hipLaunchKernelGGL(deme_inspect_kernel, dim3(inspect_blocks(n)), dim3(INSPECT_BLOCK), 0, st, P, quantity, flavor,
                       n, sph, own, vol, lbf[0], lbf[1], lbf[2], partial);
    return hipGetLastError();
}

hipError_t launch_inspect_final(int flavor, uint32_t nPart, const double* partial, double* out, hipStream_t st) {
    hipLaunchKernelGGL(deme_inspect_final_kernel, dim3(1), dim3(INSPECT_BLOCK), 0, st, flavor, nPart, partial, out);
    return hipGetLastError();
}

// dTModifyComponents (DEMMiscKernels.cu:18-33): the component spheres of marked clumps scale their offsets and radii.
__global__ void k_scale_components(SphereRec* __restrict__ sph, uint32_t n, const uint8_t* __restrict__ mark,
                                   const float* __restrict__ factor) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = sph[i].owner;
    if (!mark[o]) return;
    const float f = factor[o];
    sph[i].rx = sph[i].rx * f;
    sph[i].ry = sph[i].ry * f;
    sph[i].rz = sph[i].rz * f;
    sph[i].r = sph[i].r * f;
}

hipError_t launch_scale_components(SphereRec* sph, uint32_t n, const uint8_t* mark, const float* factor,
                                   hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_components, dim3((n + 255) / 256), dim3(256), 0, st, sph, n, mark, factor);
    return hipGetLastError();
}

}  // namespace deme_int
