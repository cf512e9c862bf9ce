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

}  // namespace deme_int
