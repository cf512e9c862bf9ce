// Host-side domain sizing for the DEM hot path (no device work).
//
// figureOutNV     src/DEM/APIPrivate.cpp:222-333 (voxel-bit split and length unit l, stored as float, API.h:508)
// decideBinSize   src/DEM/APIPrivate.cpp:335-369 (bin = 2 * smallest radius unless instructed; nb = floor(world/bin)+1)
// kernel literals src/DEM/HostSideHelpers.hpp:397-402 + APIPrivate.cpp:1358-1359 (17-decimal fixed strings)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>

#include "../../include/deme/deme.h"

namespace {
constexpr int VOXEL_RES_POWER2 = 16;    // bits of subVoxelPos_t (Defines.h:40)
constexpr int VOXEL_COUNT_POWER2 = 64;  // bits of voxelID_t (Defines.h:41)
constexpr double TINY = 1e-12;          // DEME_TINY_FLOAT
}  // namespace

extern "C" {

double deme_kernel_literal(double v) {
    std::ostringstream out;
    out.precision(17);
    out << std::fixed << v;
    return std::strtod(out.str().c_str(), nullptr);
}

int deme_figure_out_nv(float bx, float by, float bz, int exact_dir, uint8_t nv_out[3], float* l_out) {
    if (!(bx > 0.f) || !(by > 0.f) || !(bz > 0.f)) return DEME_ERR_INVALID_ARGUMENT;
    float XYZ[3] = {bx, by, bz};
    int rank[3] = {0, 1, 2};
    // ascending order of the three box sizes, ties keep X < Y < Z
    for (int i = 0; i < 2; i++)
        for (int j = i + 1; j < 3; j++)
            if (XYZ[i] > XYZ[j]) {
                std::swap(XYZ[i], XYZ[j]);
                std::swap(rank[i], rank[j]);
            }
    const float size321[3] = {XYZ[0], XYZ[1], XYZ[2]};
    // one extra bit per factor of two between consecutive sizes, a second one past 2*sqrt(2)
    int more[2] = {0, 0};
    while (XYZ[0] < XYZ[1]) {
        if (std::sqrt(2.) * XYZ[0] > XYZ[1]) break;
        more[0]++;
        XYZ[0] *= 2.;
    }
    while (XYZ[1] < XYZ[2]) {
        if (std::sqrt(2.) * XYZ[1] > XYZ[2]) break;
        more[1]++;
        XYZ[1] *= 2.;
    }
    const int budget = VOXEL_COUNT_POWER2 - more[0] - more[1];
    int left = budget % 3;
    int bits[3];
    bits[0] = budget / 3;
    bits[1] = bits[0] + more[0];
    bits[2] = bits[1] + more[1];
    while (left > 0) {
        if (bits[0] < bits[1])
            bits[0]++;
        else if (bits[1] < bits[2])
            bits[1]++;
        else
            bits[2]++;
        left--;
    }
    const double res = std::pow(2., VOXEL_RES_POWER2);
    float l;
    if (exact_dir < 0) {
        double lmax = 0;
        for (int k = 0; k < 3; k++) lmax = std::max(lmax, (double)size321[k] / res / std::pow(2., bits[k]));
        l = (float)lmax;
    } else {
        int ex = 0;
        for (int k = 0; k < 3; k++)
            if (rank[k] == exact_dir) ex = k;
        int ne[2];
        if (ex == 0) {
            ne[0] = 1;
            ne[1] = 2;
        } else if (ex == 1) {
            ne[0] = 0;
            ne[1] = 2;
        } else {
            ne[0] = 0;
            ne[1] = 1;
        }
        l = (float)((double)size321[ex] / res / std::pow(2., bits[ex]));
        for (int pass = 1; pass >= 0; pass--) {
            const int d = ne[pass];
            while ((double)l * res * std::pow(2., bits[d]) < size321[d]) {
                bits[ex] -= 1;
                bits[d] += 1;
                l = (float)((double)size321[ex] / res / std::pow(2., bits[ex]));
            }
        }
    }
    for (int k = 0; k < 3; k++) nv_out[rank[k]] = (uint8_t)bits[k];
    *l_out = l;
    return DEME_OK;
}

int deme_build_sim_params(float bx, float by, float bz, int exact_dir, const uint8_t* nv_in, float l_in,
                          const float LBF[3], float min_radius, double bin_size, const float G[3], float h, float beta,
                          int integrator, deme_sim_params* out) {
    uint8_t nv[3];
    float l;
    if (nv_in) {
        if (nv_in[0] + nv_in[1] + nv_in[2] != VOXEL_COUNT_POWER2) return DEME_ERR_INVALID_ARGUMENT;
        nv[0] = nv_in[0];
        nv[1] = nv_in[1];
        nv[2] = nv_in[2];
        l = l_in;
    } else {
        int e = deme_figure_out_nv(bx, by, bz, exact_dir, nv, &l);
        if (e) return e;
        // figureOutNV hands 64 + (extra bits of the middle axis) bits out when the two smallest sides differ by more
        // than 2*sqrt(2); such ids would overflow voxelID_t, so the split is refused.
        if (nv[0] + nv[1] + nv[2] != VOXEL_COUNT_POWER2) return DEME_ERR_UNSUPPORTED;
    }
    deme_sim_params p{};
    p.nvXp2 = nv[0];
    p.nvYp2 = nv[1];
    p.nvZp2 = nv[2];
    p.integrator = (uint8_t)integrator;
    p.l_exact = (double)l;
    p.voxelSize_exact = (double)((size_t)1 << VOXEL_RES_POWER2) * (double)l;
    p.l = deme_kernel_literal(p.l_exact);
    p.voxelSize = deme_kernel_literal(p.voxelSize_exact);
    if (bin_size > 0.0) {
        p.binSize = bin_size;
    } else {
        if (!(min_radius > TINY)) return DEME_ERR_INVALID_ARGUMENT;  // zero radius and no bin size (APIPrivate.cpp:352)
        p.binSize = 2.0 * min_radius;
    }
    const double wx = p.voxelSize_exact * (double)((size_t)1 << nv[0]);
    const double wy = p.voxelSize_exact * (double)((size_t)1 << nv[1]);
    const double wz = p.voxelSize_exact * (double)((size_t)1 << nv[2]);
    const double nbx = wx / p.binSize + 1.0, nby = wy / p.binSize + 1.0, nbz = wz / p.binSize + 1.0;
    if (nbx >= 4294967295.0 || nby >= 4294967295.0 || nbz >= 4294967295.0) return DEME_ERR_CAPACITY;
    p.nbX = (uint32_t)(wx / p.binSize) + 1;
    p.nbY = (uint32_t)(wy / p.binSize) + 1;
    p.nbZ = (uint32_t)(wz / p.binSize) + 1;
    if ((uint64_t)p.nbX * p.nbY * p.nbZ > 0xFFFFFFFFull) return DEME_ERR_CAPACITY;  // APIPrivate.cpp:156-162
    for (int k = 0; k < 3; k++) {
        p.LBF[k] = LBF[k];
        p.G[k] = G[k];
    }
    p.h = h;
    p.beta = beta;
    p.boxSize[0] = bx;
    p.boxSize[1] = by;
    p.boxSize[2] = bz;
    *out = p;
    return DEME_OK;
}

int deme_encode_positions(const deme_sim_params* p, uint64_t n, const float* xyz, uint64_t* voxelID, uint16_t* locX,
                          uint16_t* locY, uint16_t* locZ) {
    auto u64 = [](double x) -> uint64_t {
        if (!(x > 0.0)) return 0;
        if (x >= 18446744073709551616.0) return ~0ull;
        return (uint64_t)x;
    };
    auto u16 = [](double x) -> uint16_t {
        if (!(x > 0.0)) return 0;
        if (x >= 65535.0) return 65535;
        return (uint16_t)(uint32_t)x;
    };
    const double vs = p->voxelSize_exact, l = p->l_exact;
    for (uint64_t i = 0; i < n; i++) {
        // float3 - float3 first (the reference subtracts LBF in float), then double
        const double X = (double)(xyz[3 * i] - p->LBF[0]);
        const double Y = (double)(xyz[3 * i + 1] - p->LBF[1]);
        const double Z = (double)(xyz[3 * i + 2] - p->LBF[2]);
        const uint64_t nx = u64(X / vs), ny = u64(Y / vs), nz = u64(Z / vs);
        locX[i] = u16((X - (double)nx * vs) / l);
        locY[i] = u16((Y - (double)ny * vs) / l);
        locZ[i] = u16((Z - (double)nz * vs) / l);
        uint64_t id = nx;
        id += ny << p->nvXp2;
        id += nz << (p->nvXp2 + p->nvYp2);
        voxelID[i] = id;
    }
    return DEME_OK;
}

int deme_decode_positions(const deme_sim_params* p, uint64_t n, const uint64_t* voxelID, const uint16_t* locX,
                          const uint16_t* locY, const uint16_t* locZ, float* xyz) {
    const double vs = p->voxelSize_exact, l = p->l_exact;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t id = voxelID[i];
        const uint64_t vx = id & (((uint64_t)1 << p->nvXp2) - 1);
        const uint64_t vy = (id >> p->nvXp2) & (((uint64_t)1 << p->nvYp2) - 1);
        const uint64_t vz = id >> (p->nvXp2 + p->nvYp2);
        const double X = (double)vx * vs + (double)locX[i] * l;
        const double Y = (double)vy * vs + (double)locY[i] * l;
        const double Z = (double)vz * vs + (double)locZ[i] * l;
        xyz[3 * i] = (float)(X + p->LBF[0]);
        xyz[3 * i + 1] = (float)(Y + p->LBF[1]);
        xyz[3 * i + 2] = (float)(Z + p->LBF[2]);
    }
    return DEME_OK;
}

}  // extern "C"
