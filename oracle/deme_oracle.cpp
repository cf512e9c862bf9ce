// CPU oracle for the DEM per-step hot path -- TEST INFRASTRUCTURE ONLY.
//
// Host C++17 + OpenMP restatement of the reference kernels, used by tests/ as the parity checker and by bench.py as
// the reported CPU baseline. Never linked into, or called by, the product library.
//
// Build: -O2 -ffp-contract=off -fno-fast-math (x86-64 SSE2 arithmetic, no contraction), see oracle/Makefile.
// C++ overload resolution reproduces the reference's float/double promotions: sqrt/log on a float argument resolve to
// the float overloads, as they do in CUDA device code.
//
// Float->integer conversions follow the reference's device semantics: PTX cvt.rzi to an integer type saturates to
// the destination range (NaN -> 0). The sat_* helpers make that behaviour explicit on the host.
//
// Parity status of each function (pins live in tests/test_oracle_pins.py):
//   sizing (figureOutNV, binSize, literals)            pinned: closed forms for the workloads' boxes
//   voxel codec, rotation, quaternion product          pinned: round trips, identities, a 90-degree rotation
//   bin counting / filling / sorting                   pinned: brute-force enumeration of AABB bins
//   pair detection (bin rule)                          pinned: brute-force O(S^2) predicate; physical overlap set
//   canonical order + history map                      pinned: std::sort / map join on random lists
//   Hertz-Mindlin force                                pinned: closed-form Hertz law; normal and tangential damping,
//                                                      tangential spring, Coulomb cap with the back-solved history,
//                                                      rolling resistance and its collision-time switch, and the
//                                                      two-material mixing rule against their closed forms; Newton's
//                                                      3rd law (the FP64 restatement is held to the same closed forms)
//   owner reduction                                    pinned: zero net force without planes; FP64 sum
//   integration                                        pinned: free fall, constant spin, prescribed/fixed families
//   schedule                                           pinned: resting-on-plane equilibrium depth

#include "deme_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------------------------------------
// Constants (src/DEM/Defines.h:26-38)
// ---------------------------------------------------------------------------------------------------------------
constexpr double TINY_FLOAT = 1e-12;
constexpr double HUGE_FLOAT = 1e15;
constexpr double TWO_OVER_THREE = 0.666666666666667;
constexpr double FOUR_OVER_THREE = 1.333333333333333;
constexpr double FIVE_OVER_THREE = 1.666666666666667;
constexpr double TWO_TIMES_SQRT_FIVE_OVER_SIX = 1.825741858350554;
constexpr double PI = 3.1415926535897932385;
constexpr double PI_SQUARED = 9.869604401089358;

// Saturating conversions (device float->int semantics).
inline uint32_t sat_u32(double x) {
    if (!(x > 0.0)) return 0u;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
inline uint16_t sat_u16(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65535.0) return 65535;
    return (uint16_t)(uint32_t)x;
}
inline uint64_t sat_u64(double x) {
    if (!(x > 0.0)) return 0ull;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)x;
}

struct f3 {
    float x, y, z;
};
struct d3 {
    double x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
inline f3 operator*(float b, f3 a) { return mk(b * a.x, b * a.y, b * a.z); }
inline f3 operator/(f3 a, float b) { return mk(a.x / b, a.y / b, a.z / b); }
inline f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline void operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
inline void operator-=(f3& a, f3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
// CUDAMathHelpers.cu:47-57
inline f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(f3 v) { return std::sqrt(dot(v, v)); }
// CUDAMathHelpers.cu:73-75: double3 - double3 returns a float3
inline f3 sub_d(d3 a, d3 b) { return mk((float)(a.x - b.x), (float)(a.y - b.y), (float)(a.z - b.z)); }

// DEMHelperKernels.cu:137-148 applyOriQToVector3<float, float>
inline void rotq(float& X, float& Y, float& Z, float Qw, float Qx, float Qy, float Qz) {
    float oX = X, oY = Y, oZ = Z;
    X = (2.0f * (Qw * Qw + Qx * Qx) - 1.0f) * oX + (2.0f * (Qx * Qy - Qw * Qz)) * oY + (2.0f * (Qx * Qz + Qw * Qy)) * oZ;
    Y = (2.0f * (Qx * Qy + Qw * Qz)) * oX + (2.0f * (Qw * Qw + Qy * Qy) - 1.0f) * oY + (2.0f * (Qy * Qz - Qw * Qx)) * oZ;
    Z = (2.0f * (Qx * Qz - Qw * Qy)) * oX + (2.0f * (Qy * Qz + Qw * Qx)) * oY + (2.0f * (Qw * Qw + Qz * Qz) - 1.0f) * oZ;
}
inline void rotq(f3& v, float Qw, float Qx, float Qy, float Qz) { rotq(v.x, v.y, v.z, Qw, Qx, Qy, Qz); }

// DEMHelperKernels.cu:51-56
inline uint32_t locate_mask_pair(uint32_t i, uint32_t j) {
    if (i > j) return locate_mask_pair(j, i);
    return (1 + j) * j / 2 + i;
}
// DEMHelperKernels.cu:60-64 (float sqrt of a u32 converted to float, then double arithmetic, then truncation)
inline void recover_cnt_pair(uint32_t& i, uint32_t& j, uint32_t ind, uint32_t n) {
    i = n - 2 - (uint32_t)(std::sqrt((float)(4 * n * (n - 1) - 7 - 8 * ind)) / 2.0 - 0.5);
    j = ind + i + 1 + (n - i) * ((n - i) - 1) / 2 - n * (n - 1) / 2;
}

std::string fixed17(double v) {
    std::ostringstream out;
    out.precision(17);
    out << std::fixed << v;
    return out.str();
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------------------------------
struct ora_ctx {
    deme_sim_params p;
    // wall seconds per phase, under the reference's timer names (kT.h:179-181, dT.h:253-255): Discretize domain, Find
    // contact pairs, Build history map, Calculate contact forces, Collect contact forces, Integration
    double phase[6] = {0, 0, 0, 0, 0, 0};
    uint32_t nS = 0, nO = 0, nA = 0, nM = 0;
    // templates (flattened, ClumpCompAcqStratAllFlatten.cu / MassAcqStratFlatten.cu / MOIAcqStratFlatten.cu)
    std::vector<uint32_t> sphOwner;
    std::vector<float> sphR, sphRx, sphRy, sphRz;
    std::vector<uint8_t> sphMat;
    std::vector<float> mass, moiX, moiY, moiZ;
    std::vector<deme_anal_obj> anal;
    std::vector<deme_material> mats;
    std::vector<uint8_t> mask;
    std::vector<deme_family_rule> rules;
    std::vector<uint8_t> isAnalOwner;
    // owner state
    std::vector<uint64_t> voxel;
    std::vector<uint16_t> locX, locY, locZ;
    std::vector<float> qw, qx, qy, qz, vX, vY, vZ, wX, wY, wZ, aX, aY, aZ, alX, alY, alZ;
    std::vector<uint8_t> fam;
    std::vector<uint32_t> ownerScene;  // batch scene per owner (all 0 for one scene)
    std::vector<uint32_t> analScene;   // batch scene per analytical component (empty: its owner's)
    uint32_t nScenes = 1;
    // CD intermediates
    std::vector<uint16_t> numBins;
    std::vector<uint8_t> numAnal;
    std::vector<uint32_t> binScan, analScan;
    std::vector<uint32_t> pairBin, pairSph;  // sorted
    std::vector<uint32_t> activeBin, activeStart;
    std::vector<uint16_t> activeCnt;
    uint64_t nGeo = 0, maxInBin = 0;
    bool binsReady = false;
    // contacts (canonical order) and previous list
    std::vector<uint32_t> idA, idB, map;
    std::vector<uint8_t> ctype;
    std::vector<uint32_t> pidA, pidB;
    std::vector<uint8_t> pctype;
    bool havePrev = false;
    // per-contact outputs
    std::vector<float> w[4];  // delta_tan_x, delta_tan_y, delta_tan_z, delta_time (alphabetical, Models.h:241-251)
    std::vector<f3> F, Ftof, cpA, cpB;
    double time = 0.0;
    uint64_t nSteps = 0;
    bool contactsReady = false;
};

namespace {

inline void owner_pos(const ora_ctx* c, uint32_t o, double& X, double& Y, double& Z) {
    // voxelIDToPosition<double, voxelID_t, subVoxelPos_t> (DEMHelperKernels.cu:93-110) with the literal l/voxelSize
    const uint64_t id = c->voxel[o];
    const int nx = c->p.nvXp2, ny = c->p.nvYp2;
    uint64_t vx = id & (((uint64_t)1 << nx) - 1);
    uint64_t vy = (id >> nx) & (((uint64_t)1 << ny) - 1);
    uint64_t vz = id >> (nx + ny);
    X = (double)vx * c->p.voxelSize + (double)c->locX[o] * c->p.l;
    Y = (double)vy * c->p.voxelSize + (double)c->locY[o] * c->p.l;
    Z = (double)vz * c->p.voxelSize + (double)c->locZ[o] * c->p.l;
}

// positionToVoxelID<voxelID_t, subVoxelPos_t, double> (DEMHelperKernels.cu:113-135), saturating conversions
inline void pos_to_voxel(const ora_ctx* c, double X, double Y, double Z, uint64_t& id, uint16_t& sx, uint16_t& sy,
                         uint16_t& sz) {
    const double vs = c->p.voxelSize, l = c->p.l;
    uint64_t nx = sat_u64(X / vs), ny = sat_u64(Y / vs), nz = sat_u64(Z / vs);
    sx = sat_u16((X - (double)nx * vs) / l);
    sy = sat_u16((Y - (double)ny * vs) / l);
    sz = sat_u16((Z - (double)nz * vs) / l);
    id = nx;
    id += ny << c->p.nvXp2;
    id += nz << (c->p.nvXp2 + c->p.nvYp2);
}

// Sphere world position (relative to LBF) as every CD kernel computes it (DEMBinSphereKernels.cu:39-51,
// DEMContactKernels_Blockwise.cu:66-76): owner position in double plus the float-rotated offset.
inline void sphere_pos(const ora_ctx* c, uint32_t s, double& X, double& Y, double& Z) {
    const uint32_t o = c->sphOwner[s];
    double ox, oy, oz;
    owner_pos(c, o, ox, oy, oz);
    float rx = c->sphRx[s], ry = c->sphRy[s], rz = c->sphRz[s];
    rotq(rx, ry, rz, c->qw[o], c->qx[o], c->qy[o], c->qz[o]);
    X = ox + (double)rx;
    Y = oy + (double)ry;
    Z = oz + (double)rz;
}

struct BinRange {
    uint32_t x0, x1, y0, y1, z0, z1;
};

// DEMBinSphereKernels.cu:52-64 / 163-183: bins an expanded sphere's AABB touches (r' in double)
inline BinRange bin_range(const ora_ctx* c, double X, double Y, double Z, double rCD) {
    const double b = c->p.binSize;
    double bx = X / b, by = Y / b, bz = Z / b, span = rCD / b;
    BinRange r;
    r.x0 = sat_u32(bx - span);
    r.x1 = sat_u32(bx + span);
    r.y0 = sat_u32(by - span);
    r.y1 = sat_u32(by + span);
    r.z0 = sat_u32(bz - span);
    r.z1 = sat_u32(bz + span);
    return r;
}

// getPointBinID<binID_t> (DEMHelperKernels.cu:266-273)
inline uint32_t point_bin(const ora_ctx* c, double X, double Y, double Z) {
    uint32_t ix = sat_u32(X / c->p.binSize), iy = sat_u32(Y / c->p.binSize), iz = sat_u32(Z / c->p.binSize);
    return ix + iy * c->p.nbX + iz * c->p.nbX * c->p.nbY;
}

// Analytical component world pose: owner position + owner-rotated offset; owner-rotated normal (float).
inline void anal_pose(const ora_ctx* c, uint32_t a, double& X, double& Y, double& Z, float& nx, float& ny, float& nz) {
    const deme_anal_obj& ob = c->anal[a];
    const uint32_t o = ob.owner;
    double ox, oy, oz;
    owner_pos(c, o, ox, oy, oz);
    float px = ob.relPos[0], py = ob.relPos[1], pz = ob.relPos[2];
    nx = ob.rot[0];
    ny = ob.rot[1];
    nz = ob.rot[2];
    rotq(px, py, pz, c->qw[o], c->qx[o], c->qy[o], c->qz[o]);
    rotq(nx, ny, nz, c->qw[o], c->qx[o], c->qy[o], c->qz[o]);
    X = ox + (double)px;
    Y = oy + (double)py;
    Z = oz + (double)pz;
}

// checkSphereEntityOverlap<double> without contact point (DEMHelperKernels.cu:356-391)
inline bool sphere_anal_cd(const ora_ctx* c, uint32_t a, double X, double Y, double Z, double rCD) {
    const deme_anal_obj& ob = c->anal[a];
    if (ob.type != DEME_ANAL_OBJ_TYPE_PLANE) return false;
    double ax, ay, az;
    float nx, ny, nz;
    anal_pose(c, a, ax, ay, az, nx, ny, nz);
    const double dx = X - ax, dy = Y - ay, dz = Z - az;
    const double dist = dx * (double)nx + dy * (double)ny + dz * (double)nz;
    if (dist > rCD + (double)c->p.beta) return false;
    return true;
}

inline bool mask_allows(const ora_ctx* c, uint32_t famA, uint32_t famB) {
    return c->mask[locate_mask_pair(famA, famB)] == DEME_DONT_PREVENT_CONTACT;
}

// A batch keeps scenes apart: a sphere meets only the analytical components of its own scene.
inline bool same_scene(const ora_ctx* c, uint32_t sphereOwner, uint32_t a) {
    if (c->ownerScene.empty()) return true;
    const uint32_t sa = c->analScene.empty() ? c->ownerScene[c->anal[a].owner] : c->analScene[a];
    return c->ownerScene[sphereOwner] == sa;
}

// checkSpheresOverlap<double> (DEMHelperKernels.cu:195-222) -- CD flavour (float normal, half depth)
inline bool cd_overlap(double XA, double YA, double ZA, double radA, double XB, double YB, double ZB, double radB,
                       double& CPX, double& CPY, double& CPZ) {
    double d2 = (XA - XB) * (XA - XB) + (YA - YB) * (YA - YB) + (ZA - ZB) * (ZA - ZB);
    if (d2 > (radA + radB) * (radA + radB)) return false;
    float nX = (float)(XA - XB), nY = (float)(YA - YB), nZ = (float)(ZA - ZB);
    float mag = std::sqrt(nX * nX + nY * nY + nZ * nZ);
    nX /= mag;
    nY /= mag;
    nZ /= mag;
    double half = (radA + radB - std::sqrt(d2)) / 2.0;
    CPX = XB + (radB - half) * (double)nX;
    CPY = YB + (radB - half) * (double)nY;
    CPZ = ZB + (radB - half) * (double)nZ;
    return true;
}

}  // namespace

extern "C" {

void ora_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// figureOutNV (src/DEM/APIPrivate.cpp:222-333), m_box_dir_length_is_exact == NONE branch.
void ora_figure_out_nv(float bx, float by, float bz, int* nvX, int* nvY, int* nvZ, float* lout) {
    float XYZ[3] = {bx, by, bz};
    int rank[3] = {0, 1, 2};
    for (int i = 0; i < 2; i++)
        for (int j = i + 1; j < 3; j++)
            if (XYZ[i] > XYZ[j]) {
                std::swap(XYZ[i], XYZ[j]);
                std::swap(rank[i], rank[j]);
            }
    float size321[3] = {XYZ[0], XYZ[1], XYZ[2]};
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
    int base = (64 - more[0] - more[1]) / 3;
    int left = (64 - more[0] - more[1]) % 3;
    int b3 = base, b2 = b3 + more[0], b1 = b2 + more[1];
    while (left > 0) {
        if (b3 < b2)
            b3++;
        else if (b2 < b1)
            b2++;
        else
            b1++;
        left--;
    }
    int bits[3] = {b3, b2, b1};
    double l3 = (double)size321[0] / std::pow(2., 16) / std::pow(2., b3);
    double l2 = (double)size321[1] / std::pow(2., 16) / std::pow(2., b2);
    double l1 = (double)size321[2] / std::pow(2., 16) / std::pow(2., b1);
    float l = (float)std::max(l3, std::max(l2, l1));
    int out[3];
    for (int i = 0; i < 3; i++) out[rank[i]] = bits[i];
    *nvX = out[0];
    *nvY = out[1];
    *nvZ = out[2];
    *lout = l;
}

double ora_literal17(double x) { return std::strtod(fixed17(x).c_str(), nullptr); }

uint32_t ora_locate_mask_pair(uint32_t i, uint32_t j) { return locate_mask_pair(i, j); }

void ora_recover_cnt_pair(uint32_t ind, uint32_t n, uint32_t* i, uint32_t* j) {
    uint32_t a, b;
    recover_cnt_pair(a, b, ind, n);
    *i = a;
    *j = b;
}

ora_ctx* ora_create(const deme_sim_params* p, uint32_t nSpheres, const uint32_t* sphOwner, const float* sphRadius,
                    const float* sphRelPos, const uint8_t* sphMat, uint32_t nOwners, const float* ownerMass,
                    const float* ownerMoi, uint32_t nAnal, const deme_anal_obj* anal, uint32_t nMat,
                    const deme_material* mats, const uint8_t* familyMask, const deme_family_rule* rules) {
    ora_ctx* c = new ora_ctx();
    c->p = *p;
    c->nS = nSpheres;
    c->nO = nOwners;
    c->nA = nAnal;
    c->nM = nMat;
    c->sphOwner.assign(sphOwner, sphOwner + nSpheres);
    c->sphR.assign(sphRadius, sphRadius + nSpheres);
    c->sphRx.resize(nSpheres);
    c->sphRy.resize(nSpheres);
    c->sphRz.resize(nSpheres);
    for (uint32_t s = 0; s < nSpheres; s++) {
        c->sphRx[s] = sphRelPos[3 * s];
        c->sphRy[s] = sphRelPos[3 * s + 1];
        c->sphRz[s] = sphRelPos[3 * s + 2];
    }
    c->sphMat.assign(sphMat, sphMat + nSpheres);
    c->mass.assign(ownerMass, ownerMass + nOwners);
    c->moiX.resize(nOwners);
    c->moiY.resize(nOwners);
    c->moiZ.resize(nOwners);
    for (uint32_t o = 0; o < nOwners; o++) {
        c->moiX[o] = ownerMoi[3 * o];
        c->moiY[o] = ownerMoi[3 * o + 1];
        c->moiZ[o] = ownerMoi[3 * o + 2];
    }
    c->anal.assign(anal, anal + nAnal);
    c->mats.assign(mats, mats + nMat);
    c->mask.assign(familyMask, familyMask + DEME_NUM_MASK_ENTRIES);
    c->rules.assign(rules, rules + DEME_NUM_FAMILIES);
    c->isAnalOwner.assign(nOwners, 0);
    for (uint32_t a = 0; a < nAnal; a++) c->isAnalOwner[anal[a].owner] = 1;
    auto z = [&](auto& v) { v.assign(nOwners, 0); };
    z(c->voxel); z(c->locX); z(c->locY); z(c->locZ); z(c->qw); z(c->qx); z(c->qy); z(c->qz);
    z(c->vX); z(c->vY); z(c->vZ); z(c->wX); z(c->wY); z(c->wZ); z(c->aX); z(c->aY); z(c->aZ);
    z(c->alX); z(c->alY); z(c->alZ); z(c->fam);
    return c;
}

void ora_destroy(ora_ctx* c) { delete c; }

void ora_set_state(ora_ctx* c, const deme_owner_state* s) {
    const uint32_t n = c->nO;
    auto cp = [&](auto& dst, const auto* src) {
        if (src) std::copy(src, src + n, dst.begin());
    };
    cp(c->voxel, s->voxelID); cp(c->locX, s->locX); cp(c->locY, s->locY); cp(c->locZ, s->locZ);
    cp(c->qw, s->oriQw); cp(c->qx, s->oriQx); cp(c->qy, s->oriQy); cp(c->qz, s->oriQz);
    cp(c->vX, s->vX); cp(c->vY, s->vY); cp(c->vZ, s->vZ);
    cp(c->wX, s->omgX); cp(c->wY, s->omgY); cp(c->wZ, s->omgZ);
    cp(c->aX, s->aX); cp(c->aY, s->aY); cp(c->aZ, s->aZ);
    cp(c->alX, s->alphaX); cp(c->alY, s->alphaY); cp(c->alZ, s->alphaZ);
    cp(c->fam, s->family);
}

void ora_get_state(const ora_ctx* c, deme_owner_state* s) {
    auto cp = [&](const auto& src, auto* dst) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    cp(c->voxel, s->voxelID); cp(c->locX, s->locX); cp(c->locY, s->locY); cp(c->locZ, s->locZ);
    cp(c->qw, s->oriQw); cp(c->qx, s->oriQx); cp(c->qy, s->oriQy); cp(c->qz, s->oriQz);
    cp(c->vX, s->vX); cp(c->vY, s->vY); cp(c->vZ, s->vZ);
    cp(c->wX, s->omgX); cp(c->wY, s->omgY); cp(c->wZ, s->omgZ);
    cp(c->aX, s->aX); cp(c->aY, s->aY); cp(c->aZ, s->aZ);
    cp(c->alX, s->alphaX); cp(c->alY, s->alphaY); cp(c->alZ, s->alphaZ);
    cp(c->fam, s->family);
}

void ora_get_counts(const ora_ctx* c, deme_counts* out) {
    std::memset(out, 0, sizeof(*out));
    out->nSpheres = c->nS;
    out->nOwners = c->nO;
    out->nAnal = c->nA;
    out->nBinSphereTouches = c->pairBin.size();
    out->nActiveBins = c->activeBin.size();
    out->nContacts = c->idA.size();
    out->nPrevContacts = c->pidA.size();
    out->nSphereGeoContacts = c->nGeo;
    out->maxSpheresInBin = c->maxInBin;
    out->nSteps = c->nSteps;
    out->time = c->time;
}

// ---------------------------------------------------------------------------------------------------------------
// Rows a6-a8: getNumberOfBinsEachSphereTouches, prefix scans, populateBinSphereTouchingPairs, SortPairs, RLE
// (DEMBinSphereKernels.cu:12-230, DEMCubContactDetection.cu:59-166)
// ---------------------------------------------------------------------------------------------------------------
static double wall_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int ora_cd_bins_body(ora_ctx* c);

int ora_cd_bins(ora_ctx* c) {
    const double t0 = wall_now();
    const int e = ora_cd_bins_body(c);
    c->phase[0] += wall_now() - t0;
    return e;
}

static int ora_cd_bins_body(ora_ctx* c) {
    const uint32_t nS = c->nS;
    c->numBins.assign(nS, 0);
    c->numAnal.assign(nS, 0);
    std::vector<uint8_t> overflow(1, 0);
    int err = 0;
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)nS; si++) {
        const uint32_t s = (uint32_t)si;
        double X, Y, Z;
        sphere_pos(c, s, X, Y, Z);
        double rCD = (double)c->sphR[s];
        rCD += (double)c->p.beta;
        BinRange r = bin_range(c, X, Y, Z, rCD);
        uint16_t nx = (uint16_t)(r.x1 - r.x0 + 1), ny = (uint16_t)(r.y1 - r.y0 + 1), nz = (uint16_t)(r.z1 - r.z0 + 1);
        uint64_t full = (uint64_t)(r.x1 - r.x0 + 1) * (r.y1 - r.y0 + 1) * (r.z1 - r.z0 + 1);
        if (full > 65535u) {
#pragma omp atomic write
            err = DEME_ERR_CAPACITY;
        }
        c->numBins[s] = (uint16_t)((int)nx * (int)ny * (int)nz);
        const uint32_t famS = c->fam[c->sphOwner[s]];
        uint32_t cnt = 0;
        for (uint32_t a = 0; a < c->nA; a++) {
            if (!same_scene(c, c->sphOwner[s], a)) continue;
            if (!mask_allows(c, famS, c->fam[c->anal[a].owner])) continue;
            if (sphere_anal_cd(c, a, X, Y, Z, rCD)) cnt++;
        }
        c->numAnal[s] = (uint8_t)cnt;
    }
    if (err) return err;
    // exclusive scans (DEMCubContactDetection.cu:77-93)
    c->binScan.assign(nS, 0);
    c->analScan.assign(nS, 0);
    uint64_t P = 0, G = 0;
    for (uint32_t s = 0; s < nS; s++) {
        c->binScan[s] = (uint32_t)P;
        c->analScan[s] = (uint32_t)G;
        P += c->numBins[s];
        G += c->numAnal[s];
    }
    if (P > 0xFFFFFFFFull || G > 0xFFFFFFFFull) return DEME_ERR_CAPACITY;
    c->nGeo = G;
    // fill (k, j, i order) -- DEMBinSphereKernels.cu:168-183
    std::vector<uint32_t> binIDs(P), sphIDs(P);
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)nS; si++) {
        const uint32_t s = (uint32_t)si;
        double X, Y, Z;
        sphere_pos(c, s, X, Y, Z);
        double rCD = (double)c->sphR[s];
        rCD += (double)c->p.beta;
        BinRange r = bin_range(c, X, Y, Z, rCD);
        uint32_t off = c->binScan[s];
        const uint32_t nx = r.x1 - r.x0 + 1, ny = r.y1 - r.y0 + 1, nz = r.z1 - r.z0 + 1;
        for (uint32_t dk = 0; dk < nz; dk++)
            for (uint32_t dj = 0; dj < ny; dj++)
                for (uint32_t di = 0; di < nx; di++) {
                    binIDs[off] = (r.x0 + di) + (r.y0 + dj) * c->p.nbX + (r.z0 + dk) * c->p.nbX * c->p.nbY;
                    sphIDs[off] = s;
                    off++;
                }
    }
    // a bin id past the grid means a sphere left the simulation world (reported, as the device does)
    const uint64_t nBins = (uint64_t)c->p.nbX * c->p.nbY * c->p.nbZ;
    for (uint64_t e = 0; e < P; e++)
        if (binIDs[e] >= nBins) return DEME_ERR_CAPACITY;
    // stable sort by binID (DEMCubContactDetection.cu:127-129): counting sort over the key range is stable
    uint32_t maxKey = 0;
    for (uint64_t e = 0; e < P; e++) maxKey = std::max(maxKey, binIDs[e]);
    std::vector<uint32_t> start((size_t)maxKey + 2, 0);
    for (uint64_t e = 0; e < P; e++) start[(size_t)binIDs[e] + 1]++;
    for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
    c->pairBin.assign(P, 0);
    c->pairSph.assign(P, 0);
    {
        std::vector<uint32_t> cur(start.begin(), start.end() - 1);
        for (uint64_t e = 0; e < P; e++) {
            uint32_t pos = cur[binIDs[e]]++;
            c->pairBin[pos] = binIDs[e];
            c->pairSph[pos] = sphIDs[e];
        }
    }
    // unique / run-length encode / scan (DEMCubContactDetection.cu:135-166)
    c->activeBin.clear();
    c->activeCnt.clear();
    c->activeStart.clear();
    c->maxInBin = 0;
    for (uint64_t e = 0; e < P;) {
        uint64_t f = e;
        while (f < P && c->pairBin[f] == c->pairBin[e]) f++;
        c->activeBin.push_back(c->pairBin[e]);
        c->activeCnt.push_back((uint16_t)(f - e));
        c->activeStart.push_back((uint32_t)e);
        c->maxInBin = std::max<uint64_t>(c->maxInBin, f - e);
        e = f;
    }
    c->binsReady = true;
    if (c->maxInBin > DEME_MAX_SPHERES_PER_BIN) return DEME_ERR_BIN_OVERFLOW;
    return DEME_OK;
}

void ora_get_bin_counts(const ora_ctx* c, uint16_t* nb, uint8_t* na) {
    if (nb) std::copy(c->numBins.begin(), c->numBins.end(), nb);
    if (na) std::copy(c->numAnal.begin(), c->numAnal.end(), na);
}
void ora_get_bin_pairs(const ora_ctx* c, uint32_t* b, uint32_t* s) {
    if (b) std::copy(c->pairBin.begin(), c->pairBin.end(), b);
    if (s) std::copy(c->pairSph.begin(), c->pairSph.end(), s);
}
void ora_get_active_bins(const ora_ctx* c, uint32_t* b, uint16_t* n) {
    if (b) std::copy(c->activeBin.begin(), c->activeBin.end(), b);
    if (n) std::copy(c->activeCnt.begin(), c->activeCnt.end(), n);
}

// ---------------------------------------------------------------------------------------------------------------
// Row a9: per-bin pair detection (DEMContactKernels_Blockwise.cu:15-309), a10 sphere-plane contacts
// (DEMBinSphereKernels.cu:186-228), a11 canonical sort (DEMCubContactDetection.cu:236-258), a12 history map
// (DEMHistoryMappingKernels.cu:16-60).
//
// Canonical order: by idA; for equal idA, sphere-plane contacts first (ascending plane index) and then sphere-sphere
// contacts by ascending idB. This is the reference's stable-sort-by-idA order with its atomic within-bin order
// (DEMContactKernels_Blockwise.cu:258) replaced by ascending idB.
// ---------------------------------------------------------------------------------------------------------------
namespace {

void remap_wildcards(ora_ctx* c) {
    // rearrangeContactWildcards (DEMPrepForceKernels.cu:33-55)
    const size_t n = c->idA.size();
    for (int k = 0; k < 4; k++) {
        std::vector<float> nw(n, 0.0f);
        for (size_t i = 0; i < n; i++) {
            uint32_t m = c->map[i];
            if (m != DEME_NULL_MAPPING_PARTNER) nw[i] = c->w[k][m];
        }
        c->w[k].swap(nw);
    }
    c->F.assign(n, f3{0, 0, 0});
    c->Ftof.assign(n, f3{0, 0, 0});
    c->cpA.assign(n, f3{0, 0, 0});
    c->cpB.assign(n, f3{0, 0, 0});
}

void build_history_map(ora_ctx* c) {
    // buildPersistentMap (DEMHistoryMappingKernels.cu:16-60): same (A, B, type) in the previous list, else NULL.
    const size_t n = c->idA.size();
    c->map.assign(n, DEME_NULL_MAPPING_PARTNER);
    if (!c->havePrev) return;
    const size_t np = c->pidA.size();
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const size_t i = (size_t)ii;
        const uint32_t A = c->idA[i];
        size_t lo = std::lower_bound(c->pidA.begin(), c->pidA.end(), A) - c->pidA.begin();
        for (size_t j = lo; j < np && c->pidA[j] == A; j++) {
            if (c->pidB[j] == c->idB[i] && c->pctype[j] == c->ctype[i]) {
                c->map[i] = (uint32_t)j;
                break;
            }
        }
    }
}

void canonicalize_and_map(ora_ctx* c, std::vector<uint32_t>& A, std::vector<uint32_t>& B, std::vector<uint8_t>& T) {
    const size_t n = A.size();
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
        if (A[x] != A[y]) return A[x] < A[y];
        int gx = T[x] == DEME_SPHERE_SPHERE_CONTACT ? 1 : 0, gy = T[y] == DEME_SPHERE_SPHERE_CONTACT ? 1 : 0;
        if (gx != gy) return gx < gy;
        return B[x] < B[y];
    });
    // keep the previous list for the history map (kT copies new -> previous, DEMCubContactDetection.cu:358-370)
    if (c->contactsReady) {
        c->pidA.swap(c->idA);
        c->pidB.swap(c->idB);
        c->pctype.swap(c->ctype);
        c->havePrev = true;
    }
    c->idA.resize(n);
    c->idB.resize(n);
    c->ctype.resize(n);
    for (size_t k = 0; k < n; k++) {
        c->idA[k] = A[idx[k]];
        c->idB[k] = B[idx[k]];
        c->ctype[k] = T[idx[k]];
    }
    build_history_map(c);
    c->contactsReady = true;
}

}  // namespace

int ora_cd_pairs(ora_ctx* c) {
    const double t0 = wall_now();
    if (!c->binsReady) return DEME_ERR_NOT_READY;
    if (c->maxInBin > DEME_MAX_SPHERES_PER_BIN) return DEME_ERR_BIN_OVERFLOW;
    const size_t nB = c->activeBin.size();
    std::vector<std::vector<uint32_t>> outA(nB), outB(nB);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t bi = 0; bi < (int64_t)nB; bi++) {
        const uint32_t n = c->activeCnt[bi];
        if (n <= 1) continue;
        const uint32_t binID = c->activeBin[bi];
        const uint32_t st = c->activeStart[bi];
        // stage (DEMContactKernels_Blockwise.cu:178-209): float radius + beta, double position
        std::vector<double> bx(n), by(n), bz(n);
        std::vector<float> rr(n);
        std::vector<uint32_t> own(n), ids(n);
        std::vector<uint32_t> famv(n);
        for (uint32_t t = 0; t < n; t++) {
            uint32_t s = c->pairSph[st + t];
            ids[t] = s;
            own[t] = c->sphOwner[s];
            famv[t] = c->fam[own[t]];
            float r = c->sphR[s];
            r += c->p.beta;
            rr[t] = r;
            sphere_pos(c, s, bx[t], by[t], bz[t]);
        }
        const uint32_t nPairs = n * (n - 1) / 2;
        for (uint32_t ind = 0; ind < nPairs; ind++) {
            uint32_t a, b;
            recover_cnt_pair(a, b, ind, n);
            if (own[a] == own[b]) continue;
            if (!mask_allows(c, famv[a], famv[b])) continue;
            double cx, cy, cz;
            bool in = cd_overlap(bx[a], by[a], bz[a], (double)rr[a], bx[b], by[b], bz[b], (double)rr[b], cx, cy, cz);
            if (!in) continue;
            if (point_bin(c, cx, cy, cz) == binID) {
                outA[bi].push_back(ids[a]);
                outB[bi].push_back(ids[b]);
            }
        }
    }
    // geo contacts first (DEMBinSphereKernels.cu:186-228 at numAnalGeoSphereTouchesScan offsets), then bin pairs
    std::vector<uint32_t> A, B;
    std::vector<uint8_t> T;
    A.reserve(c->nGeo);
    for (uint32_t s = 0; s < c->nS; s++) {
        if (!c->numAnal[s]) continue;
        double X, Y, Z;
        sphere_pos(c, s, X, Y, Z);
        double rCD = (double)c->sphR[s];
        rCD += (double)c->p.beta;
        const uint32_t famS = c->fam[c->sphOwner[s]];
        for (uint32_t a = 0; a < c->nA; a++) {
            if (!same_scene(c, c->sphOwner[s], a)) continue;
            if (!mask_allows(c, famS, c->fam[c->anal[a].owner])) continue;
            if (sphere_anal_cd(c, a, X, Y, Z, rCD)) {
                A.push_back(s);
                B.push_back(a);
                T.push_back(DEME_SPHERE_PLANE_CONTACT);
            }
        }
    }
    for (size_t bi = 0; bi < nB; bi++) {
        for (size_t k = 0; k < outA[bi].size(); k++) {
            A.push_back(outA[bi][k]);
            B.push_back(outB[bi][k]);
            T.push_back(DEME_SPHERE_SPHERE_CONTACT);
        }
    }
    if (A.size() > 0xFFFFFFFFull) return DEME_ERR_CAPACITY;
    const double t1 = wall_now();
    c->phase[1] += t1 - t0;
    canonicalize_and_map(c, A, B, T);
    c->phase[2] += wall_now() - t1;
    return DEME_OK;
}

int ora_cd(ora_ctx* c) {
    int e = ora_cd_bins(c);
    if (e) return e;
    e = ora_cd_pairs(c);
    if (e) return e;
    const double t0 = wall_now();
    remap_wildcards(c);
    c->phase[2] += wall_now() - t0;
    return DEME_OK;
}

int ora_set_contacts(ora_ctx* c, uint64_t n, const uint32_t* idA, const uint32_t* idB, const uint8_t* type) {
    std::vector<uint32_t> A(idA, idA + n), B(idB, idB + n);
    std::vector<uint8_t> T(type, type + n);
    canonicalize_and_map(c, A, B, T);
    remap_wildcards(c);
    return DEME_OK;
}

void ora_get_contacts(const ora_ctx* c, uint32_t* idA, uint32_t* idB, uint8_t* type, uint32_t* map) {
    if (idA) std::copy(c->idA.begin(), c->idA.end(), idA);
    if (idB) std::copy(c->idB.begin(), c->idB.end(), idB);
    if (type) std::copy(c->ctype.begin(), c->ctype.end(), type);
    if (map) std::copy(c->map.begin(), c->map.end(), map);
}

// Brute force: a sphere pair is a contact iff the expanded spheres overlap, and the bin holding the contact point is
// touched by both spheres' AABBs (the only bin whose per-bin test can accept it), owners differ and families allow.
uint64_t ora_brute_force_pairs(ora_ctx* c, uint32_t* outA, uint32_t* outB, uint64_t cap) {
    const uint32_t nS = c->nS;
    std::vector<double> X(nS), Y(nS), Z(nS);
    std::vector<BinRange> R(nS);
    for (uint32_t s = 0; s < nS; s++) {
        sphere_pos(c, s, X[s], Y[s], Z[s]);
        double rCD = (double)c->sphR[s] + (double)c->p.beta;
        R[s] = bin_range(c, X[s], Y[s], Z[s], rCD);
    }
    uint64_t n = 0;
    for (uint32_t a = 0; a < nS; a++) {
        for (uint32_t b = a + 1; b < nS; b++) {
            if (c->sphOwner[a] == c->sphOwner[b]) continue;
            if (!mask_allows(c, c->fam[c->sphOwner[a]], c->fam[c->sphOwner[b]])) continue;
            float ra = c->sphR[a];
            ra += c->p.beta;
            float rb = c->sphR[b];
            rb += c->p.beta;
            double cx, cy, cz;
            if (!cd_overlap(X[a], Y[a], Z[a], (double)ra, X[b], Y[b], Z[b], (double)rb, cx, cy, cz)) continue;
            uint32_t ix = sat_u32(cx / c->p.binSize), iy = sat_u32(cy / c->p.binSize), iz = sat_u32(cz / c->p.binSize);
            auto inside = [&](const BinRange& r) {
                return ix >= r.x0 && ix <= r.x1 && iy >= r.y0 && iy <= r.y1 && iz >= r.z0 && iz <= r.z1;
            };
            if (inside(R[a]) && inside(R[b])) {
                if (n < cap) {
                    outA[n] = a;
                    outB[n] = b;
                }
                n++;
            }
        }
    }
    return n;
}

// ---------------------------------------------------------------------------------------------------------------
// Rows a14-a15: prepareForceArrays + calculateContactForces with the frictional Hertz-Mindlin model
// (DEMPrepForceKernels.cu:23-31, DEMCalcForceKernels.cu:14-179, FullHertzianForceModel.cu:1-124,
//  DEMHelperKernels.cu:233-264, 316-339, 394-445)
// ---------------------------------------------------------------------------------------------------------------
namespace {

struct BodyGeom {
    d3 ownerPos, bodyPos;
    float mass, radius;
    float qw, qx, qy, qz;
    uint32_t owner;
    uint8_t mat;
    f3 linVel, rotVel;
};

inline void sphere_body(const ora_ctx* c, uint32_t s, BodyGeom& g) {
    const uint32_t o = c->sphOwner[s];
    g.owner = o;
    float rx = c->sphRx[s], ry = c->sphRy[s], rz = c->sphRz[s];
    g.radius = c->sphR[s];
    g.mass = c->mass[o];
    owner_pos(c, o, g.ownerPos.x, g.ownerPos.y, g.ownerPos.z);
    g.qw = c->qw[o];
    g.qx = c->qx[o];
    g.qy = c->qy[o];
    g.qz = c->qz[o];
    rotq(rx, ry, rz, g.qw, g.qx, g.qy, g.qz);
    g.bodyPos = d3{g.ownerPos.x + (double)rx, g.ownerPos.y + (double)ry, g.ownerPos.z + (double)rz};
    g.mat = c->sphMat[s];
    g.linVel = mk(c->vX[o], c->vY[o], c->vZ[o]);
    g.rotVel = mk(c->wX[o], c->wY[o], c->wZ[o]);
}

// matProxy2ContactParam<float> (DEMHelperKernels.cu:316-339): double intermediates, float results
inline void mat_proxy(const deme_material& A, const deme_material& B, float& E, float& G, float& CoR, float& mu,
                      float& Crr) {
    float invE = (1. - A.nu * A.nu) / A.E + (1. - B.nu * B.nu) / B.E;
    E = 1. / invE;
    float invG = 2. * (2. - A.nu) * (1. + A.nu) / A.E + 2. * (2. - B.nu) * (1. + B.nu) / B.E;
    G = 1. / invG;
    CoR = std::min(A.CoR, B.CoR);
    mu = std::max(A.mu, B.mu);
    Crr = std::max(A.Crr, B.Crr);
}

}  // namespace

int ora_force(ora_ctx* c) {
    const size_t n = c->idA.size();
    if (!c->contactsReady || c->w[0].size() != n) return DEME_ERR_NOT_READY;
    // prepareForceArrays: zero a/alpha (DEMPrepForceKernels.cu:16-24)
    std::fill(c->aX.begin(), c->aX.end(), 0.f);
    std::fill(c->aY.begin(), c->aY.end(), 0.f);
    std::fill(c->aZ.begin(), c->aZ.end(), 0.f);
    std::fill(c->alX.begin(), c->alX.end(), 0.f);
    std::fill(c->alY.begin(), c->alY.end(), 0.f);
    std::fill(c->alZ.begin(), c->alZ.end(), 0.f);
    const float ts = c->p.h;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const size_t i = (size_t)ii;
        uint8_t type = c->ctype[i];
        BodyGeom A, B;
        sphere_body(c, c->idA[i], A);
        d3 cp{0, 0, 0};
        f3 B2A{0, 0, 0};
        double overlapDepth = 0;
        if (type == DEME_SPHERE_SPHERE_CONTACT) {
            sphere_body(c, c->idB[i], B);
            // checkSpheresOverlap<double, float> (DEMHelperKernels.cu:233-264)
            const double radA = A.radius, radB = B.radius;
            const double XA = A.bodyPos.x, YA = A.bodyPos.y, ZA = A.bodyPos.z;
            const double XB = B.bodyPos.x, YB = B.bodyPos.y, ZB = B.bodyPos.z;
            double d2 = (XA - XB) * (XA - XB) + (YA - YB) * (YA - YB) + (ZA - ZB) * (ZA - ZB);
            if (d2 > (radA + radB) * (radA + radB)) {
                type = DEME_NOT_A_CONTACT;
            } else {
                B2A.x = (float)(XA - XB);
                B2A.y = (float)(YA - YB);
                B2A.z = (float)(ZA - ZB);
                float mag = std::sqrt(B2A.x * B2A.x + B2A.y * B2A.y + B2A.z * B2A.z);
                B2A.x /= mag;
                B2A.y /= mag;
                B2A.z /= mag;
                overlapDepth = radA + radB - std::sqrt(d2);
                cp.x = XB + (radB - overlapDepth / 2.0) * (double)B2A.x;
                cp.y = YB + (radB - overlapDepth / 2.0) * (double)B2A.y;
                cp.z = ZB + (radB - overlapDepth / 2.0) * (double)B2A.z;
            }
        } else {
            // analytical B (DEMCalcForceKernels.cu:112-151)
            const uint32_t ab = c->idB[i];
            const deme_anal_obj& ob = c->anal[ab];
            const uint32_t o = ob.owner;
            B.owner = o;
            B.mat = ob.material;
            B.mass = ob.mass;
            B.radius = (float)HUGE_FLOAT;
            owner_pos(c, o, B.ownerPos.x, B.ownerPos.y, B.ownerPos.z);
            float rx = ob.relPos[0], ry = ob.relPos[1], rz = ob.relPos[2];
            B.qw = c->qw[o];
            B.qx = c->qx[o];
            B.qy = c->qy[o];
            B.qz = c->qz[o];
            rotq(rx, ry, rz, B.qw, B.qx, B.qy, B.qz);
            B.bodyPos = d3{B.ownerPos.x + (double)rx, B.ownerPos.y + (double)ry, B.ownerPos.z + (double)rz};
            f3 rot = mk(ob.rot[0], ob.rot[1], ob.rot[2]);
            rotq(rot, B.qw, B.qx, B.qy, B.qz);
            B.linVel = mk(c->vX[o], c->vY[o], c->vZ[o]);
            B.rotVel = mk(c->wX[o], c->wY[o], c->wZ[o]);
            if (ob.type == DEME_ANAL_OBJ_TYPE_PLANE) {
                // checkSphereEntityOverlap<double, float> (DEMHelperKernels.cu:394-445), beta4Entity = 0
                const double radA = A.radius;
                const double dx = A.bodyPos.x - B.bodyPos.x, dy = A.bodyPos.y - B.bodyPos.y,
                             dz = A.bodyPos.z - B.bodyPos.z;
                const double dist = dx * (double)rot.x + dy * (double)rot.y + dz * (double)rot.z;
                overlapDepth = (radA + (double)0.0f - dist) / 2.0;
                if (overlapDepth < 0.0) {
                    type = DEME_NOT_A_CONTACT;
                } else {
                    cp.x = A.bodyPos.x - (double)rot.x * (dist + overlapDepth);
                    cp.y = A.bodyPos.y - (double)rot.y * (dist + overlapDepth);
                    cp.z = A.bodyPos.z - (double)rot.z * (dist + overlapDepth);
                    B2A = rot;
                }
            } else {
                type = DEME_NOT_A_CONTACT;
            }
        }
        f3 force{0, 0, 0}, tof{0, 0, 0};
        float delta_tan_x = c->w[0][i], delta_tan_y = c->w[1][i], delta_tan_z = c->w[2][i], delta_time = c->w[3][i];
        if (type != DEME_NOT_A_CONTACT) {
            f3 locCPA = sub_d(cp, A.ownerPos);
            f3 locCPB = sub_d(cp, B.ownerPos);
            rotq(locCPA, A.qw, -A.qx, -A.qy, -A.qz);
            rotq(locCPB, B.qw, -B.qx, -B.qy, -B.qz);
            // ---- FullHertzianForceModel.cu ----
            float E_cnt, G_cnt, CoR_cnt, mu_cnt, Crr_cnt;
            mat_proxy(c->mats[A.mat], c->mats[B.mat], E_cnt, G_cnt, CoR_cnt, mu_cnt, Crr_cnt);
            f3 rotVelCPA = cross(A.rotVel, locCPA);
            f3 rotVelCPB = cross(B.rotVel, locCPB);
            rotq(rotVelCPA, A.qw, A.qx, A.qy, A.qz);
            rotq(rotVelCPB, B.qw, B.qx, B.qy, B.qz);
            float mass_eff, sqrt_Rd, beta;
            f3 vrel_tan;
            f3 delta_tan = mk(delta_tan_x, delta_tan_y, delta_tan_z);
            {
                const f3 velB2A = (A.linVel + rotVelCPA) - (B.linVel + rotVelCPB);
                const float projection = dot(velB2A, B2A);
                vrel_tan = velB2A - projection * B2A;
                {
                    delta_tan += ts * vrel_tan;
                    const float disp_proj = dot(delta_tan, B2A);
                    delta_tan -= disp_proj * B2A;
                    delta_time += ts;
                }
                mass_eff = (A.mass * B.mass) / (A.mass + B.mass);
                sqrt_Rd = std::sqrt(overlapDepth * (A.radius * B.radius) / (A.radius + B.radius));
                const float Sn = 2. * E_cnt * sqrt_Rd;
                const float loge = (CoR_cnt < TINY_FLOAT) ? std::log(TINY_FLOAT) : std::log(CoR_cnt);
                beta = loge / std::sqrt(loge * loge + PI_SQUARED);
                const float k_n = TWO_OVER_THREE * Sn;
                const float gamma_n = TWO_TIMES_SQRT_FIVE_OVER_SIX * beta * std::sqrt(Sn * mass_eff);
                force += (float)(k_n * overlapDepth + gamma_n * projection) * B2A;
            }
            if (Crr_cnt > 0.0) {
                bool should_add_rolling_resistance = true;
                {
                    const float R_eff = std::sqrt((A.radius * B.radius) / (A.radius + B.radius));
                    const float kn_simple = FOUR_OVER_THREE * E_cnt * std::sqrt(R_eff);
                    const float gn_simple = -2.f * std::sqrt((float)(FIVE_OVER_THREE * mass_eff * E_cnt)) * beta *
                                            std::pow(R_eff, 0.25f);
                    const float d_coeff = gn_simple / (2.f * std::sqrt(kn_simple * mass_eff));
                    if (d_coeff < 1.0) {
                        float t_collision = PI * std::sqrt(mass_eff / (kn_simple * (1.f - d_coeff * d_coeff)));
                        if (delta_time <= t_collision) should_add_rolling_resistance = false;
                    }
                }
                if (should_add_rolling_resistance) {
                    const f3 v_rot = rotVelCPB - rotVelCPA;
                    const float v_rot_mag = length(v_rot);
                    if (v_rot_mag > TINY_FLOAT) tof = (v_rot / v_rot_mag) * (Crr_cnt * length(force));
                }
            }
            if (mu_cnt > 0.0) {
                const float kt = 8. * G_cnt * sqrt_Rd;
                const float gt = -TWO_TIMES_SQRT_FIVE_OVER_SIX * beta * std::sqrt(mass_eff * kt);
                f3 tangent_force = (-kt) * delta_tan - gt * vrel_tan;
                const float ft = length(tangent_force);
                if (ft > TINY_FLOAT) {
                    const float ft_max = length(force) * mu_cnt;
                    if (ft > ft_max) {
                        tangent_force = (ft_max / ft) * tangent_force;
                        delta_tan = (tangent_force + gt * vrel_tan) / (-kt);
                    }
                } else {
                    tangent_force = mk(0, 0, 0);
                }
                force += tangent_force;
            }
            delta_tan_x = delta_tan.x;
            delta_tan_y = delta_tan.y;
            delta_tan_z = delta_tan.z;
            c->cpA[i] = locCPA;
            c->cpB[i] = locCPB;
        } else {
            // _forceModelContactWildcardDestroy_ (Models.h:248-249); contact points are cleared to zero here (the
            // reference leaves stale values, which only ever multiply a zero force).
            delta_tan_x = delta_tan_y = delta_tan_z = delta_time = 0.f;
            c->cpA[i] = mk(0, 0, 0);
            c->cpB[i] = mk(0, 0, 0);
        }
        c->F[i] = force;
        c->Ftof[i] = tof;
        c->w[0][i] = delta_tan_x;
        c->w[1][i] = delta_tan_y;
        c->w[2][i] = delta_tan_z;
        c->w[3][i] = delta_time;
    }
    return DEME_OK;
}

void ora_get_contact_forces(const ora_ctx* c, float* F, float* T, float* cA, float* cB) {
    const size_t n = c->F.size();
    for (size_t i = 0; i < n; i++) {
        if (F) { F[3 * i] = c->F[i].x; F[3 * i + 1] = c->F[i].y; F[3 * i + 2] = c->F[i].z; }
        if (T) { T[3 * i] = c->Ftof[i].x; T[3 * i + 1] = c->Ftof[i].y; T[3 * i + 2] = c->Ftof[i].z; }
        if (cA) { cA[3 * i] = c->cpA[i].x; cA[3 * i + 1] = c->cpA[i].y; cA[3 * i + 2] = c->cpA[i].z; }
        if (cB) { cB[3 * i] = c->cpB[i].x; cB[3 * i + 1] = c->cpB[i].y; cB[3 * i + 2] = c->cpB[i].z; }
    }
}

// Per-contact records given from outside (the device's), so that the owner reduction can be compared on identical
// inputs; the list must be the oracle's current one (n = its contact count).
void ora_set_contact_forces(ora_ctx* c, const float* F, const float* T, const float* cA, const float* cB) {
    const size_t n = c->idA.size();
    c->F.resize(n);
    c->Ftof.resize(n);
    c->cpA.resize(n);
    c->cpB.resize(n);
    for (size_t i = 0; i < n; i++) {
        c->F[i] = {F[3 * i], F[3 * i + 1], F[3 * i + 2]};
        c->Ftof[i] = {T[3 * i], T[3 * i + 1], T[3 * i + 2]};
        c->cpA[i] = {cA[3 * i], cA[3 * i + 1], cA[3 * i + 2]};
        c->cpB[i] = {cB[3 * i], cB[3 * i + 1], cB[3 * i + 2]};
    }
}

void ora_get_wildcards(const ora_ctx* c, float* w) {
    const size_t n = c->idA.size();
    for (int k = 0; k < 4; k++)
        if (c->w[k].size() == n) std::copy(c->w[k].begin(), c->w[k].end(), w + k * n);
}

void ora_set_wildcards(ora_ctx* c, const float* w) {
    const size_t n = c->idA.size();
    for (int k = 0; k < 4; k++) c->w[k].assign(w + k * n, w + (k + 1) * n);
}

// FP64 restatement: identical formulas with every intermediate in double (tolerance reference for a15).
void ora_force_fp64(const ora_ctx* c, double* Fout, double* Tout) {
    const size_t n = c->idA.size();
    const double ts = c->p.h;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const size_t i = (size_t)ii;
        // geometry in double
        auto body = [&](uint32_t s, double& px, double& py, double& pz, double& ox, double& oy, double& oz, double q[4],
                        double& r, double& m, double v[3], double w[3], uint8_t& mat) {
            uint32_t o = c->sphOwner[s];
            owner_pos(c, o, ox, oy, oz);
            q[0] = c->qw[o]; q[1] = c->qx[o]; q[2] = c->qy[o]; q[3] = c->qz[o];
            double rx = c->sphRx[s], ry = c->sphRy[s], rz = c->sphRz[s];
            double X = (2 * (q[0] * q[0] + q[1] * q[1]) - 1) * rx + 2 * (q[1] * q[2] - q[0] * q[3]) * ry +
                       2 * (q[1] * q[3] + q[0] * q[2]) * rz;
            double Y = 2 * (q[1] * q[2] + q[0] * q[3]) * rx + (2 * (q[0] * q[0] + q[2] * q[2]) - 1) * ry +
                       2 * (q[2] * q[3] - q[0] * q[1]) * rz;
            double Z = 2 * (q[1] * q[3] - q[0] * q[2]) * rx + 2 * (q[2] * q[3] + q[0] * q[1]) * ry +
                       (2 * (q[0] * q[0] + q[3] * q[3]) - 1) * rz;
            px = ox + X; py = oy + Y; pz = oz + Z;
            r = c->sphR[s];
            m = c->mass[o];
            v[0] = c->vX[o]; v[1] = c->vY[o]; v[2] = c->vZ[o];
            w[0] = c->wX[o]; w[1] = c->wY[o]; w[2] = c->wZ[o];
            mat = c->sphMat[s];
        };
        auto rot = [](const double q[4], double v[3], bool inv) {
            double qw = q[0], qx = inv ? -q[1] : q[1], qy = inv ? -q[2] : q[2], qz = inv ? -q[3] : q[3];
            double x = v[0], y = v[1], z = v[2];
            v[0] = (2 * (qw * qw + qx * qx) - 1) * x + 2 * (qx * qy - qw * qz) * y + 2 * (qx * qz + qw * qy) * z;
            v[1] = 2 * (qx * qy + qw * qz) * x + (2 * (qw * qw + qy * qy) - 1) * y + 2 * (qy * qz - qw * qx) * z;
            v[2] = 2 * (qx * qz - qw * qy) * x + 2 * (qy * qz + qw * qx) * y + (2 * (qw * qw + qz * qz) - 1) * z;
        };
        double pa[3], oa[3], qa[4], ra, ma, va[3], wa[3];
        double pb[3], ob[3], qb[4], rb, mb, vb[3], wb[3];
        uint8_t matA, matB;
        body(c->idA[i], pa[0], pa[1], pa[2], oa[0], oa[1], oa[2], qa, ra, ma, va, wa, matA);
        double nrm[3], cp[3], delta;
        bool contact = true;
        if (c->ctype[i] == DEME_SPHERE_SPHERE_CONTACT) {
            body(c->idB[i], pb[0], pb[1], pb[2], ob[0], ob[1], ob[2], qb, rb, mb, vb, wb, matB);
            double d[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
            double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            delta = ra + rb - dist;
            if (delta < 0) contact = false;
            for (int k = 0; k < 3; k++) {
                nrm[k] = d[k] / dist;
                cp[k] = pb[k] + (rb - delta / 2) * nrm[k];
            }
        } else {
            const deme_anal_obj& a = c->anal[c->idB[i]];
            uint32_t o = a.owner;
            owner_pos(c, o, ob[0], ob[1], ob[2]);
            qb[0] = c->qw[o]; qb[1] = c->qx[o]; qb[2] = c->qy[o]; qb[3] = c->qz[o];
            double rel[3] = {a.relPos[0], a.relPos[1], a.relPos[2]};
            rot(qb, rel, false);
            double dir[3] = {a.rot[0], a.rot[1], a.rot[2]};
            rot(qb, dir, false);
            for (int k = 0; k < 3; k++) pb[k] = ob[k] + rel[k];
            rb = (double)(float)HUGE_FLOAT;
            mb = a.mass;
            vb[0] = c->vX[o]; vb[1] = c->vY[o]; vb[2] = c->vZ[o];
            wb[0] = c->wX[o]; wb[1] = c->wY[o]; wb[2] = c->wZ[o];
            matB = a.material;
            double dist = (pa[0] - pb[0]) * dir[0] + (pa[1] - pb[1]) * dir[1] + (pa[2] - pb[2]) * dir[2];
            delta = (ra - dist) / 2;
            if (delta < 0) contact = false;
            for (int k = 0; k < 3; k++) {
                nrm[k] = dir[k];
                cp[k] = pa[k] - dir[k] * (dist + delta);
            }
        }
        double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        if (contact) {
            const deme_material &MA = c->mats[matA], &MB = c->mats[matB];
            double E = 1.0 / ((1 - (double)MA.nu * MA.nu) / MA.E + (1 - (double)MB.nu * MB.nu) / MB.E);
            double G = 1.0 / (2.0 * (2 - (double)MA.nu) * (1 + (double)MA.nu) / MA.E +
                              2.0 * (2 - (double)MB.nu) * (1 + (double)MB.nu) / MB.E);
            double CoR = std::min(MA.CoR, MB.CoR), mu = std::max(MA.mu, MB.mu), Crr = std::max(MA.Crr, MB.Crr);
            double lA[3] = {cp[0] - oa[0], cp[1] - oa[1], cp[2] - oa[2]};
            double lB[3] = {cp[0] - ob[0], cp[1] - ob[1], cp[2] - ob[2]};
            rot(qa, lA, true);
            rot(qb, lB, true);
            double rA[3] = {wa[1] * lA[2] - wa[2] * lA[1], wa[2] * lA[0] - wa[0] * lA[2], wa[0] * lA[1] - wa[1] * lA[0]};
            double rB[3] = {wb[1] * lB[2] - wb[2] * lB[1], wb[2] * lB[0] - wb[0] * lB[2], wb[0] * lB[1] - wb[1] * lB[0]};
            rot(qa, rA, false);
            rot(qb, rB, false);
            double vrel[3], vt[3];
            for (int k = 0; k < 3; k++) vrel[k] = (va[k] + rA[k]) - (vb[k] + rB[k]);
            double proj = vrel[0] * nrm[0] + vrel[1] * nrm[1] + vrel[2] * nrm[2];
            for (int k = 0; k < 3; k++) vt[k] = vrel[k] - proj * nrm[k];
            double dt[3] = {c->w[0][i], c->w[1][i], c->w[2][i]};
            double dtime = c->w[3][i];
            for (int k = 0; k < 3; k++) dt[k] += ts * vt[k];
            double dp = dt[0] * nrm[0] + dt[1] * nrm[1] + dt[2] * nrm[2];
            for (int k = 0; k < 3; k++) dt[k] -= dp * nrm[k];
            dtime += ts;
            double meff = ma * mb / (ma + mb);
            double sRd = std::sqrt(delta * (ra * rb) / (ra + rb));
            double Sn = 2 * E * sRd;
            double loge = CoR < TINY_FLOAT ? std::log(TINY_FLOAT) : std::log(CoR);
            double beta = loge / std::sqrt(loge * loge + PI_SQUARED);
            double kn = TWO_OVER_THREE * Sn, gn = TWO_TIMES_SQRT_FIVE_OVER_SIX * beta * std::sqrt(Sn * meff);
            double fn = kn * delta + gn * proj;
            for (int k = 0; k < 3; k++) F[k] = fn * nrm[k];
            if (Crr > 0) {
                bool add = true;
                double Reff = std::sqrt(ra * rb / (ra + rb));
                double kns = FOUR_OVER_THREE * E * std::sqrt(Reff);
                double gns = -2 * std::sqrt(FIVE_OVER_THREE * meff * E) * beta * std::pow(Reff, 0.25);
                double dc = gns / (2 * std::sqrt(kns * meff));
                if (dc < 1.0) {
                    double tc = PI * std::sqrt(meff / (kns * (1 - dc * dc)));
                    if (dtime <= tc) add = false;
                }
                if (add) {
                    double vr[3] = {rB[0] - rA[0], rB[1] - rA[1], rB[2] - rA[2]};
                    double vm = std::sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
                    double fl = std::sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
                    if (vm > TINY_FLOAT)
                        for (int k = 0; k < 3; k++) T[k] = vr[k] / vm * (Crr * fl);
                }
            }
            if (mu > 0) {
                double kt = 8 * G * sRd, gt = -TWO_TIMES_SQRT_FIVE_OVER_SIX * beta * std::sqrt(meff * kt);
                double tf[3];
                for (int k = 0; k < 3; k++) tf[k] = -kt * dt[k] - gt * vt[k];
                double ft = std::sqrt(tf[0] * tf[0] + tf[1] * tf[1] + tf[2] * tf[2]);
                if (ft > TINY_FLOAT) {
                    double fmax = std::sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * mu;
                    if (ft > fmax)
                        for (int k = 0; k < 3; k++) tf[k] *= fmax / ft;
                } else {
                    tf[0] = tf[1] = tf[2] = 0;
                }
                for (int k = 0; k < 3; k++) F[k] += tf[k];
            }
        }
        for (int k = 0; k < 3; k++) {
            Fout[3 * i + k] = F[k];
            Tout[3 * i + k] = T[k];
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Row a16: collectContactForces (DEMCubForceCollection.cu:20-158, DEMCollectForceKernels.cu:11-133).
//
// Per-contact terms are forceToAcc (F * modifier / mass) and forceToAngAcc (cross(cp, R^T((F + Ftof) * modifier)) /
// MOI). The reference sums them with CUB ReduceByKey after a stable sort of [A-side entries; B-side entries] by owner.
// Order fixed here: each clump owner sums its A-side terms in ascending contact index, then its B-side terms ordered
// by (B sphere id, A sphere id), sequentially in float. Analytical owners (planes) are summed in double and rounded
// once.
// ---------------------------------------------------------------------------------------------------------------
namespace {

inline f3 ang_term(const ora_ctx* c, uint32_t o, f3 cpt, f3 F, f3 T, float mod) {
    f3 myF = (F + T) * mod;
    rotq(myF, c->qw[o], -c->qx[o], -c->qy[o], -c->qz[o]);
    return cross(cpt, myF) / mk(c->moiX[o], c->moiY[o], c->moiZ[o]);
}

inline uint32_t ownerB_of(const ora_ctx* c, size_t i) {
    return c->ctype[i] == DEME_SPHERE_SPHERE_CONTACT ? c->sphOwner[c->idB[i]] : c->anal[c->idB[i]].owner;
}

}  // namespace

int ora_collect(ora_ctx* c) {
    const size_t n = c->idA.size();
    const uint32_t nO = c->nO;
    // owner -> A contacts and B contacts, ascending contact index
    std::vector<uint32_t> aCnt(nO + 1, 0), bCnt(nO + 1, 0);
    for (size_t i = 0; i < n; i++) {
        aCnt[c->sphOwner[c->idA[i]] + 1]++;
        bCnt[ownerB_of(c, i) + 1]++;
    }
    for (uint32_t o = 0; o < nO; o++) {
        aCnt[o + 1] += aCnt[o];
        bCnt[o + 1] += bCnt[o];
    }
    std::vector<uint32_t> aList(n), bList(n);
    {
        std::vector<uint32_t> ca(aCnt.begin(), aCnt.end() - 1), cb(bCnt.begin(), bCnt.end() - 1);
        for (size_t i = 0; i < n; i++) {
            aList[ca[c->sphOwner[c->idA[i]]]++] = (uint32_t)i;
            bList[cb[ownerB_of(c, i)]++] = (uint32_t)i;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t oo = 0; oo < (int64_t)nO; oo++) {
        const uint32_t o = (uint32_t)oo;
        if (c->isAnalOwner[o]) continue;
        std::sort(bList.begin() + bCnt[o], bList.begin() + bCnt[o + 1], [c](uint32_t x, uint32_t y) {
            return c->idB[x] != c->idB[y] ? c->idB[x] < c->idB[y] : c->idA[x] < c->idA[y];
        });
    }
#pragma omp parallel for schedule(static)
    for (int64_t oo = 0; oo < (int64_t)nO; oo++) {
        const uint32_t o = (uint32_t)oo;
        const float m = c->mass[o];
        if (!c->isAnalOwner[o]) {
            f3 acc{0, 0, 0}, ang{0, 0, 0};
            for (uint32_t k = aCnt[o]; k < aCnt[o + 1]; k++) {
                const uint32_t i = aList[k];
                acc += c->F[i] * 1.f / m;
                ang += ang_term(c, o, c->cpA[i], c->F[i], c->Ftof[i], 1.f);
            }
            for (uint32_t k = bCnt[o]; k < bCnt[o + 1]; k++) {
                const uint32_t i = bList[k];
                acc += c->F[i] * -1.f / m;
                ang += ang_term(c, o, c->cpB[i], c->F[i], c->Ftof[i], -1.f);
            }
            if (aCnt[o + 1] > aCnt[o] || bCnt[o + 1] > bCnt[o]) {
                c->aX[o] = 0.f + acc.x;
                c->aY[o] = 0.f + acc.y;
                c->aZ[o] = 0.f + acc.z;
                c->alX[o] = 0.f + ang.x;
                c->alY[o] = 0.f + ang.y;
                c->alZ[o] = 0.f + ang.z;
            }
        } else {
            double acc[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
            for (uint32_t k = bCnt[o]; k < bCnt[o + 1]; k++) {
                const uint32_t i = bList[k];
                f3 t = c->F[i] * -1.f / m;
                f3 g = ang_term(c, o, c->cpB[i], c->F[i], c->Ftof[i], -1.f);
                acc[0] += t.x; acc[1] += t.y; acc[2] += t.z;
                ang[0] += g.x; ang[1] += g.y; ang[2] += g.z;
            }
            if (bCnt[o + 1] > bCnt[o]) {
                c->aX[o] = (float)acc[0]; c->aY[o] = (float)acc[1]; c->aZ[o] = (float)acc[2];
                c->alX[o] = (float)ang[0]; c->alY[o] = (float)ang[1]; c->alZ[o] = (float)ang[2];
            }
        }
    }
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Row a17: integrateOwners (DEMIntegrationKernels.cu:6-179) with the pass-on rules of
// IntegrationVelPassOn{ForwardEuler,CenteredDiff,ExtendedTaylor}.cu and numeric prescriptions
// (APIPrivate.cpp:1005-1048).
// ---------------------------------------------------------------------------------------------------------------
int ora_integrate(ora_ctx* c) {
    const float h = c->p.h;
    const float t = (float)c->time;
    (void)t;  // numeric prescriptions do not depend on t
#pragma omp parallel for schedule(static)
    for (int64_t oo = 0; oo < (int64_t)c->nO; oo++) {
        const uint32_t o = (uint32_t)oo;
        const deme_family_rule& R = c->rules[c->fam[o]];
        // integrateVel
        bool LinPrescribed = false, RotPrescribed = false;
        f3 old_v = mk(c->vX[o], c->vY[o], c->vZ[o]);
        f3 old_w = mk(c->wX[o], c->wY[o], c->wZ[o]);
        if (R.used) {
            if (R.velSet & 1) c->vX[o] = R.vel[0];
            if (R.velSet & 2) c->vY[o] = R.vel[1];
            if (R.velSet & 4) c->vZ[o] = R.vel[2];
            if (R.velSet & 8) c->wX[o] = R.vel[3];
            if (R.velSet & 16) c->wY[o] = R.vel[4];
            if (R.velSet & 32) c->wZ[o] = R.vel[5];
            LinPrescribed = R.linVelPrescribed;
            RotPrescribed = R.rotVelPrescribed;
        }
        f3 v_update{0, 0, 0}, w_update{0, 0, 0};
        if (!LinPrescribed) {
            v_update.x = (c->aX[o] + c->p.G[0]) * h;
            v_update.y = (c->aY[o] + c->p.G[1]) * h;
            v_update.z = (c->aZ[o] + c->p.G[2]) * h;
            c->vX[o] += v_update.x;
            c->vY[o] += v_update.y;
            c->vZ[o] += v_update.z;
        }
        if (!RotPrescribed) {
            w_update.x = c->alX[o] * h;
            w_update.y = c->alY[o] * h;
            w_update.z = c->alZ[o] * h;
            c->wX[o] += w_update.x;
            c->wY[o] += w_update.y;
            c->wZ[o] += w_update.z;
        }
        f3 v, w;
        switch (c->p.integrator) {
            case DEME_INTEGRATOR_FORWARD_EULER:
                v = old_v;
                w = old_w;
                break;
            case DEME_INTEGRATOR_CENTERED_DIFFERENCE:
                v = old_v + v_update;
                w = old_w + w_update;
                break;
            default:
                v = old_v + v_update * 0.5f;
                w = old_w + w_update * 0.5f;
        }
        // integratePos
        double X, Y, Z;
        owner_pos(c, o, X, Y, Z);
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
            X += (double)v.x * h;
            Y += (double)v.y * h;
            Z += (double)v.z * h;
        }
        pos_to_voxel(c, X, Y, Z, c->voxel[o], c->locX[o], c->locY[o], c->locZ[o]);
        if (!RotPrescribed) {
            double dq0 = 1, dq1 = w.x, dq2 = w.y, dq3 = w.z;
            double len = std::sqrt(dq1 * dq1 + dq2 * dq2 + dq3 * dq3);
            double theta = 0.5 * h * len;
            if (len > 0) {
                dq0 = std::cos(theta);
                double s = std::sin(theta) / len;
                dq1 *= s;
                dq2 *= s;
                dq3 *= s;
            }
            // HamiltonProduct<float> (DEMHelperKernels.cu:169-186)
            const float a1 = c->qw[o], b1 = c->qx[o], c1 = c->qy[o], d1 = c->qz[o];
            const float a2 = (float)dq0, b2 = (float)dq1, c2 = (float)dq2, d2 = (float)dq3;
            c->qw[o] = a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2;
            c->qx[o] = a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2;
            c->qy[o] = a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2;
            c->qz[o] = a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2;
        }
    }
    return DEME_OK;
}

// Row a18: deterministic schedule. Contact detection runs on the state at steps k * f' (f' = max(f, 1)) and its
// list serves steps k * f' .. k * f' + f' - 1; each step is force -> collect -> integrate, then t += h
// (dT.cpp:1527-1585 with the kT/dT staleness made deterministic).
int ora_step(ora_ctx* c, int nSteps, int updateFreq) {
    const uint64_t f = updateFreq > 0 ? (uint64_t)updateFreq : 1;
    for (int k = 0; k < nSteps; k++) {
        if (c->nSteps % f == 0 || !c->contactsReady) {
            int e = ora_cd(c);
            if (e) return e;
        }
        const double t0 = wall_now();
        int e = ora_force(c);
        if (e) return e;
        const double t1 = wall_now();
        ora_collect(c);
        const double t2 = wall_now();
        ora_integrate(c);
        const double t3 = wall_now();
        c->phase[3] += t1 - t0;
        c->phase[4] += t2 - t1;
        c->phase[5] += t3 - t2;
        c->time += (double)c->p.h;
        c->nSteps++;
    }
    return DEME_OK;
}

void ora_get_time(const ora_ctx* c, double* t, uint64_t* n) {
    if (t) *t = c->time;
    if (n) *n = c->nSteps;
}
void ora_set_time(ora_ctx* c, double t) { c->time = t; }
void ora_phase_times(ora_ctx* c, double* out, int reset) {
    for (int k = 0; k < 6; k++) {
        if (out) out[k] = c->phase[k];
        if (reset) c->phase[k] = 0.0;
    }
}

}  // extern "C"

void ora_set_scenes(ora_ctx* c, uint32_t nScenes, const uint32_t* ownerScene, const uint32_t* analScene) {
    c->nScenes = nScenes;
    c->ownerScene.assign(ownerScene, ownerScene + c->nO);
    if (analScene) c->analScene.assign(analScene, analScene + c->nA);
    else c->analScene.clear();
}
