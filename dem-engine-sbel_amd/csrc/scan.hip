// Device-wide exclusive prefix sums (the occupancy scans of contact detection: per-bin occupancy -> bin start,
// per-sphere contact counts -> row start, per-owner counts -> list start).
//
// Three phases per level, tiles of 4096 elements (256 lanes x 16 items, four wave64s): tile reduce -> scan of tile
// sums (recursively) -> tile scan plus offset. Deterministic; sums are carried in 64 bits so overflow of the 32-bit
// outputs is detectable from the returned total.
#include <mutex>

#include "deme_internal.h"

namespace deme_int {
namespace {

// Up to this many tiles (8.4M elements) the apply pass folds each tile's offset from the raw tile sums itself.
constexpr uint64_t SCAN_FOLD_TILES = 2048;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per lane across a 256-lane block; returns exclusive prefix, writes block total.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t& total) {
    __shared__ uint64_t waveTot[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t incl = wave_incl_scan(v, lane);
    if (lane == 63) waveTot[wid] = incl;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; w++) {
        const uint64_t x = waveTot[w];
        if (w < wid) pre += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return pre + incl - v;
}

// Element i of the scanned sequence: a[idx ? idx[i] : i] + (b ? b[i] : 0).
template <typename TIn>
__device__ __forceinline__ uint64_t scan_elem(const TIn* __restrict__ a, const uint32_t* __restrict__ idx,
                                              const TIn* __restrict__ b, uint64_t i) {
    uint64_t v = (uint64_t)a[idx ? (uint64_t)idx[i] : i];
    if (b) v += (uint64_t)b[i];
    return v;
}

template <typename TIn, int ITEMS = SCAN_ITEMS>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const TIn* __restrict__ a, const uint32_t* __restrict__ gidx,
                                                            const TIn* __restrict__ b, uint64_t n,
                                                            uint64_t* __restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * (uint64_t)(SCAN_BLOCK * ITEMS);
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint64_t idx = base + (uint64_t)i * SCAN_BLOCK + threadIdx.x;
        if (idx < n) s += scan_elem(a, gidx, b, idx);
    }
    uint64_t tot;
    block_excl_scan(s, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// tile scan: loads striped (coalesced), scans per-lane contiguous runs, stores striped.
template <typename TIn, typename TOut, int ITEMS = SCAN_ITEMS>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const TIn* a, const uint32_t* __restrict__ gidx,
                                                           const TIn* __restrict__ b, uint64_t n,
                                                           const uint64_t* __restrict__ offsets, TOut* out,
                                                           unsigned long long* totalOut,
                                                           const uint64_t* __restrict__ rawSums = nullptr) {
    // +1 padding per 16 entries keeps the per-lane contiguous walk free of 64-bit bank conflicts
    __shared__ uint64_t tile[SCAN_BLOCK * ITEMS + SCAN_BLOCK * ITEMS / 16];
    auto at = [](uint32_t k) { return k + (k >> 4); };
    const uint64_t base = (uint64_t)blockIdx.x * (uint64_t)(SCAN_BLOCK * ITEMS);
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = (uint32_t)i * SCAN_BLOCK + threadIdx.x;
        const uint64_t idx = base + k;
        tile[at(k)] = idx < n ? scan_elem(a, gidx, b, idx) : 0;
    }
    __syncthreads();
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) s += tile[at(threadIdx.x * ITEMS + i)];
    uint64_t tot;
    uint64_t run = block_excl_scan(s, tot);
    uint64_t off = offsets ? offsets[blockIdx.x] : 0;
    if (rawSums) {
        // the tile's offset folded here from the unscanned tile sums of the tiles before it (fixed order)
        uint64_t part = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += SCAN_BLOCK) part += rawSums[t];
        block_excl_scan(part, off);
    }
    run += off;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = threadIdx.x * ITEMS + i;
        const uint64_t v = tile[at(k)];
        tile[at(k)] = run;
        run += v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = (uint32_t)i * SCAN_BLOCK + threadIdx.x;
        const uint64_t idx = base + k;
        if (idx < n) out[idx] = (TOut)tile[at(k)];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint64_t total = off + tot;
        out[n] = (TOut)total;
        if (totalOut) *totalOut = total;
    }
}

// In-place exclusive scan of u64 values (the per-tile sums); out has n+1 slots.
hipError_t scan_u64_inplace(uint64_t* v, uint64_t n, uint64_t* scratch, hipStream_t st) {
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
        k_scan_apply<uint64_t, uint64_t><<<1, SCAN_BLOCK, 0, st>>>(v, nullptr, nullptr, n, nullptr, v, nullptr);
        return hipGetLastError();
    }
    uint64_t* sums = scratch;
    k_scan_reduce<uint64_t><<<(uint32_t)nb, SCAN_BLOCK, 0, st>>>(v, nullptr, nullptr, n, sums);
    hipError_t e = scan_u64_inplace(sums, nb, scratch + nb + 1, st);
    if (e != hipSuccess) return e;
    k_scan_apply<uint64_t, uint64_t><<<(uint32_t)nb, SCAN_BLOCK, 0, st>>>(v, nullptr, nullptr, n, sums, v, nullptr);
    return hipGetLastError();
}

// ---- two scans over one gather (the row and mirror-row starts of the cell path) ----
// Pass 1 reads each permuted pair of counts once, stores it in id order (tmp) and folds both tile sums; pass 2 scans
// the tile sums of both components in one block; pass 3 scans the tiles from tmp (sequential reads) and adds them.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan2_reduce(const uint32_t* __restrict__ a1,
                                                             const uint32_t* __restrict__ b1,
                                                             const uint32_t* __restrict__ a2,
                                                             const uint32_t* __restrict__ idx, uint64_t n,
                                                             uint2* __restrict__ tmp, uint64_t* __restrict__ sums1,
                                                             uint64_t* __restrict__ sums2) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN2_TILE;
    uint64_t s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < SCAN2_ITEMS; i++) {
        const uint64_t k = base + (uint64_t)i * SCAN_BLOCK + threadIdx.x;
        if (k < n) {
            const uint32_t p = idx ? idx[k] : (uint32_t)k;
            const uint32_t v1 = a1[p] + (b1 ? b1[k] : 0u), v2 = a2[p];
            tmp[k] = make_uint2(v1, v2);
            s1 += v1;
            s2 += v2;
        }
    }
    uint64_t t1, t2;
    block_excl_scan(s1, t1);
    block_excl_scan(s2, t2);
    if (threadIdx.x == 0) {
        sums1[blockIdx.x] = t1;
        sums2[blockIdx.x] = t2;
    }
}

// Pass 1 of the candidate-list pass's row scans: the counts are formed here, not read -- per sphere the touching
// candidates of its row plus its plane contacts, and the touching candidates with it as partner (through its
// transposed row) -- stored in tmp and folded into the tile sums. Each touching transposed slot t also gets its rank
// in the mirror row, tRank[t] (the touching slots before it in its row), from which the fill pass links the contact to
// its mirror slot (a rank by slot: consecutive slots are written together, where ranks by candidate would scatter).
// A failed pass (ctr->err) leaves them alone.
// Rows are short (a few candidates), so a lane does not walk its ITEMS spheres one after another: it loads the row
// bounds of all of them, then the first CAND_SPEC entries of every row and transposed row, then the flags of those,
// and only longer rows finish in a loop -- three rounds of loads in flight instead of a chain per sphere.
constexpr uint32_t CAND_SPEC = 2;
template <int ITEMS>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan2_reduce_cand(const uint32_t* __restrict__ candStart,
                                                                  const uint32_t* __restrict__ tStart,
                                                                  const uint32_t* __restrict__ tC,
                                                                  const uint8_t* __restrict__ acc,
                                                                  const uint32_t* __restrict__ numAnal, uint64_t n,
                                                                  uint32_t* __restrict__ tRank, uint2* __restrict__ tmp,
                                                                  uint64_t* __restrict__ sums1,
                                                                  uint64_t* __restrict__ sums2,
                                                                  const DevCounters* ctr) {
    if (ctr->err) return;
    const uint64_t base = (uint64_t)blockIdx.x * (uint64_t)(SCAN_BLOCK * ITEMS);
    // spheres in groups of G: with one group's loads in flight (70 VGPRs, seven waves per SIMD) a launch over 4M
    // spheres is resident at once; all eight spheres at a time take 118 VGPRs, four waves, and run slower
    constexpr int G = ITEMS < 4 ? ITEMS : 4;
    static_assert(ITEMS % G == 0, "whole groups");
    uint64_t s1 = 0, s2 = 0;
#pragma unroll
    for (int g = 0; g < ITEMS; g += G) {
        uint32_t c0[G], c1[G], t0[G], t1[G], v1[G], v2[G];
#pragma unroll
        for (int i = 0; i < G; i++) {
            const uint64_t k = base + (uint64_t)(g + i) * SCAN_BLOCK + threadIdx.x;
            const bool in = k < n;  // (a sphere past the end has empty rows)
            c0[i] = in ? candStart[k] : 0u;
            c1[i] = in ? candStart[k + 1] : 0u;
            t0[i] = in ? tStart[k] : 0u;
            t1[i] = in ? tStart[k + 1] : 0u;
            v1[i] = in ? numAnal[k] : 0u;
            v2[i] = 0u;
        }
        uint32_t q[G][CAND_SPEC];
#pragma unroll
        for (int i = 0; i < G; i++)
#pragma unroll
            for (uint32_t j = 0; j < CAND_SPEC; j++) {
                if (c0[i] + j < c1[i]) v1[i] += acc[c0[i] + j];
                q[i][j] = t0[i] + j < t1[i] ? tC[t0[i] + j] : 0u;
            }
        uint32_t f[G][CAND_SPEC];
#pragma unroll
        for (int i = 0; i < G; i++)
#pragma unroll
            for (uint32_t j = 0; j < CAND_SPEC; j++) f[i][j] = t0[i] + j < t1[i] ? acc[q[i][j]] : 0u;
#pragma unroll
        for (int i = 0; i < G; i++)
#pragma unroll
            for (uint32_t j = 0; j < CAND_SPEC; j++)
                if (f[i][j]) tRank[t0[i] + j] = v2[i]++;
#pragma unroll
        for (int i = 0; i < G; i++) {
            for (uint32_t c = c0[i] + CAND_SPEC; c < c1[i]; c++) v1[i] += acc[c];
            for (uint32_t t = t0[i] + CAND_SPEC; t < t1[i]; t++)
                if (acc[tC[t]]) tRank[t] = v2[i]++;
            const uint64_t k = base + (uint64_t)(g + i) * SCAN_BLOCK + threadIdx.x;
            if (k < n) tmp[k] = make_uint2(v1[i], v2[i]);
            s1 += v1[i];
            s2 += v2[i];
        }
    }
    uint64_t tot1, tot2;
    block_excl_scan(s1, tot1);
    block_excl_scan(s2, tot2);
    if (threadIdx.x == 0) {
        sums1[blockIdx.x] = tot1;
        sums2[blockIdx.x] = tot2;
    }
}

// One block: exclusive scans, in place, of both tile-sum arrays (nb <= SCAN_TILE entries each).
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan2_tiles(uint64_t* __restrict__ sums1, uint64_t* __restrict__ sums2,
                                                            uint32_t nb) {
    for (int c = 0; c < 2; c++) {
        uint64_t* v = c ? sums2 : sums1;
        const uint32_t b = threadIdx.x * SCAN_ITEMS;
        uint64_t x[SCAN_ITEMS];
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            x[i] = b + i < nb ? v[b + i] : 0;
            s += x[i];
        }
        uint64_t tot;
        uint64_t run = block_excl_scan(s, tot);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            if (b + i < nb) v[b + i] = run;
            run += x[i];
        }
    }
}

template <int ITEMS>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan2_apply(const uint2* __restrict__ tmp, uint64_t n,
                                                            const uint64_t* __restrict__ offs1,
                                                            const uint64_t* __restrict__ offs2,
                                                            uint32_t* __restrict__ out1, uint32_t* __restrict__ out2,
                                                            unsigned long long* total1, unsigned long long* total2,
                                                            bool fold) {
    // +1 padding per 16 entries: the per-lane contiguous walk hits distinct banks
    constexpr uint32_t TILE = SCAN_BLOCK * ITEMS;
    __shared__ uint32_t t1[TILE + TILE / 16], t2[TILE + TILE / 16];
    auto at = [](uint32_t k) { return k + (k >> 4); };
    const uint64_t base = (uint64_t)blockIdx.x * TILE;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = (uint32_t)i * SCAN_BLOCK + threadIdx.x;
        const uint2 v = base + k < n ? tmp[base + k] : make_uint2(0u, 0u);
        t1[at(k)] = v.x;
        t2[at(k)] = v.y;
    }
    __syncthreads();
    uint64_t s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        s1 += t1[at(threadIdx.x * ITEMS + i)];
        s2 += t2[at(threadIdx.x * ITEMS + i)];
    }
    uint64_t tot1, tot2;
    uint64_t o1, o2;
    if (fold) {
        // offs1 / offs2 hold the raw tile sums: fold the tiles before this one (fixed order)
        uint64_t p1 = 0, p2 = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += SCAN_BLOCK) {
            p1 += offs1[t];
            p2 += offs2[t];
        }
        block_excl_scan(p1, o1);
        block_excl_scan(p2, o2);
    } else {
        o1 = offs1[blockIdx.x];
        o2 = offs2[blockIdx.x];
    }
    uint64_t r1 = block_excl_scan(s1, tot1) + o1;
    uint64_t r2 = block_excl_scan(s2, tot2) + o2;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = at(threadIdx.x * ITEMS + i);
        const uint32_t v1 = t1[k], v2 = t2[k];
        t1[k] = (uint32_t)r1;
        t2[k] = (uint32_t)r2;
        r1 += v1;
        r2 += v2;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t k = (uint32_t)i * SCAN_BLOCK + threadIdx.x;
        if (base + k < n) {
            out1[base + k] = t1[at(k)];
            out2[base + k] = t2[at(k)];
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        out1[n] = (uint32_t)(o1 + tot1);
        out2[n] = (uint32_t)(o2 + tot2);
        *total1 = o1 + tot1;
        *total2 = o2 + tot2;
    }
}

}  // namespace

hipError_t scan2_u32_gather(const uint32_t* a1, const uint32_t* b1, const uint32_t* a2, const uint32_t* idx,
                            uint64_t n, uint32_t* out1, uint32_t* out2, unsigned long long* total1,
                            unsigned long long* total2, uint2* tmp, uint64_t* sums, hipStream_t st) {
    const uint64_t nb64 = n ? (n + SCAN2_TILE - 1) / SCAN2_TILE : 1;
    if (nb64 > SCAN_TILE) return hipErrorInvalidValue;  // callers scan such sizes with scan_u32
    const uint32_t nb = (uint32_t)nb64;
    k_scan2_reduce<<<nb, SCAN_BLOCK, 0, st>>>(a1, b1, a2, idx, n, tmp, sums, sums + nb);
    const bool fold = nb <= SCAN_FOLD_TILES;
    if (!fold) k_scan2_tiles<<<1, SCAN_BLOCK, 0, st>>>(sums, sums + nb, nb);
    k_scan2_apply<SCAN2_ITEMS><<<nb, SCAN_BLOCK, 0, st>>>(tmp, n, sums, sums + nb, out1, out2, total1, total2, fold);
    return hipGetLastError();
}

hipError_t scan2_cand_rows(const uint32_t* candStart, const uint32_t* tStart, const uint32_t* tC, const uint8_t* acc,
                           const uint32_t* numAnal, uint64_t n, uint32_t* tRank, uint32_t* out1, uint32_t* out2,
                           unsigned long long* total1, unsigned long long* total2, uint2* tmp, uint64_t* sums,
                           const DevCounters* ctr, hipStream_t st) {
    if (n <= SCAN2_SMALL_N) {
        // small lists: one sphere per lane, so the latency chain of each lane is one row
        const uint32_t nb = (uint32_t)(n ? (n + SCAN_BLOCK - 1) / SCAN_BLOCK : 1);
        k_scan2_reduce_cand<1><<<nb, SCAN_BLOCK, 0, st>>>(candStart, tStart, tC, acc, numAnal, n, tRank, tmp, sums,
                                                          sums + nb, ctr);
        k_scan2_apply<1><<<nb, SCAN_BLOCK, 0, st>>>(tmp, n, sums, sums + nb, out1, out2, total1, total2, true);
        return hipGetLastError();
    }
    const uint64_t nb64 = (n + SCAN2_TILE - 1) / SCAN2_TILE;
    if (nb64 > SCAN_TILE) return hipErrorInvalidValue;
    const uint32_t nb = (uint32_t)nb64;
    k_scan2_reduce_cand<SCAN2_ITEMS><<<nb, SCAN_BLOCK, 0, st>>>(candStart, tStart, tC, acc, numAnal, n, tRank, tmp,
                                                                sums, sums + nb, ctr);
    const bool fold = nb <= SCAN_FOLD_TILES;
    if (!fold) k_scan2_tiles<<<1, SCAN_BLOCK, 0, st>>>(sums, sums + nb, nb);
    k_scan2_apply<SCAN2_ITEMS><<<nb, SCAN_BLOCK, 0, st>>>(tmp, n, sums, sums + nb, out1, out2, total1, total2, fold);
    return hipGetLastError();
}

// ---- workspace of the allocation-free scan (deme_prefix_scan_u32_noalloc) ----
// The tile sums of scans of up to NOALLOC_SCAN_MAX elements, plus one slot for a total nobody asked for. One copy per
// GPU, held until the process ends. In host code the name `scratch` is a host-side stand-in, not a device address:
// noalloc_scan_scratch() looks the address up once per device and every launch takes that address.
constexpr uint64_t NOALLOC_SCRATCH_ELEMS = scan_scratch_elems(NOALLOC_SCAN_MAX) + 1;
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_ELEMS];

hipError_t noalloc_scan_scratch(int device, uint64_t** ptr, uint64_t* elems) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static uint64_t* table[MAX_DEVICES] = {};
    *ptr = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!table[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the lookup gives the current device's copy
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        table[device] = static_cast<uint64_t*>(p);
    }
    *ptr = table[device];
    if (elems) *elems = NOALLOC_SCRATCH_ELEMS;
    return hipSuccess;
}

hipError_t scan_u32(const uint32_t* a, const uint32_t* b, uint64_t n, uint32_t* out, unsigned long long* totalDev,
                    uint64_t* scratch, hipStream_t st, const uint32_t* idx) {
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (n == 0) {
        // out[0] = 0 and total = 0
        k_scan_apply<uint32_t, uint32_t><<<1, SCAN_BLOCK, 0, st>>>(a, idx, b, 0, nullptr, out, totalDev);
        return hipGetLastError();
    }
    if (nb <= 1) {
        k_scan_apply<uint32_t, uint32_t><<<1, SCAN_BLOCK, 0, st>>>(a, idx, b, n, nullptr, out, totalDev);
        return hipGetLastError();
    }
    uint64_t* sums = scratch;
    const uint64_t nbHalf = (n + SCAN2_TILE - 1) / SCAN2_TILE;
    if (nbHalf <= SCAN_FOLD_TILES) {
        // up to SCAN_FOLD_TILES half-size tiles: twice the workgroups of full tiles, and each tile folds its own
        // offset from the raw tile sums (no separate scan of the sums)
        k_scan_reduce<uint32_t, SCAN2_ITEMS><<<(uint32_t)nbHalf, SCAN_BLOCK, 0, st>>>(a, idx, b, n, sums);
        k_scan_apply<uint32_t, uint32_t, SCAN2_ITEMS><<<(uint32_t)nbHalf, SCAN_BLOCK, 0, st>>>(a, idx, b, n, nullptr,
                                                                                              out, totalDev, sums);
        return hipGetLastError();
    }
    k_scan_reduce<uint32_t><<<(uint32_t)nb, SCAN_BLOCK, 0, st>>>(a, idx, b, n, sums);
    hipError_t e = scan_u64_inplace(sums, nb, scratch + nb + 1, st);
    if (e != hipSuccess) return e;
    k_scan_apply<uint32_t, uint32_t><<<(uint32_t)nb, SCAN_BLOCK, 0, st>>>(a, idx, b, n, sums, out, totalDev);
    return hipGetLastError();
}

}  // namespace deme_int
