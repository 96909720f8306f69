// mpm_sort.hip -- stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Same contract as molla's RadixSortKernel (molla-compute-primitives/src/
// radix_sort.rs:1-30): 8-bit digits, stable, so ties keep their input
// order.  Only the key bits that can be non-zero are sorted (ceil(bits/8)
// passes).  Each pass is three launches:
//   k_hist    per-block digit histogram (blocks of 4096 elements, LDS atomics)
//   scan      device-wide exclusive scan of the digit-major histogram
//   k_scatter stable scatter: within a block, elements are ranked in index
//             order with 64-lane ballots (8 ballots give the lanes that share
//             a digit), per-wave digit counts combined in LDS.
#include "mpm_internal.hpp"

namespace mpm {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 4096
constexpr int RADIX = 256;

static inline uint32_t sort_blocks(uint32_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

size_t radix_scratch_words(uint32_t n) {
    // ping-pong keys + values, histogram (RADIX * blocks), its scan scratch
    const size_t h = (size_t)RADIX * sort_blocks(n);
    return 2 * (size_t)n + 2 * h + scan_scratch_words(h) + 64;
}

__global__ __launch_bounds__(SORT_THREADS) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                       int shift, uint32_t nblocks,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * SORT_TILE;
    for (int r = 0; r < SORT_ITEMS; ++r) {
        uint32_t i = base + r * SORT_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(SORT_THREADS) void k_scatter(const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint32_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, uint32_t n,
                                                          int shift, uint32_t nblocks,
                                                          const uint32_t* __restrict__ offs) {
    constexpr int W = SORT_THREADS / 64;
    __shared__ uint32_t base[RADIX];
    __shared__ uint32_t cnt[W][RADIX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t blk = blockIdx.x * SORT_TILE;
    for (int r = 0; r < SORT_ITEMS; ++r) {
#pragma unroll
        for (int w = 0; w < W; ++w) cnt[w][tid] = 0u;
        __syncthreads();
        const uint32_t i = blk + r * SORT_THREADS + tid;
        const bool valid = i < n;
        uint32_t key = 0u, val = 0u, d = 0u;
        if (valid) {
            key = kin[i];
            val = vin[i];
            d = (key >> shift) & 0xffu;
        }
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t m = __ballot(on);
            same &= on ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(same & lt);
        const bool leader = valid && (same & lt) == 0ull;
        if (leader) cnt[wave][d] = (uint32_t)__popcll(same);
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wave; ++w) before += cnt[w][d];
            pos = base[d] + before + rank;
        }
        __syncthreads();
        {
            uint32_t tot = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) tot += cnt[w][tid];
            base[tid] += tot;
        }
        if (valid) {
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
    }
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits,
                            uint32_t* scratch, hipStream_t s) {
    if (n <= 1 || bits <= 0) return hipSuccess;
    const uint32_t nb = sort_blocks(n);
    uint32_t* k2 = scratch;
    uint32_t* v2 = scratch + n;
    uint32_t* hist = scratch + 2 * (size_t)n;
    uint32_t* offs = hist + (size_t)RADIX * nb;
    uint32_t* sscr = offs + (size_t)RADIX * nb;
    const int passes = (bits + 7) / 8;
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_hist, dim3(nb), dim3(SORT_THREADS), 0, s, ka, n, shift, nb, hist);
        hipError_t e = exclusive_scan_u32(hist, offs, (size_t)RADIX * nb, sscr, nullptr, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(SORT_THREADS), 0, s, ka, va, kb, vb, n, shift, nb,
                           offs);
        uint32_t* t = ka;
        ka = kb;
        kb = t;
        t = va;
        va = vb;
        vb = t;
    }
    if (ka != keys) {
        hipError_t e = hipMemcpyAsync(keys, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace mpm
