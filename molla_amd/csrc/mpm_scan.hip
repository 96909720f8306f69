// mpm_scan.hip -- device-wide exclusive scan of u32 and the device-built
// chunk list of the fused particle pass.
//
// Chunks (one block of the particle pass each, `chunk_cap` = 256 threads x
// particles per thread) are cut from the sorted particle array so that a
// chunk never leaves its *segment*: consecutive tiles along x in one tile row
// of one scene.  With eight-tile segments allowed, a group of eight tiles
// that holds at least two chunks' worth of particles is one segment; a
// sparser group is two segments of four.  A segment's last chunk is partly
// empty, and eight-tile segments halve the idle lanes this costs on dense
// scenes, where a chunk covers about one tile anyway.  In a sparse row a
// chunk may spread over its whole segment, and four tiles keep its P2G box
// (18 x 6 x 6 cells at the sort, plus the motion since) well inside the
// block's LDS budget.
#include <algorithm>

#include "mpm_internal.hpp"

namespace mpm {

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

size_t scan_scratch_words(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + SCAN_THREADS + 2; }

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_warp, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_warp[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const int nw = blockDim.x >> 6;
        uint32_t w = lane < nw ? s_warp[lane] : 0u;
        uint32_t wi = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t t = (uint32_t)__shfl_up((int)wi, o, 64);
            if (lane >= o) wi += t;
        }
        if (lane < nw) s_warp[lane] = wi - w;
        if (lane == nw - 1) s_warp[63] = wi;
    }
    __syncthreads();
    const uint32_t r = incl - v + s_warp[wave];
    if (total) *total = s_warp[63];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                              uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_warp[64];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot = 0;
    block_excl_scan(s, s_warp, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of `nb` block sums in place by one block; total in sums[nb].
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nb) {
    __shared__ uint32_t s_warp[64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t off = 0; off < nb; off += SCAN_THREADS) {
        const uint32_t i = off + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot = 0;
        const uint32_t e = block_excl_scan(v, s_warp, &tot);
        const uint32_t c = carry;
        if (i < nb) sums[i] = c + e;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t* __restrict__ in, size_t n,
                                                            const uint32_t* __restrict__ sums,
                                                            uint32_t* __restrict__ out) {
    __shared__ uint32_t s_warp[64];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t run = block_excl_scan(s, s_warp, nullptr) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// out[i] = sum(in[0..i)); total (sum of all n) is left in scratch[nblocks].
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch,
                              uint32_t** total, hipStream_t s) {
    const uint32_t nb = uint32_t((n + SCAN_TILE - 1) / SCAN_TILE);
    if (total) *total = scratch + nb;
    if (n == 0) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, scratch, 0u);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, scratch);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, scratch, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, scratch, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Chunk list of the sorted particle array.  Two rules, both keeping a chunk
// inside one tile row (so its scatter box is at most the row's 4 x 4 cross
// section plus the stencil reach in y and z):
//   greedy rows (span 4 or 6 tiles; 6 is the default): a chunk starts at
//     each row's first particle and runs for chunk_cap particles or until
//     the next particle would make it span more than `span` tiles, so chunks
//     start anywhere inside a tile and only a row's last chunk (or one cut by
//     the span) is partly filled;
//   aligned segments: chunks of chunk_cap from the start of each aligned
//     group of four tiles, or of eight where those eight hold at least two
//     chunks' worth.
// Greedy rows of four tiles make ~10 % fewer chunks than the aligned rule on
// the 2 M scene and ~14 % fewer on the 512-scene batch; six tiles ~11 % and
// ~18 % (boxes up to 27 x 7 x 7 cells at the sort, inside the 1600-cell LDS
// box with room for the motion until the next sort).
//
// The aligned segment of tile `key`: its first tile and its length along x.
// lb is the tile lower-bound table of the sorted keys.
// dense_group: particles that make an eight-tile group one segment (~0u: never).
__device__ __forceinline__ void segment_of(const SceneDev& sc, const uint32_t* lb, uint32_t key, uint32_t dense_group,
                                           uint32_t& first, uint32_t& ntiles) {
    const uint32_t tl = key - sc.tile_base;
    const uint32_t tx = tl % sc.W;
    const uint32_t g8 = tx & ~7u, n8 = min(8u, sc.W - g8);
    const uint32_t first8 = key - (tx - g8);
    if (lb[first8 + n8] - lb[first8] >= dense_group) {
        first = first8;
        ntiles = n8;
        return;
    }
    const uint32_t g4 = tx & ~3u;
    first = key - (tx - g4);
    ntiles = min(4u, sc.W - g4);
}

__global__ void k_chunk_flags(const SceneDev* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                              const uint32_t* __restrict__ keys, const uint32_t* __restrict__ lb, uint32_t n,
                              uint32_t chunk_cap, uint32_t dense_group, uint32_t* __restrict__ flags) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t key = keys[j];
        const SceneDev& sc = scenes[tile_scene[key]];
        uint32_t first, nt;
        segment_of(sc, lb, key, dense_group, first, nt);
        const uint32_t seg_begin = lb[first];
        flags[j] = ((j - seg_begin) % chunk_cap) == 0u ? 1u : 0u;
    }
}

// Greedy rows: one thread per tile row marks its chunks' first slots in
// `flags` (zero elsewhere, cleared before).  Each step is one table lookup:
// the chunk ends at chunk_cap particles or at the first slot of the tile
// `span` tiles past the tile holding its first particle, whichever is first.
__global__ void k_row_chunks(const SceneDev* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                             const uint32_t* __restrict__ lb, uint32_t total_tiles, uint32_t chunk_cap, uint32_t span,
                             uint32_t* __restrict__ flags) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total_tiles; t += gridDim.x * blockDim.x) {
        const SceneDev& sc = scenes[tile_scene[t]];
        if ((t - sc.tile_base) % sc.W) continue;  // not a row's first tile
        const uint32_t row_end = t + sc.W;        // one past the row's last tile
        const uint32_t e_row = lb[row_end];
        uint32_t s = lb[t], tt = t;
        while (s < e_row) {
            flags[s] = 1u;
            while (lb[tt + 1] <= s) ++tt;  // the tile holding slot s
            s = min(s + chunk_cap, lb[min(tt + span, row_end)]);
        }
    }
}

// Greedy rows: a chunk's count is the distance to the next chunk's start.
__global__ void k_chunk_counts(Chunk* __restrict__ chunks, const uint32_t* __restrict__ nch, uint32_t n) {
    const uint32_t m = *nch;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x)
        chunks[c].count = (c + 1 < m ? chunks[c + 1].start : n) - chunks[c].start;
}

__global__ void k_chunk_emit(const SceneDev* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                             const uint32_t* __restrict__ keys, const uint32_t* __restrict__ lb, uint32_t n,
                             uint32_t chunk_cap, uint32_t dense_group, bool greedy, const uint32_t* __restrict__ flags,
                             const uint32_t* __restrict__ excl,
                             Chunk* __restrict__ chunks, uint32_t* __restrict__ chunk_of_slot) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t c = excl[j] + flags[j] - 1u;  // chunk holding slot j
        chunk_of_slot[j] = c;
        if (!flags[j]) continue;
        const uint32_t key = keys[j];
        const uint32_t scn = tile_scene[key];
        const SceneDev& sc = scenes[scn];
        Chunk ch;
        ch.scene = scn;
        ch.start = j;
        ch.count = 0u;  // greedy rows: k_chunk_counts
        if (!greedy) {
            uint32_t first, nt;
            segment_of(sc, lb, key, dense_group, first, nt);
            ch.count = min(chunk_cap, lb[first + nt] - j);
        }
        // the tile row of the chunk's segment (read by stencil_escapes)
        const uint32_t r = (key - sc.tile_base) / sc.W;
        ch.pad = chunk_row_tag(r % sc.W, r / sc.W);
        chunks[c] = ch;
    }
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) { *dst = *src; }

// One device word to another, on the stream (no host round trip).
hipError_t launch_copy_u32(const uint32_t* src, uint32_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(1), 0, s, src, dst);
    return hipGetLastError();
}

hipError_t build_sorted_chunks(const SceneDev* scenes, const uint32_t* tile_scene, const uint32_t* keys,
                               const uint32_t* lb, uint32_t n, uint32_t* flags, uint32_t* excl,
                               uint32_t* scan_scratch, Chunk* chunks, uint32_t* n_chunks,
                               uint32_t* chunk_of_slot, uint32_t chunk_cap, uint32_t seg_tiles, uint32_t total_tiles,
                               hipStream_t s) {
    if (chunk_cap == 0 || chunk_cap > CHUNK_PARTICLES * MAX_PPT || (seg_tiles != 4 && seg_tiles != 6 && seg_tiles != 8))
        return hipErrorInvalidValue;
    const bool greedy = seg_tiles != 8;  // 4, 6: greedy rows of at most that many tiles
    const uint32_t dense_group = 2 * chunk_cap;
    const unsigned blocks = std::max(1u, std::min<unsigned>((n + 255) / 256, 8192u));
    if (greedy) {
        hipError_t e = launch_fill_u32(flags, 0u, n, s);
        if (e != hipSuccess) return e;
        const unsigned tb = std::max(1u, std::min<unsigned>((total_tiles + 255) / 256, 8192u));
        hipLaunchKernelGGL(k_row_chunks, dim3(tb), dim3(256), 0, s, scenes, tile_scene, lb, total_tiles, chunk_cap,
                           seg_tiles, flags);
    } else {
        hipLaunchKernelGGL(k_chunk_flags, dim3(blocks), dim3(256), 0, s, scenes, tile_scene, keys, lb, n, chunk_cap,
                           dense_group, flags);
    }
    uint32_t* total = nullptr;
    hipError_t e = exclusive_scan_u32(flags, excl, n, scan_scratch, &total, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_chunk_emit, dim3(blocks), dim3(256), 0, s, scenes, tile_scene, keys, lb, n, chunk_cap,
                       dense_group, greedy, flags, excl, chunks, chunk_of_slot);
    if (greedy) hipLaunchKernelGGL(k_chunk_counts, dim3(blocks), dim3(256), 0, s, chunks, total, n);
    hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(1), 0, s, total, n_chunks);
    return hipGetLastError();
}

}  // namespace mpm
