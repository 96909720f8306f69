// mpm_scatter.hpp -- device helpers shared by the forward and adjoint
// particle kernels: tile-major cell indexing, active-tile marking, and the
// block-level P2G scatter (LDS-reduced per-chunk partials with a global
// atomic fallback).
#pragma once
#include <type_traits>
#include "mpm_device.hpp"
#include "mpm_internal.hpp"

namespace mpm {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

// The wave's index in its block, as a wave-uniform value: the compiler cannot
// tell that threadIdx.x >> 6 is the same on every lane, and a per-tile index
// built from it would turn the tile's record and scene loads into vector
// loads (one dependent VGPR load per field) instead of scalar ones.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)); }

// XCD-aware walk of a chunk list by a persistent grid.  Workgroups are
// dispatched round-robin over the 8 XCDs of an MI355X, each with its own L2,
// so block b runs on XCD b % 8.  Giving every XCD one contiguous range of the
// (spatially sorted) chunk list keeps the grid tiles its chunks touch in that
// XCD's L2 instead of replicating the whole grid in all eight.
constexpr uint32_t NUM_XCD = 8;
struct ChunkWalk {
    uint32_t ci, end, step;
};
__device__ __forceinline__ ChunkWalk chunk_walk(uint32_t nch) {
    const uint32_t g = gridDim.x;
    if (g < NUM_XCD || g % NUM_XCD != 0) return {blockIdx.x, nch, g};
    const uint32_t x = blockIdx.x % NUM_XCD;
    const uint32_t lo = uint32_t((uint64_t)nch * x / NUM_XCD);
    const uint32_t hi = uint32_t((uint64_t)nch * (x + 1) / NUM_XCD);
    return {lo + blockIdx.x / NUM_XCD, hi, g / NUM_XCD};
}

// Plane c of a particle array at one lane's slot, as the wave-uniform plane
// base plus the slot's 32-bit byte offset `boff` = 4 slot (capacities stay
// below 2^30 particles, create() enforces it), which lets the compiler use
// the SGPR-base form of global loads and stores.
__device__ __forceinline__ float plane_ld(const float* ps, uint32_t cap, int c, uint32_t boff) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(ps + (size_t)c * cap) + boff);
}
__device__ __forceinline__ void plane_st(float* ps, uint32_t cap, int c, uint32_t boff, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(ps + (size_t)c * cap) + boff) = v;
}

__device__ __forceinline__ uint32_t cell_index(const SceneDev& sc, int ix, int iy, int iz) {
    uint32_t t = sc.tile_base + uint32_t(ix >> 2) + sc.W * (uint32_t(iy >> 2) + sc.W * uint32_t(iz >> 2));
    return (t << 6) | uint32_t(ix & 3) | (uint32_t(iy & 3) << 2) | (uint32_t(iz & 3) << 4);
}

// Storage of tile t: the tile itself in a dense grid; in a sparse pool its
// slot of the current substep, or -1 when the pool gave it none (it then
// holds no P2G mass and no particle reads it).
__device__ __forceinline__ int storage_tile(const uint32_t* tile_slot, uint32_t t) {
    return tile_slot ? int(tile_slot[t]) - 1 : int(t);
}

// Storage index of cell (ix, iy, iz) in `c`; false when its tile has none.
__device__ __forceinline__ bool cell_storage(const SceneDev& sc, const uint32_t* tile_slot, int ix, int iy, int iz,
                                             uint32_t& c) {
    const uint32_t t = sc.tile_base + uint32_t(ix >> 2) + sc.W * (uint32_t(iy >> 2) + sc.W * uint32_t(iz >> 2));
    const int st = storage_tile(tile_slot, t);
    if (st < 0) return false;
    c = (uint32_t(st) << 6) | uint32_t(ix & 3) | (uint32_t(iy & 3) << 2) | (uint32_t(iz & 3) << 4);
    return true;
}

// The tiles one particle marks for the sparse pool: those of cells
// base-1 .. base+3 on each axis, clamped to the grid (mark_particles,
// mpm_active_tiles.wgsl:90-106).
__device__ __forceinline__ void mark_stencil_tiles(uint32_t* tile_mark, const SceneDev& sc, const int* b) {
    const int last = int(sc.G) - 1;
    int lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = max(b[k] - 1, 0);
        hi[k] = min(b[k] + 3, last);
        if (lo[k] > hi[k]) return;
        lo[k] >>= 2;
        hi[k] >>= 2;
    }
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x)
                tile_mark[sc.tile_base + uint32_t(x) + sc.W * (uint32_t(y) + sc.W * uint32_t(z))] = 1u;
}

// The CPU-diff floor (mpm_cpu_diff.rs:527-531): node row iy lies strictly
// below floor_y.  Evaluated in f64 from the scene's f32 values, where the
// product is exact, so a row whose f32 position would round onto floor_y is
// decided as the f64 oracle decides it (e.g. dx = 1/80, floor_y = 3 dx).
__device__ __forceinline__ bool below_floor(const SceneDev& sc, int iy) {
    return (double)sc.origin[1] + (double)iy * (double)sc.dx < (double)sc.floor_y;
}

__device__ __forceinline__ bool in_grid(const SceneDev& sc, int ix, int iy, int iz) {
    int n = int(sc.G);
    return ix >= 0 && iy >= 0 && iz >= 0 && ix < n && iy < n && iz < n;
}

__device__ __forceinline__ void global_add(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void mark_tile(uint32_t* flag, uint32_t* list, uint32_t* count, uint32_t t) {
    if ((__hip_atomic_load(&flag[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & (FLAG_DYN | FLAG_STATIC)) == 0u) {
        if ((atomicOr(&flag[t], FLAG_DYN) & (FLAG_DYN | FLAG_STATIC)) == 0u) {
            uint32_t slot = atomicAdd(count, 1u);
            list[slot] = t;
        }
    }
}

__device__ __forceinline__ void mark_gmp(uint32_t* flag, uint32_t t) {
    if ((__hip_atomic_load(&flag[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_GMP) == 0u)
        atomicOr(&flag[t], FLAG_GMP);
}

__device__ __forceinline__ void latch(uint32_t* status, uint32_t bits) {
    atomicOr(&status[ST_FAULT], bits);
}

// Wave-wide reductions with DPP lane moves (VALU only; a shuffle would be a
// ds_bpermute through the LDS pipe, which the scatter keeps busy).  Within
// each row of 16 lanes: swap neighbours, swap pairs, half-row mirror, row
// mirror; then rows 1 and 3 take rows 0 and 2 (row_bcast:15), rows 2 and 3
// take lane 31 (row_bcast:31), and lane 63 holds the result, read back as a
// wave-uniform value.  Lanes a step leaves out read `identity`, so each step
// fuses into one DPP min/max.  Every lane of the wave must be active.
template <typename Op>
__device__ __forceinline__ int wave_reduce_dpp(int v, int identity, Op op) {
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x142, 0xA, 0xF, false));  // row_bcast:15 into rows 1, 3
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x143, 0xC, 0xF, false));  // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min(int v) {
    return wave_reduce_dpp(v, 0x7fffffff, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max(int v) {
    return wave_reduce_dpp(v, int(0x80000000u), [](int a, int b) { return max(a, b); });
}

// Material row of a particle (mpm.wgsl:152-157): global table at scene base.
__device__ __forceinline__ const float* mat_row(const StepArgs& a, const SceneDev& sc, uint32_t m) {
    if (m >= sc.n_mat) m = 0;
    return a.mats + (size_t)(sc.mat_base + m) * MAT_STRIDE;
}

struct Scatter {
    bool on;   // the particle scatters this pass
    bool esc;  // ... with global atomics (stencil left its sort neighbourhood)
    uint32_t slot;  // its particle slot (a sparse pool defers its global-atomic scatter)
    int b[3];
    float fx[3];
    float m;
    float mv[3];
    float A[9];  // m C + 4/dx^2 stress
};

// Stencil base cell (window-relative for the shader model) and fractional
// offset of position x.
template <int MODEL>
__device__ __forceinline__ void stencil_of(const SceneDev& sc, const float* x, int* b, float* fx) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float xp = (x[k] - sc.origin[k]) * sc.inv_dx;
        const float bf = floorf(xp - 0.5f);
        b[k] = int(bf) - (MODEL == MODEL_SHADER ? sc.window[k] : 0);
        fx[k] = xp - bf;
    }
}

// A wave's scatter box and contribution bounds over particles already
// reduced (wave-uniform values): box cells lo..hi, and bounds on any single
// mass and momentum contribution (w <= 1 and |d| <= 1.5 dx).
struct WaveBounds {
    int lo[3] = {0x3fffffff, 0x3fffffff, 0x3fffffff};
    int hi[3] = {-0x3fffffff, -0x3fffffff, -0x3fffffff};
    float bm = 0.f, bp = 0.f;
};

// Maximum of non-negative floats (+inf included), whose bit patterns order
// as integers; a NaN lane counts as 0, as fmaxf would drop it.
__device__ __forceinline__ float wave_maxf(float v) {
    return __int_as_float(wave_max(v == v ? __float_as_int(v) : 0));
}

// Merges one particle per lane into w.  Every lane of the wave must be active.
__device__ __forceinline__ void wave_bounds_add(const Scatter& p, float dx, WaveBounds& w) {
    int lo[3] = {0x3fffffff, 0x3fffffff, 0x3fffffff}, hi[3] = {-0x3fffffff, -0x3fffffff, -0x3fffffff};
    float bm = 0.f, bp = 0.f;
    if (p.on && !p.esc) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = p.b[k];
            hi[k] = p.b[k] + 2;
        }
        bm = fabsf(p.m);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            bp = fmaxf(bp, fabsf(p.mv[r]) + 1.5f * dx * (fabsf(p.A[3 * r]) + fabsf(p.A[3 * r + 1]) +
                                                         fabsf(p.A[3 * r + 2])));
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w.lo[k] = min(w.lo[k], wave_min(lo[k]));
        w.hi[k] = max(w.hi[k], wave_max(hi[k]));
    }
    w.bm = fmaxf(w.bm, wave_maxf(bm));
    w.bp = fmaxf(w.bp, wave_maxf(bp));
}

// The scatter inputs of a thread's earlier particles of a chunk (passes with
// more than one particle per thread): position (the stencil is recomputed
// from it), mass, momentum, affine and the on/esc flags, one word per lane.
struct ScatterStash {
    float x[3][BLOCK];
    float m[BLOCK];
    float mv[3][BLOCK];
    float A[9][BLOCK];
    uint32_t flags[BLOCK];
    __device__ __forceinline__ void put(const Scatter& p, const float* x_) {
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k][t] = x_[k];
            mv[k][t] = p.mv[k];
        }
        m[t] = p.m;
#pragma unroll
        for (int k = 0; k < 9; ++k) A[k][t] = p.A[k];
        flags[t] = (p.on ? 1u : 0u) | (p.esc ? 2u : 0u);
    }
    // Relaxed atomic loads (plain ds_read): the compiler does not merge them
    // with a copy of a register-held Scatter on the other side of a branch,
    // a merge that would move that Scatter into scratch memory.
    template <class T>
    __device__ static __forceinline__ T ld(const T* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ uint32_t flag_bits() const { return ld(&flags[threadIdx.x]); }
    __device__ __forceinline__ void get(const SceneDev& sc, uint32_t slot, Scatter& p) const {
        const int t = threadIdx.x;
        const uint32_t f = ld(&flags[t]);
        p.on = (f & 1u) != 0u;
        p.esc = (f & 2u) != 0u;
        p.slot = slot;
        float xx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xx[k] = ld(&x[k][t]);
            p.mv[k] = ld(&mv[k][t]);
        }
        stencil_of<MODEL_SHADER>(sc, xx, p.b, p.fx);
        p.m = ld(&m[t]);
#pragma unroll
        for (int k = 0; k < 9; ++k) p.A[k] = ld(&A[k][t]);
    }
};

// LDS workspace of the block scatter.  The block's box is accumulated in
// fixed point: on gfx950 LDS integer atomics run at about the LDS read rate
// while LDS f32 atomic adds are ~20x slower, and integer sums are independent
// of the order in which lanes and waves arrive.
//   narrow (the shader model's forward P2G): two 64-bit words and one
//     32-bit word per cell, each added with one LDS atomic of raw conversion
//     bit patterns (no per-contribution offset removal; see NarrowWords):
//       acc[0]  mass, 2^-40 of the block's largest particle mass, so a cell's
//               mass keeps better than f32 relative precision however small
//               it is (the grid update's m <= 1e-9 test, mpm.wgsl:1019, is
//               then decided as in f32 arithmetic); its top 16 bits count
//               the contributions, which removes the conversion offsets;
//       acc[1]  (px, py) as two int32 halves, 2^-23 of the block's bound on
//               any single momentum value;
//       accz    pz as int32 (a 32-bit atomic conflicts on half as many LDS
//               banks as a 64-bit one);
//   wide (int64, four words): each f32 contribution converts exactly (2^-50
//     of the largest), for sums whose terms span many magnitudes and cancel
//     (the adjoint's dJ/d grid v, the differentiable model's P2G).
template <bool WIDE>
struct ScatterLdsT {
    // before the scatter the same space stages the gather's grid velocities
    // (LDS_CELLS float4 = the first two rows of acc)
    alignas(16) unsigned long long acc[WIDE ? 4 : 2][LDS_CELLS];
    uint32_t accz[WIDE ? 1 : LDS_CELLS];
    int red[WAVES * 6];
    float redf[WAVES * 2];
    int box[16];
    // sparse pool: the tiles the box's touched cells mark (bit per tile of
    // the box's tile range, box[8..13]); cleared after each flush
    uint32_t tmask[64];
    double scale[2];  // power-of-two fixed-point scale of the mass and momentum channels
};
using ScatterLds = ScatterLdsT<false>;

// The narrow scatter's words.  A momentum value q (|w q| < 2^22 after the
// scale) converts with one FMA: the f32 bits of w q + 1.5 * 2^23 are
// MAGIC32 + round(w q).  The mass converts with one f64 FMA: the f64 bits of
// w m + 1.5 * 2^52 are MAGIC64 + round(w m) for 0 <= w m < 2^51.  A cell's
// sums of N such terms carry N times each offset.  MAGIC64 is 0x4338 << 48,
// and the mass scale keeps every contribution below 2^39, so the 512 of a
// chunk of two particles per thread stay below 2^48:
//   acc[0] = (N 0x4338 mod 2^16) << 48 + sum(m),
// whose low 48 bits are the mass sum and whose top 16 bits give N
// (0x4338 = 8 * 0x867, and 0x867 * 0x1357 = 1 mod 2^13, N <= 512 < 2^13).
// N then removes N MAGIC32 from each int32 momentum sum.
struct NarrowWords {
    static constexpr uint32_t MAGIC32 = 0x4B400000u;  // bits of 1.5 * 2^23
    static constexpr int MASS_TOP = 40;               // mass scale: largest particle mass -> 2^40
    // N, the contributions a cell received, from its mass word (0: untouched)
    __device__ static __forceinline__ uint32_t count(unsigned long long wm) {
        return ((uint32_t(wm >> 48) >> 3) * 0x1357u) & 0x1fffu;
    }
    __device__ static __forceinline__ void unpack(unsigned long long wm, unsigned long long wxy, uint32_t wz,
                                                  long long& m, int& x, int& y, int& z) {
        const uint32_t n = count(wm);
        m = (long long)(wm & ((1ull << 48) - 1ull));
        z = (int)(wz - n * MAGIC32);
        unpack_lo_hi_raw(wxy - (unsigned long long)n * ((unsigned long long)MAGIC32 * ((1ull << 32) + 1ull)), x, y);
    }
    // the two int32 halves of a 64-bit sum lo + hi * 2^32 (each |.| < 2^31)
    __device__ static __forceinline__ void unpack_lo_hi_raw(unsigned long long s, int& lo, int& hi) {
        lo = (int)(uint32_t)s;
        hi = (int)(uint32_t)((s - (unsigned long long)(long long)lo) >> 32);
    }
};

// round(a * b) + MAGIC32 as u32 bits for |a * b| < 2^22.
__device__ __forceinline__ uint32_t raw32(float a, float b) {
    return __float_as_uint(__builtin_fmaf(a, b, 12582912.0f));
}

// round(a * b) + MAGIC64 as u64 bits for 0 <= a * b < 2^51 (one rounding).
__device__ __forceinline__ unsigned long long raw64(double a, double b) {
    return (unsigned long long)__double_as_longlong(__builtin_fma(a, b, 6755399441055744.0));
}

__device__ __forceinline__ unsigned long long pair32(uint32_t lo, uint32_t hi) {
    return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

__device__ __forceinline__ void lds_add_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// round(v * s) as a two's-complement int64 for |v * s| < 2^51: adding
// 1.5 * 2^52 leaves the rounded integer in the low mantissa bits.
__device__ __forceinline__ unsigned long long fixed64(float v, double s) {
    const double t = (double)v * s + 6755399441055744.0;
    return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000ll);
}

__device__ __forceinline__ void lds_add_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Smallest power of two >= |x| (x > 0), as its exponent.
__device__ __forceinline__ int exp2_ceil(float x) {
    int e;
    const float m = frexpf(x, &e);  // x = m * 2^e, m in [0.5, 1)
    return m == 0.5f ? e - 1 : e;
}

// One particle's 27-node P2G straight to global f32 atomics.
__device__ __forceinline__ void scatter_global(const StepArgs& a, const SceneDev& sc, const Scatter& p) {
    float w[3][3];
    quad_weights(p.fx[0], w[0]);
    quad_weights(p.fx[1], w[1]);
    quad_weights(p.fx[2], w[2]);
    const float dx = sc.dx;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                int ix = p.b[0] + i, iy = p.b[1] + j, iz = p.b[2] + k;
                if (!in_grid(sc, ix, iy, iz)) continue;
                float wt = w[0][i] * w[1][j] * w[2][k];
                float d0 = (float(i) - p.fx[0]) * dx, d1 = (float(j) - p.fx[1]) * dx,
                      d2 = (float(k) - p.fx[2]) * dx;
                const uint32_t t = cell_index(sc, ix, iy, iz) >> 6;
                uint32_t c;
                if (!cell_storage(sc, a.tile_slot, ix, iy, iz, c)) continue;  // pool overflow (latched)
                float* g = reinterpret_cast<float*>(&a.gmp[c]);
                global_add(g + 0, wt * p.m);
                global_add(g + 1, wt * (p.mv[0] + p.A[0] * d0 + p.A[1] * d1 + p.A[2] * d2));
                global_add(g + 2, wt * (p.mv[1] + p.A[3] * d0 + p.A[4] * d1 + p.A[5] * d2));
                global_add(g + 3, wt * (p.mv[2] + p.A[6] * d0 + p.A[7] * d1 + p.A[8] * d2));
                mark_tile(a.tile_flag, a.list_out, a.count_out, t);
                mark_gmp(a.tile_flag, t);
            }
}

// P2G of one block's particles (chunk `chunk`).  Every thread of the block
// must call this, with its last particle in `p` and, for PPT > 1 (shader
// model only), its PPT - 1 earlier ones in `st` (slots p.slot - BLOCK, ...)
// with their wave bounds already reduced into `pre`.  Contributions are reduced with LDS int64 atomics over the
// bounding box of the block's stencils; the box is stored as the chunk's
// partial with plain stores (no global atomics, no registration: the grid
// update finds the chunk through the sort).  Escaped particles, and blocks
// whose box exceeds LDS_CELLS, scatter with global atomics.  MASS = false
// skips the mass channel (left 0 in the partials) for scatters whose m is 0.
// SPARSE: the grid is a tile pool (a.tile_mark etc. are set): the chunk marks
// its tiles, and global-atomic scatters wait for the pool's slots.
template <bool MASS = true, bool WIDE = false, bool SPARSE = false, int PPT = 1>
__device__ __forceinline__ void block_scatter(const StepArgs& a, const SceneDev& sc, const Scatter& last,
                                              uint32_t chunk, ScatterLdsT<WIDE>& L,
                                              const ScatterStash* st = nullptr, const WaveBounds* pre = nullptr) {
    static_assert(PPT == 1 || !WIDE, "several particles per thread: shader model only");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float dx = sc.dx;
    // particle q of this thread (q = PPT - 1 is `last`)
    // (field by field: an aggregate copy here keeps the structs in scratch)
    auto particle = [&](int q, Scatter& p) __attribute__((always_inline)) {
        if (PPT == 1 || q == PPT - 1) {
            p.on = last.on;
            p.esc = last.esc;
            p.slot = last.slot;
            p.m = last.m;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                p.b[k] = last.b[k];
                p.fx[k] = last.fx[k];
                p.mv[k] = last.mv[k];
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) p.A[k] = last.A[k];
        } else {
            st[q].get(sc, last.slot - uint32_t(PPT - 1 - q) * BLOCK, p);
        }
    };
    // whether particle q scatters through LDS / needs the global path
    auto flags_of = [&](int q, bool& on, bool& esc) __attribute__((always_inline)) {
        if (PPT == 1 || q == PPT - 1) {
            on = last.on;
            esc = last.esc;
        } else {
            const uint32_t f = st[q].flag_bits();
            on = (f & 1u) != 0u;
            esc = (f & 2u) != 0u;
        }
    };
    WaveBounds wb;
    if (PPT > 1 && pre) wb = *pre;
    wave_bounds_add(last, dx, wb);
    const int* lo = wb.lo;
    const int* hi = wb.hi;
    const float bm = wb.bm, bp = wb.bp;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            L.red[wave * 6 + k] = lo[k];
            L.red[wave * 6 + 3 + k] = hi[k];
        }
        L.redf[wave * 2] = bm;
        L.redf[wave * 2 + 1] = bp;
    }
    __syncthreads();
    if (tid == 0) {
        int l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = L.red[k];
            h[k] = L.red[3 + k];
            for (int w = 1; w < WAVES; ++w) {
                l[k] = min(l[k], L.red[w * 6 + k]);
                h[k] = max(h[k], L.red[w * 6 + 3 + k]);
            }
            l[k] = max(l[k], 0);
            h[k] = min(h[k], int(sc.G) - 1);
        }
        float mm = L.redf[0], mp = L.redf[1];
        for (int w = 1; w < WAVES; ++w) {
            mm = fmaxf(mm, L.redf[w * 2]);
            mp = fmaxf(mp, L.redf[w * 2 + 1]);
        }
        // A quadratic B-spline weight product is at most 0.75^3 < 0.422, so
        // with narrow momentum values scaled below 2^23 every contribution w q
        // stays below 0.422 * 2^23 < 2^22 (raw32's range) and the 512 of a
        // two-particle-per-thread chunk below 2^31.  Narrow masses: below
        // 2^40, contributions below 2^39, sums of 512 below 2^48.  Wide values
        // (256 per chunk): below 2^50, contributions below 2^49, sums of 256
        // below 2^57.
        const int top = WIDE ? 50 : 23;
        L.scale[0] = mm > 0.f && mm < 1e30f ? ldexp(1.0, (WIDE ? 50 : NarrowWords::MASS_TOP) - exp2_ceil(mm)) : 1.0;
        L.scale[1] = mp > 0.f && mp < 1e30f ? ldexp(1.0, top - exp2_ceil(mp)) : 1.0;
        const int bx = h[0] - l[0] + 1, by = h[1] - l[1] + 1, bz = h[2] - l[2] + 1;
        int use = 2;  // nothing reduced in LDS
        if (bx > 0 && by > 0 && bz > 0)
            use = (a.use_lds && bx * by * bz <= LDS_CELLS && chunk < a.part_capacity) ? 1 : 0;
        L.box[0] = l[0];
        L.box[1] = l[1];
        L.box[2] = l[2];
        L.box[3] = bx;
        L.box[4] = by;
        L.box[5] = bz;
        L.box[6] = use;
        // sparse pool: the tile range of the box grown by one cell (the
        // range of tiles its touched cells can mark)
        const int last = int(sc.G) - 1;
        L.box[8] = max(l[0] - 1, 0) >> 2;
        L.box[9] = max(l[1] - 1, 0) >> 2;
        L.box[10] = max(l[2] - 1, 0) >> 2;
        L.box[11] = (min(h[0] + 1, last) >> 2) - L.box[8] + 1;
        L.box[12] = (min(h[1] + 1, last) >> 2) - L.box[9] + 1;
        L.box[13] = (min(h[2] + 1, last) >> 2) - L.box[10] + 1;
        if (use == 0 && a.use_lds) atomicAdd(&a.status[ST_FALLBACK], 1u);
        if (a.use_lds && chunk < a.part_capacity) {
            a.chunk_box[2 * (size_t)chunk] = make_int4(l[0], l[1], l[2], 0);
            a.chunk_box[2 * (size_t)chunk + 1] =
                use == 1 ? make_int4(bx, by, bz, bx * by * bz) : make_int4(0, 0, 0, 0);
        }
    }
    __syncthreads();
    const int use = L.box[6];
    if (use == 1) {
        const int x0 = L.box[0], y0 = L.box[1], z0 = L.box[2], bx = L.box[3], by = L.box[4];
        const int vol = bx * by * L.box[5];
        const double sm64 = L.scale[0], sp64 = L.scale[1];
        // narrow: momentum values are pre-scaled and rounded by raw32; wide:
        // node values stay unscaled and fixed64 scales them exactly
        const float sp = WIDE ? 1.0f : (float)sp64;
        for (int i = tid; i < vol; i += BLOCK) {
            if (WIDE && MASS) L.acc[0][i] = 0;
            if (!WIDE) L.acc[0][i] = 0;
            L.acc[1][i] = 0;
            if constexpr (WIDE) {
                L.acc[2][i] = 0;
                L.acc[3][i] = 0;
            } else {
                L.accz[i] = 0u;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int qi = 0; qi < PPT; ++qi) {
          bool q_on, q_esc;
          flags_of(qi, q_on, q_esc);
          if (q_on && !q_esc) {
            Scatter p;
            particle(qi, p);
            float w[3][3];
            quad_weights(p.fx[0], w[0]);
            quad_weights(p.fx[1], w[1]);
            quad_weights(p.fx[2], w[2]);
            // With d = ((i, j, k) - fx) dx the momentum term is linear in the
            // node offset: mv + A d = c + i a0 + j a1 + k a2, c = mv - dx A fx,
            // a_q = dx A[:, q]; all pre-scaled to fixed point.
            float a0[3], a1[3], a2[3], c[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                a0[r] = dx * p.A[3 * r] * sp;
                a1[r] = dx * p.A[3 * r + 1] * sp;
                a2[r] = dx * p.A[3 * r + 2] * sp;
                c[r] = p.mv[r] * sp - (a0[r] * p.fx[0] + a1[r] * p.fx[1] + a2[r] * p.fx[2]);
            }
            // narrow mass: w0 * (w1 w2 m s) in f64, where the row factor is an
            // exact product (24-bit by 24-bit) and raw64 rounds once
            const double msd = (double)p.m * sm64;
            const double w0d[3] = {(double)w[0][0], (double)w[0][1], (double)w[0][2]};
            const int G = int(sc.G);
            const int li0 = (p.b[0] - x0) + bx * ((p.b[1] - y0) + by * (p.b[2] - z0));
            // CHECK: test every node against the grid; a wave whose stencils
            // all lie inside (every particle after the position clamp) skips it
            auto nodes = [&](auto check) {
                constexpr bool CHECK = decltype(check)::value;
                // node values c + i a0 + j a1 + k a2, stepped by additions
                float zk[3] = {c[0], c[1], c[2]};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int iz = p.b[2] + k;
                    float yj[3] = {zk[0], zk[1], zk[2]};
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int iy = p.b[1] + j;
                        const bool row_in = !CHECK || (iz >= 0 && iz < G && iy >= 0 && iy < G);
                        if (row_in) {
                            const float wjk = w[1][j] * w[2][k];
                            const double wm = (double)wjk * msd;
                            const int lrow = li0 + bx * (j + by * k);
                            float q0 = yj[0], q1 = yj[1], q2 = yj[2];
#pragma unroll
                            for (int i = 0; i < 3; ++i) {
                                const int ix = p.b[0] + i;
                                if (!CHECK || (ix >= 0 && ix < G)) {
                                    const float wt = w[0][i] * wjk;
                                    if constexpr (WIDE) {
                                        if (MASS) lds_add_u64(&L.acc[0][lrow + i], fixed64(wt * p.m, sm64));
                                        lds_add_u64(&L.acc[1][lrow + i], fixed64(wt * q0, sp64));
                                        lds_add_u64(&L.acc[2][lrow + i], fixed64(wt * q1, sp64));
                                        lds_add_u64(&L.acc[3][lrow + i], fixed64(wt * q2, sp64));
                                    } else {
                                        lds_add_u64(&L.acc[0][lrow + i], raw64(w0d[i], MASS ? wm : 0.0));
                                        lds_add_u64(&L.acc[1][lrow + i], pair32(raw32(wt, q0), raw32(wt, q1)));
                                        lds_add_u32(&L.accz[lrow + i], raw32(wt, q2));
                                    }
                                }
                                if (i < 2) {
                                    q0 += a0[0];
                                    q1 += a0[1];
                                    q2 += a0[2];
                                }
                            }
                        }
                        if (j < 2) {
                            yj[0] += a1[0];
                            yj[1] += a1[1];
                            yj[2] += a1[2];
                        }
                    }
                    if (k < 2) {
                        zk[0] += a2[0];
                        zk[1] += a2[1];
                        zk[2] += a2[2];
                    }
                }
            };
            const bool interior = p.b[0] >= 0 && p.b[1] >= 0 && p.b[2] >= 0 && p.b[0] + 2 < G && p.b[1] + 2 < G &&
                                  p.b[2] + 2 < G;
            if (__ballot(!interior) == 0) nodes(std::false_type{});
            else nodes(std::true_type{});
          }
        }
        __syncthreads();
        float4* part = a.part + (size_t)chunk * LDS_CELLS;
        if constexpr (WIDE) {
            // int64 sums: exact in f64 up to 2^53, one rounding to f32 after the scale
            const double im = 1.0 / sm64, ip = 1.0 / sp64;
            for (int i = tid; i < vol; i += BLOCK)
                part[i] = make_float4(MASS ? (float)((double)(long long)L.acc[0][i] * im) : 0.f,
                                      (float)((double)(long long)L.acc[1][i] * ip),
                                      (float)((double)(long long)L.acc[2][i] * ip),
                                      (float)((double)(long long)L.acc[3][i] * ip));
        } else {
            // exact integer sums (NarrowWords): the mass through f64, the
            // int32 momenta with one rounding to f32, then power-of-two scales
            const double im = 1.0 / sm64;
            const float ip = (float)(1.0 / sp64);
            constexpr bool marking = SPARSE;
            const int tb[3] = {L.box[8], L.box[9], L.box[10]}, tw = L.box[11], th = L.box[12];
            for (int i = tid; i < vol; i += BLOCK) {
                long long sm;
                int sx, sy, sz;
                NarrowWords::unpack(L.acc[0][i], L.acc[1][i], L.accz[i], sm, sx, sy, sz);
                part[i] = make_float4(MASS ? (float)((double)sm * im) : 0.f, (float)sx * ip, (float)sy * ip,
                                      (float)sz * ip);
                if (marking && NarrowWords::count(L.acc[0][i]) != 0u) {
                    // a touched cell (inside some stencil) marks every tile
                    // within one cell of it: the union over the chunk's
                    // particles of their base-1..base+3 tiles
                    const int row = i / bx;
                    const int c[3] = {x0 + (i - row * bx), y0 + row % by, z0 + row / by};
                    int lo[3], hi[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = (max(c[k] - 1, 0) >> 2) - tb[k];
                        hi[k] = (min(c[k] + 1, int(sc.G) - 1) >> 2) - tb[k];
                    }
                    for (int z = lo[2]; z <= hi[2]; ++z)
                        for (int y = lo[1]; y <= hi[1]; ++y)
                            for (int x = lo[0]; x <= hi[0]; ++x) {
                                const int bit = x + tw * (y + th * z);
                                atomicOr(&L.tmask[bit >> 5], 1u << (bit & 31));
                            }
                }
            }
        }
        __syncthreads();
        if (!WIDE && SPARSE && tid < 64) {
            // the marked tiles of this chunk to the pool's mark table
            const int tw = L.box[11], th = L.box[12];
            uint32_t m = L.tmask[tid];
            L.tmask[tid] = 0u;
            while (m) {
                const int bit = tid * 32 + (__ffs(m) - 1);
                m &= m - 1u;
                const int x = bit % tw, y = (bit / tw) % th, z = bit / (tw * th);
                a.tile_mark[sc.tile_base + uint32_t(L.box[8] + x) +
                            sc.W * (uint32_t(L.box[9] + y) + sc.W * uint32_t(L.box[10] + z))] = 1u;
            }
        }
    }
#pragma unroll 1
    for (int qi = 0; qi < PPT; ++qi) {
        bool q_on, q_esc;
        flags_of(qi, q_on, q_esc);
        if (!q_on || (!SPARSE && !q_esc && use != 0)) continue;  // nothing but the LDS scatter
        Scatter p;
        particle(qi, p);
        if (SPARSE && p.on && !p.esc && use != 0) {
            // a stencil with no node inside the grid touches no cell of the box,
            // but its clamped window may still reach a face tile (:97-106)
            const int G = int(sc.G);
            if (p.b[0] + 2 < 0 || p.b[1] + 2 < 0 || p.b[2] + 2 < 0 || p.b[0] >= G || p.b[1] >= G || p.b[2] >= G)
                mark_stencil_tiles(a.tile_mark, sc, p.b);
        }
        if (p.on && (p.esc || use == 0)) {
            if (p.esc) atomicAdd(&a.status[ST_ESCAPES], 1u);
            if (SPARSE) {
                // sparse pool: the slots of these tiles are assigned after this
                // pass, so the global-atomic scatter waits for them
                mark_stencil_tiles(a.tile_mark, sc, p.b);
                a.defer_list[atomicAdd(&a.status[ST_DEFER], 1u)] = p.slot;
            } else {
                scatter_global(a, sc, p);
            }
        }
    }
}

// True when a stencil starting at cell `b` leaves the 3x3x3 tiles around the
// particle's sorted tile (tile key `key` at its slot), i.e. the chunk
// partials of the last sort cannot hold its contribution.  The sorted chunk's
// tile row (ty, tz) comes with the chunk, so the particle's tile x is its key
// minus the row's first key (no division); a chunk without a row tag takes
// the division.
__device__ __forceinline__ bool stencil_escapes(const SceneDev& sc, const Chunk& ch, uint32_t key, const int* b) {
    int tk[3];
    if (ch.pad & 0x80000000u) {
        tk[1] = int(ch.pad & 0x7fffu);
        tk[2] = int((ch.pad >> 15) & 0xffffu);
        const uint32_t tx = key - (sc.tile_base + sc.W * (uint32_t(tk[1]) + sc.W * uint32_t(tk[2])));
        if (tx >= sc.W) return true;  // not in the chunk's row: keep it off the partials
        tk[0] = int(tx);
    } else {
        const uint32_t tl = key - sc.tile_base;
        tk[0] = int(tl % sc.W);
        tk[1] = int((tl / sc.W) % sc.W);
        tk[2] = int(tl / (sc.W * sc.W));
    }
    bool esc = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (b[k] < 4 * tk[k] - 4 || b[k] > 4 * tk[k] + 5) esc = true;
    return esc;
}

// Loads the grid velocities of chunk `ci`'s last P2G box into L.acc (as
// float4, box-local index lx + bx (ly + by lz)) when that box describes this
// pass's stencils: the previous pass scattered the same chunk list from the
// current positions (a.box_gather) and reduced it in LDS (box extent > 0).
// Block-uniform; every thread calls it.  Returns whether the box is staged
// (origin in o, extent in d); the caller must not overwrite L.acc before
// every thread has finished its gather.
template <bool WIDE, bool SPARSE = false>
__device__ __forceinline__ bool stage_gather_box(const StepArgs& a, const SceneDev& sc, uint32_t ci,
                                                 ScatterLdsT<WIDE>& L, int4& o, int4& d) {
    if (!a.box_gather || !a.use_lds || ci >= a.part_capacity) return false;
    o = a.chunk_box[2 * (size_t)ci];
    d = a.chunk_box[2 * (size_t)ci + 1];
    if (d.x <= 0 || d.y <= 0 || d.z <= 0 || d.x * d.y * d.z > LDS_CELLS) return false;
    float4* stage = reinterpret_cast<float4*>(&L.acc[0][0]);
    // Box cell i = lx + bx (ly + by lz) of thread t: i = t + u BLOCK.  All of
    // a thread's loads are issued before its LDS stores, so the staging costs
    // one load latency rather than one per cell (nothing of the particle
    // pass is live yet, so the registers are free).
    constexpr int PER = 4;  // loads in flight per thread
    const int vol = d.x * d.y * d.z;
    const float inv_bx = 1.0f / float(d.x), inv_by = 1.0f / float(d.y);
#pragma unroll 1
    for (int i0 = 0; i0 < vol; i0 += PER * BLOCK) {
        float4 v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = i0 + int(threadIdx.x) + u * BLOCK;
            v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < vol) {
                // exact for the box's small integers: (n + 0.5) / b is never
                // within f32 rounding of an integer
                const int r = int((float(i) + 0.5f) * inv_bx), lx = i - r * d.x;
                const int lz = int((float(r) + 0.5f) * inv_by), ly = r - lz * d.y;
                uint32_t c;
                if (cell_storage(sc, SPARSE ? a.tile_slot : nullptr, o.x + lx, o.y + ly, o.z + lz, c)) v[u] = a.gv[c];
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = i0 + int(threadIdx.x) + u * BLOCK;
            if (i < vol) stage[i] = v[u];
        }
    }
    __syncthreads();
    return true;
}

// Adds to q the P2G partials of every chunk that may have touched cell
// (ix, iy, iz) of tile (tx, ty, tz): the chunks holding particles sorted into
// the tile's 3x3x3 neighbourhood, i.e. nine x-rows of consecutive tile keys
// whose slot ranges come from the tile lower-bound table.  A chunk spanning
// two rows is added once, and chunks are added in ascending id, so the sum
// order is fixed.  Called by all 64 lanes of a wave for one tile:
//   1. lanes 0-8 resolve the nine rows' chunk ranges at once,
//   2. the deduplicated ranges are laid out as one candidate list, one
//      candidate per lane, whose boxes load in parallel,
//   3. a ballot keeps the boxes overlapping the tile, and the wave adds
//      their partials in ascending order, U loads in flight at a time.
// part / nparts: this wave adds only the part-th of nparts contiguous ranges
// of the candidate list, in ascending order, starting from q (the waves of a
// block can share a tile's candidates; the caller adds their sums in part
// order).  nparts = 1: the whole list.  use_rng: rng holds the nine rows'
// candidate ranges from the tile's static record (GridDev::srec), lane r < 9
// holding row r's (first, last), read by the caller.
// U: overlapping chunks whose partial loads are in flight together.
template <int U = 8>
__device__ __forceinline__ void add_partials(const GridDev& g, const SceneDev& sc, int tx, int ty, int tz,
                                             int ix, int iy, int iz, float4& q, int part = 0, int nparts = 1,
                                             bool use_rng = false, int2 rng = int2{0, -1}) {
    const int lane = threadIdx.x & 63;
    const int W = int(sc.W);
    const int X0 = tx * 4, Y0 = ty * 4, Z0 = tz * 4;
    const int xa = max(tx - 1, 0), xb = min(tx + 1, W - 1);
    int rc0 = 0, rc1 = -1;  // empty row
    if (use_rng) {
        rc0 = rng.x;
        rc1 = rng.y;
    } else if (lane < 9) {
        const int nz = tz + lane / 3 - 1, ny = ty + lane % 3 - 1;
        if (nz >= 0 && nz < W && ny >= 0 && ny < W) {
            const uint32_t row = sc.tile_base + uint32_t(W * (ny + W * nz));
            const uint32_t s0 = g.tile_lb[row + xa], s1 = g.tile_lb[row + xb + 1];
            if (s0 < s1) {
                rc0 = int(g.chunk_of_slot[s0]);
                rc1 = min(int(g.chunk_of_slot[s1 - 1]), int(g.part_capacity) - 1);
            }
        }
    }
    int c0[9], off[10];
    int last = -1, total = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int a0 = __builtin_amdgcn_readlane(rc0, r), a1 = __builtin_amdgcn_readlane(rc1, r);
        const int lo = max(a0, last + 1);
        c0[r] = lo;
        off[r] = total;
        total += a1 >= lo ? a1 - lo + 1 : 0;
        last = max(last, a1);
    }
    off[9] = total;
    const int beg = total * part / nparts, end = total * (part + 1) / nparts;
    for (int base = beg; base < end; base += 64) {
        const int j = base + lane;
        int mine = -1;
        if (j < end) {
#pragma unroll
            for (int r = 0; r < 9; ++r)
                if (j >= off[r] && j < off[r + 1]) mine = c0[r] + (j - off[r]);
        }
        int4 o = make_int4(0, 0, 0, 0), d = make_int4(0, 0, 0, 0);
        bool hit = false;
        if (mine >= 0) {
            o = g.chunk_box[2 * (size_t)mine];
            d = g.chunk_box[2 * (size_t)mine + 1];
            hit = d.x > 0 && o.x < X0 + 4 && o.x + d.x > X0 && o.y < Y0 + 4 && o.y + d.y > Y0 && o.z < Z0 + 4 &&
                  o.z + d.z > Z0;
        }
        uint64_t m = __ballot(hit);
        while (m) {
            // up to U overlapping chunks per round: their loads are all in
            // flight before the (ascending-order) adds
            int ls[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ls[u] = m ? __ffsll((unsigned long long)m) - 1 : -1;
                m &= m - 1;
            }
            float4 r[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                r[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (ls[u] >= 0) {
                    const int l = ls[u];
                    const int ox = __builtin_amdgcn_readlane(o.x, l), oy = __builtin_amdgcn_readlane(o.y, l),
                              oz = __builtin_amdgcn_readlane(o.z, l);
                    const int bx = __builtin_amdgcn_readlane(d.x, l), by = __builtin_amdgcn_readlane(d.y, l),
                              bz = __builtin_amdgcn_readlane(d.z, l);
                    const int cid = __builtin_amdgcn_readlane(mine, l);
                    const int lx = ix - ox, ly = iy - oy, lz = iz - oz;
                    if (lx >= 0 && ly >= 0 && lz >= 0 && lx < bx && ly < by && lz < bz)
                        r[u] = g.part[(size_t)cid * LDS_CELLS + lx + bx * (ly + by * lz)];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (ls[u] >= 0) {
                    q.x += r[u].x;
                    q.y += r[u].y;
                    q.z += r[u].z;
                    q.w += r[u].w;
                }
            }
        }
    }
}

// A static tile's record as the grid kernels use it: tile, scene and (lanes
// 0-8) the nine rows' candidate ranges; valid = false for entries without one.
struct TileRec {
    uint32_t t, scene;
    int2 rng;
    bool valid;
};
__device__ __forceinline__ void load_tile_rec(const GridDev& g, uint32_t ns, uint32_t i, TileRec& r) {
    r.valid = g.srec && i < ns;
    r.t = 0u;
    r.scene = 0u;
    r.rng = make_int2(0, -1);
    if (r.valid) {
        const int* rec = g.srec + (size_t)i * SREC_WORDS;
        r.t = uint32_t(rec[0]);
        r.scene = uint32_t(rec[1]);
        const int lane = threadIdx.x & 63;
        if (lane < 9) r.rng = make_int2(rec[2 + 2 * lane], rec[3 + 2 * lane]);
    }
}

// The tiles the grid kernels visit, one wave per tile: the static list of
// the last sort and the dynamic list of a dense grid, or the slots in use of
// a sparse pool (whose storage tile is then the slot itself).
struct TileWalk {
    uint32_t n, ns;
    __device__ __forceinline__ TileWalk(const GridDev& g, const uint32_t* dcount) {
        if (g.tile_slot) {
            n = min(*g.pool_count, g.pool_capacity);
            ns = 0;
        } else {
            ns = *g.scount;
            n = ns + *dcount;
        }
    }
    // tile id and storage tile of entry i
    __device__ __forceinline__ void at(const GridDev& g, const uint32_t* dlist, uint32_t i, uint32_t& t,
                                       uint32_t& st) const {
        if (g.tile_slot) {
            t = g.slot_tile[i];
            st = i;
        } else {
            t = i < ns ? g.slist[i] : dlist[i - ns];
            st = t;
        }
    }
};

// Tile coordinates and this lane's cell of global tile t (one lane per cell).
__device__ __forceinline__ void tile_cell(const SceneDev& sc, uint32_t t, int lane, int& tx, int& ty, int& tz,
                                          int& ix, int& iy, int& iz) {
    const uint32_t tl = t - sc.tile_base;
    tx = int(tl % sc.W);
    ty = int((tl / sc.W) % sc.W);
    tz = int(tl / (sc.W * sc.W));
    ix = tx * 4 + (lane & 3);
    iy = ty * 4 + ((lane >> 2) & 3);
    iz = tz * 4 + (lane >> 4);
}

}  // namespace mpm
