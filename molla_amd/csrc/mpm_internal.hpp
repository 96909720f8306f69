// mpm_internal.hpp -- structures shared by the host side and the kernels of
// the MI355X MPM substep library.
//
// Storage layout in HBM (all device-resident for the life of a context):
//   * Particles: component-major SoA, NCOMP float planes of length `cap`
//     (plane k at ps + k * cap).  Planes: x(0-2) v(3-5) C(6-14, row-major)
//     F(15-23, row-major) jp(24) mass(25) vol(26) E(27) mat(28, u32 bits)
//     id(29, u32 bits: original particle index).  Two copies (ping-pong) so
//     the tile reorder is a pure gather.
//   * Grid: dense, tile-major.  Every scene owns W^3 tiles of 4x4x4 cells
//     (W = ceil(G/4)) starting at its `tile_base`; cell (ix,iy,iz) of scene s
//     lives at ((tile_base + tx + W*(ty + W*tz)) << 6) | lx | ly<<2 | lz<<4.
//     `gmp` holds the P2G accumulation (m, px, py, pz) and `gv` the grid
//     velocity (vx, vy, vz, m) after the grid update.  Only tiles on the
//     active list are touched in a substep.
//   * Sparse grid (a tile pool of fixed capacity, mpm_create_sparse): gmp and
//     gv hold `capacity` tiles; tile_slot[t] is 1 + the pool slot of tile t in
//     the current substep, 0 when t has none, and cell storage is
//     (slot << 6) | local.  Slots go to the tiles marked by the substep's
//     particles (every tile within one cell of a stencil, the reference's
//     base-1..base+3 rule) and rigid spheres.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int NCOMP = 30;
enum Comp : int {
    CX = 0, CV = 3, CC = 6, CF = 15, CJP = 24, CMASS = 25, CVOL = 26, CE = 27, CMAT = 28, CID = 29
};

constexpr int MODEL_SHADER = 0;
constexpr int MODEL_CPUDIFF = 1;

constexpr int RHEO_SAND = 0, RHEO_SNOW = 1, RHEO_VISCOUS = 2, RHEO_FLUID = 3;

constexpr uint32_t FAULT_OVERFLOW = 1u;
constexpr uint32_t FAULT_NUMERICAL = 2u;

// Status words in device memory.
// ST_COUNT0/1: the two dynamic active-tile lists; ST_SCOUNT: the static list.
// Sparse tile pool: ST_POOL the tiles marked by the last substep (slots in
// use, or the demand that overflowed the pool), ST_DEFER the particles whose
// global-atomic scatter waits for the slots.
enum StatusWord : int {
    ST_FAULT = 0, ST_GATHERS = 1, ST_COUNT0 = 2, ST_COUNT1 = 3, ST_SCOUNT = 4, ST_ESCAPES = 5, ST_NCHUNKS = 6,
    ST_FALLBACK = 7, ST_POOL = 8, ST_DEFER = 9, ST_WORDS = 10
};

struct SceneDev {
    float origin[3];
    float dx;
    float inv_dx;
    float gravity_y;
    float youngs;
    float poisson;
    float boundary_friction;
    float floor_y;
    float floor_friction;
    float mu;   // youngs / (2 (1 + nu)), f32 like mpm.wgsl:542
    float lam;  // youngs nu / ((1 + nu)(1 - 2 nu)), mpm.wgsl:543
    uint32_t G;
    uint32_t W;
    uint32_t tile_base;
    int32_t window[3];
    uint32_t model;
    uint32_t mat_base;
    uint32_t n_mat;
    uint32_t pbase;
    uint32_t n;
    uint32_t cbase;  // first chunk of the scene
};

// A contiguous run of particle slots of one scene, processed by one block.
struct Chunk {
    uint32_t scene;
    uint32_t start;
    uint32_t count;
    uint32_t pad;  // chunks of the sorted layout: chunk_row_tag of their tile row; else 0
};

// Chunks of the sorted layout lie in one tile row (ty, tz) of their scene.
__host__ __device__ inline uint32_t chunk_row_tag(uint32_t ty, uint32_t tz) {
    return 0x80000000u | ty | (tz << 15);
}

constexpr uint32_t CHUNK_PARTICLES = 256;  // one particle per thread of a block
// Particles per thread of the shader model's sorted-layout passes: a chunk of
// the sorted layout holds up to CHUNK_PARTICLES * ppt particles.
constexpr int MAX_PPT = 2;

constexpr int MAT_STRIDE = 8;  // rheology, dp_alpha, theta_c, theta_s, hardening, vdamp, adamp, bulk

enum Mode : int { MODE_P2G = 0, MODE_G2P = 1, MODE_G2P2G = 2 };

// Tile flag bits.
constexpr uint32_t FLAG_DYN = 1u;     // on the dynamic active list being built
constexpr uint32_t FLAG_GMP = 2u;     // gmp holds atomically accumulated data
constexpr uint32_t FLAG_STATIC = 4u;  // on the static active list of the last sort

// P2G partials.  Each block of the fused particle pass reduces its scatter in
// LDS over the bounding box of its particles' stencils and stores that box as
// its chunk's partial (plain stores).  Particles are kept sorted by 4^3 tile,
// so the chunks that can touch tile T are the chunks holding particles sorted
// into T's 3x3x3 tile neighbourhood; the grid update finds them from the
// tile lower-bound table of the last sort and sums their partials per cell.
// A particle whose stencil has left that neighbourhood since the sort is
// scattered with global atomics instead (and counted in ST_ESCAPES).
// 1600 cells of two 64-bit words and one 32-bit word (32 KB, plus a 17 KB
// stash with two particles per thread; the differentiable model's four int64
// channels: 51 KB) let three blocks share a CU's 160 KB LDS; a chunk's box (4
// or 8 tiles of one tile row plus the stencil reach and the motion since the
// last sort) stays well inside it, and a larger box falls back to global
// atomics (counted in ST_FALLBACK).
constexpr int LDS_CELLS = 1600;
constexpr int SREC_WORDS = 20;

struct GridDev {
    // sparse pool (null for a dense grid): 1 + slot per tile (0: none), the
    // tile of each slot, and the number of slots in use (status[ST_POOL],
    // capped at pool_capacity)
    const uint32_t* tile_slot;
    const uint32_t* slot_tile;
    const uint32_t* pool_count;
    uint32_t pool_capacity;
    float4* gmp;          // atomic accumulation / final P2G kept for readback
    float4* gv;           // grid velocity (vx, vy, vz, m)
    uint32_t* tile_flag;  // FLAG_* bits
    const uint32_t* tile_scene;
    const uint32_t* tile_lb;  // first sorted slot with key >= tile (total_tiles + 1)
    int4* chunk_box;      // per chunk: (x0, y0, z0, -), (bx, by, bz, vol); bx = 0: empty
    float4* part;         // per chunk: LDS_CELLS cells (m, px, py, pz)
    const uint32_t* slist;  // static active list
    const uint32_t* scount;
    // per static-list entry (SREC_WORDS ints, built at the sort; null: none):
    // tile, scene, then (first, last) candidate chunk of the nine tile rows
    // around the tile (last < first: none)
    const int* srec;
    const uint32_t* chunk_of_slot;  // sorted slot -> chunk of the fused pass
    uint32_t part_capacity;         // chunks with a partial slot
};

struct StepArgs {
    const SceneDev* scenes;
    const float* mats;
    float* ps;      // particle planes read by the pass
    float* ps_out;  // where a G2P pass writes x, v, C, F, jp (ps itself, or the
                    // next tape slot of the adjoint's recompute)
    uint32_t cap;
    float4* gmp;
    float4* gv;
    uint32_t* tile_flag;
    const uint32_t* sorted_keys;  // tile key of each slot at the last sort
    int4* chunk_box;
    float4* part;
    uint32_t* list_out;   // active-tile list being built by the scatter
    uint32_t* count_out;  // its counter
    uint32_t* status;
    const Chunk* chunks;
    uint32_t n_chunks;              // launch bound (and count when n_chunks_dev is null)
    const uint32_t* n_chunks_dev;   // device-built chunk count of the sorted layout
    uint32_t part_capacity;         // chunks >= this scatter with global atomics
    float dt;
    int use_lds;   // 0: every node goes straight to global atomics
    int write_vc;  // 1: store v and C every substep
    int has_E;     // per-particle Young's modulus plane is valid
    int box_gather;  // 1: chunk_box holds, per chunk of this pass, the P2G box of
                     // the current positions (the previous pass scattered them)
    float* aff_out;  // or null: 9 planes of a.cap floats that receive each scattered
                     // particle's P2G affine A = m C + stress term (the diff tape)
    // sparse pool (all null for a dense grid)
    const uint32_t* tile_slot;  // 1 + slot of each tile in the current substep
    uint32_t* tile_mark;        // tiles the scattering particles mark (nonzero) for the next compaction
    uint32_t* defer_list;       // slots of particles whose global-atomic scatter waits for the compaction
};

struct RigidDev;  // rigid bodies on the grid (mpm_rigid.hpp)

// ---- sparse tile pool (mpm_kernels.hip) ----
// Slots for the marked tiles (marks cleared), in arrival order; more marked
// tiles than `capacity` latch FAULT_OVERFLOW.  status[ST_POOL] = marked tiles.
hipError_t launch_pool_compact(uint32_t* tile_mark, uint32_t total_tiles, uint32_t* tile_slot, uint32_t* slot_tile,
                               uint32_t capacity, uint32_t* status, hipStream_t s);
// Marks of legacy rigid spheres (mark_rigids, mpm_active_tiles.wgsl:108-124).
hipError_t launch_mark_spheres(const SceneDev* scenes, const RigidDev& rg, uint32_t* tile_mark, uint32_t capacity,
                               uint32_t* status, hipStream_t s);
// The deferred global-atomic P2G of the particles listed in a.defer_list.
hipError_t launch_deferred_p2g(const StepArgs& a, int model, hipStream_t s);
// Mass and momentum of the first `cells` storage cells (pool order).
hipError_t launch_storage_to_host_layout(const float4* gmp, uint32_t cells, float* mass, float* mom, hipStream_t s);

// ---- launchers (mpm_kernels.hip) ----
hipError_t launch_particle_pass(int mode, int model, const StepArgs& a, int ppt, int blocks,
                                hipStream_t s);
// keep (may be null): per-cell copy of the P2G total (m, p) of the listed tiles.
// split: four waves per tile instead of one (same results; for scenes with no
// more active tiles than blocks).
hipError_t launch_grid_update(const SceneDev* scenes, const GridDev& g, const uint32_t* dlist,
                              const uint32_t* dcount, uint32_t* next_count, float dt,
                              int zero_after, int use_partials, int blocks, hipStream_t s,
                              float4* keep = nullptr, const RigidDev* rigid = nullptr, int split = 0);
// Shell force of legacy rigid spheres on the active tiles (after the update).
hipError_t launch_rigid_shell(const SceneDev* scenes, const GridDev& g, const uint32_t* dlist,
                              const uint32_t* dcount, float dt, const RigidDev& rg, int blocks, hipStream_t s);
hipError_t launch_zero_f64(double* p, uint32_t n, hipStream_t s);
hipError_t launch_clear_tiles(const GridDev& g, const uint32_t* dlist, uint32_t* dcount, int blocks,
                              hipStream_t s, float4* also = nullptr);
// `iters` semi-implicit Jacobi sweeps of g.gv over the active tiles (scratch:
// a second cell buffer); the result is left in g.gv.  Tile flags must still
// mark this substep's dynamic tiles (FLAG_DYN).
hipError_t launch_implicit_sweeps(const SceneDev* scenes, const GridDev& g, float4* scratch, uint32_t iters,
                                  const uint32_t* dlist, const uint32_t* dcount, float dt, int blocks,
                                  hipStream_t s, const RigidDev* rigid = nullptr);
hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t launch_snap_counters(const uint32_t* status, uint32_t* dst, hipStream_t s);
hipError_t launch_copy_u32(const uint32_t* src, uint32_t* dst, hipStream_t s);
hipError_t launch_sum_pairs(const double* v, const uint32_t* count, uint32_t cap, double* out, hipStream_t s);
// Static active list of a sort: the 3x3x3 dilation of occupied tiles, in
// ascending tile order (mark/pos: total_tiles u32 each; scan_scratch:
// scan_scratch_words(total_tiles)).  Also builds the tile lower-bound table.
hipError_t launch_static_rebuild(const SceneDev* scenes, const uint32_t* tile_scene,
                                 uint32_t* tile_flag, uint32_t* slist, uint32_t* scount,
                                 const uint32_t* sorted_keys, uint32_t n, uint32_t* tile_lb,
                                 uint32_t total_tiles, uint32_t* mark, uint32_t* pos, uint32_t* scan_scratch,
                                 hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// The static tiles' records (GridDev::srec) for the sorted chunk list:
// max_static bounds the static count (read on the device).
hipError_t launch_static_records(const SceneDev* scenes, const uint32_t* tile_scene, const uint32_t* slist,
                                 const uint32_t* scount, uint32_t max_static, const uint32_t* tile_lb,
                                 const uint32_t* chunk_of_slot, uint32_t part_capacity, int* srec, hipStream_t s);
hipError_t launch_tile_keys(const StepArgs& a, uint32_t* keys_out, int by_slot, hipStream_t s);
hipError_t launch_chunk_spread(const SceneDev* scenes, const Chunk* chunks, const uint32_t* n_chunks,
                               uint32_t max_chunks, const float* ps, uint32_t cap, const uint32_t* src_in,
                               uint32_t* src_out, uint32_t* keys, uint32_t chunk_cap, hipStream_t s);
hipError_t launch_reorder(const float* src, float* dst, uint32_t cap, const uint32_t* src_slot,
                          uint32_t n, hipStream_t s);
hipError_t launch_index_slots(const float* ps, uint32_t cap, uint32_t* slot_of_id, uint32_t n,
                              hipStream_t s);
hipError_t launch_unpermute(const float* ps, uint32_t cap, int comp0, int ncomp,
                            const uint32_t* slot_of_id, float* out, uint32_t n, hipStream_t s);
hipError_t launch_permute_in(float* ps, uint32_t cap, int comp0, int ncomp,
                             const uint32_t* slot_of_id, const float* in, uint32_t n,
                             hipStream_t s);
// Soil checkpoint: planes -> 40-word streams (ids in original order), and
// streams -> planes in identity slot order; aux = 12 words per particle.
hipError_t launch_ckpt_pack(const float* ps, uint32_t cap, const uint32_t* slot_of_id, const uint32_t* aux,
                            int zero_w, uint32_t n, uint32_t* out, hipStream_t s);
hipError_t launch_ckpt_unpack(const uint32_t* in, uint32_t n, float* ps, uint32_t cap, float youngs,
                              uint32_t* aux, hipStream_t s);
hipError_t launch_grid_to_dense(const float4* gmp, const SceneDev* scenes, uint32_t scene,
                                uint32_t G, float* mass, float* mom, hipStream_t s);
// Deterministic gather-order P2G.
hipError_t launch_gather_prepare(const StepArgs& a, float* affine, int blocks, hipStream_t s);
hipError_t launch_gather_p2g(const StepArgs& a, const uint32_t* tile_scene, const uint32_t* tile_lb,
                             const float* affine, const uint32_t* slist, const uint32_t* scount,
                             int blocks, hipStream_t s);
hipError_t launch_reorder_index(const uint32_t* sorted_ids, const uint32_t* slot_of_id,
                                uint32_t* src_slot, uint32_t n, hipStream_t s);

// ---- reverse pass of the CPUDIFF substep (mpm_diff.hip) ----
// Adjoint particle planes (same slot order and plane stride as the state).
enum AdjComp : int { AX = 0, AV = 3, AC = 6, AF = 15, AE = 24, NADJ = 25 };

struct AdjArgs {
    StepArgs a;    // a.ps: the substep's pre-state; a.gv: its grid velocity;
                   // a.gmp: atomic target of the reverse G2P scatter
    float* adj;    // NADJ planes of a.cap floats
    float4* gadj;  // per cell: (dJ/dm, dJ/dp) after the grid reverse
    double* gml;   // per chunk: accumulated (dJ/dmu, dJ/dlambda)
    const float* aff;  // the substep's P2G affine A (9 planes), recorded by its forward
};
// Reverse G2P + x/F advance: gathers grid v, scatters dJ/d(grid v) (partials),
// folds the G2P part of dJ/dx into adj x and replaces adj F by (I + dt C)^T adj F.
hipError_t launch_adj_g2p(const AdjArgs& q, int blocks, hipStream_t s);
// Grid reverse over the active tiles: dJ/d(grid v) -> (dJ/dm, dJ/dp) in gadj;
// g.gmp holds the substep's P2G total (m, p).
// use_esc = 0: gadj holds no escaped-particle part (it is overwritten).
// split: four waves per tile (small scenes; same choice as the forward's).
hipError_t launch_adj_grid(const SceneDev* scenes, const GridDev& g, float4* gadj, const uint32_t* dlist,
                           const uint32_t* dcount, float dt, int use_esc, int blocks, hipStream_t s,
                           int split = 0);
// Reverse P2G + stress: gathers gadj; writes adj x, v, C, F of the pre-state
// and accumulates adj E and the per-chunk (mu, lambda) sums.
hipError_t launch_adj_p2g(const AdjArgs& q, int blocks, hipStream_t s);
// Reverse P2G of the substep of q.a.ps fused with the reverse G2P of the
// previous substep (pre-state ps_prev, grid velocity gv_prev).  Only valid
// when neither scatters with global atomics.
hipError_t launch_adj_fused(const AdjArgs& q, const float* ps_prev, const float4* gv_prev, int blocks,
                            hipStream_t s);

// ---- scans and the sorted chunk list (mpm_scan.hip) ----
size_t scan_scratch_words(size_t n);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch,
                              uint32_t** total, hipStream_t s);
hipError_t build_sorted_chunks(const SceneDev* scenes, const uint32_t* tile_scene, const uint32_t* keys,
                               const uint32_t* lb, uint32_t n, uint32_t* flags, uint32_t* excl,
                               uint32_t* scan_scratch, Chunk* chunks, uint32_t* n_chunks,
                               uint32_t* chunk_of_slot, uint32_t chunk_cap, uint32_t seg_tiles, uint32_t total_tiles,
                               hipStream_t s);

// ---- radix sort (mpm_sort.hip) ----
size_t radix_scratch_words(uint32_t n);
// Stable LSD sort of (key, value) pairs on the low `bits` bits of the keys.
// keys/vals are sorted in place; scratch holds radix_scratch_words(n) u32.
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits,
                            uint32_t* scratch, hipStream_t s);

}  // namespace mpm
