// mpm_capi.hip -- host runtime behind include/molla_mpm.h.
//
// One context (`mpm_sim`) owns a HIP stream and every device buffer of one or
// more scenes (a single scene is a batch of one).  Stepping never allocates
// and never copies to or from the host; the only host<->device traffic is the
// explicit set/read calls, mirroring the reference's persistent-buffer design
// (book/src/compute/persistent.md; mpm.rs:1030-1063).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/molla_mpm.h"
#include "mpm_internal.hpp"
#include "mpm_rigid.hpp"

using mpm::Chunk;
using mpm::SceneDev;

namespace {

constexpr uint32_t CHUNK = 256;

// Everything a captured step_n bakes in: the launches' arguments and the
// positions of its sorts (which follow from the interval and the substeps
// since the last sort).
struct GraphKey {
    uint32_t dt_bits, n, mode, ps_cur, list_cur, dirty, sort_start, write_vc, pblocks, gblocks, sort_every, has_E,
        chunking;
    bool operator<(const GraphKey& o) const {
        return std::tie(dt_bits, n, mode, ps_cur, list_cur, dirty, sort_start, write_vc, pblocks, gblocks, sort_every,
                        has_E, chunking) < std::tie(o.dt_bits, o.n, o.mode, o.ps_cur, o.list_cur, o.dirty,
                                                    o.sort_start, o.write_vc, o.pblocks, o.gblocks, o.sort_every,
                                                    o.has_E, o.chunking);
    }
};

struct ProfRec {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

struct mpm_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<mpm_scene_config> cfg;
    std::vector<SceneDev> sh;
    uint32_t n_scenes = 0;
    uint32_t total_tiles = 0;
    uint32_t cap = 0;
    uint32_t n = 0;
    std::vector<uint32_t> counts, pbase;
    // device buffers
    SceneDev* d_scenes = nullptr;
    float* d_mats = nullptr;
    float* ps[2] = {nullptr, nullptr};
    int ps_cur = 0;
    float4* gmp = nullptr;
    float4* gv = nullptr;
    uint32_t* tile_flag = nullptr;
    uint32_t* tile_scene = nullptr;
    uint32_t* tile_lb = nullptr;
    uint32_t* slist = nullptr;
    int* srec = nullptr;  // per static tile: id, scene, candidate chunk ranges (dense grids)
    bool srec_valid = false;  // srec describes the current static list and chunk list
    uint32_t* src_slot = nullptr;
    int4* chunk_box = nullptr;
    float4* part = nullptr;
    uint32_t max_chunks = 0;   // host chunk list (scene-aligned, unsorted layout)
    Chunk* schunks = nullptr;  // device-built chunk list of the sorted layout
    uint32_t max_schunks = 0;
    uint32_t part_cap = 0;     // chunks with an LDS partial slot
    uint32_t* chunk_of_slot = nullptr;
    uint32_t* cflags = nullptr;
    uint32_t* cexcl = nullptr;
    uint32_t* scan_scratch = nullptr;
    uint32_t* smark = nullptr;  // static-list rebuild: per-tile mark, its scan, scan scratch
    uint32_t* spos = nullptr;
    uint32_t* sscan = nullptr;
    uint32_t* list[2] = {nullptr, nullptr};
    uint32_t* status = nullptr;  // ST_WORDS u32
    uint32_t* slot_of_id = nullptr;
    uint32_t* keys = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* sort_scratch = nullptr;
    Chunk* d_chunks = nullptr;
    uint32_t n_chunks = 0;
    float* affine = nullptr;
    float* staging = nullptr;  // cap * 9 floats
    // sparse tile pool (mpm_create_sparse; pool_cap = 0: dense grid)
    uint32_t pool_cap = 0;        // tiles of storage
    uint32_t* tile_slot = nullptr;   // per tile: 1 + slot in the current substep, 0: none
    uint32_t* tile_mark = nullptr;   // per tile: marked by the scattering pass (cleared by the compaction)
    uint32_t* slot_tile = nullptr;   // per slot: its tile
    uint32_t* defer_list = nullptr;  // particles whose global-atomic scatter waits for the slots
    int list_cur = 0;
    bool grid_dirty = false;
    bool sorted = false;
    bool boxes_valid = false;  // chunk_box holds the P2G boxes of the current positions
    uint32_t since_sort = 0;
    uint32_t sorts = 0;
    // Bumped by everything that rewrites the sort state (keys, slot order,
    // chunk list, tile table) or the planes wholesale: a sort of any kind and
    // the autotuner's trials.  Unlike `sorts`, nothing ever puts it back, so
    // the adjoint can tell that its forward's layout is gone.
    uint64_t layout_gen = 0;
    int has_E = 0;
    // options
    int p2g_mode = MPM_P2G_FUSED;
    uint32_t jacobi_iters = 0;  // semi-implicit sweeps per substep (MpmTimeStep::Implicit)
    float4* gv2 = nullptr;      // their ping-pong buffer
    uint32_t sort_every = 16;
    int write_vc = 1;
    uint32_t pblocks = 0;
    uint32_t gblocks = 0;
    uint32_t ppt = 2;         // particles per thread asked for (shader model)
    uint32_t seg_tiles = 6;   // chunk rule: 4 or 6 greedy rows of at most that many tiles, 8 aligned segments
    uint32_t chunk_ppt = 1;   // particles per thread of the current sorted chunk list
    // profiling
    int profiling = 0;         // MPM_PROFILE_* flags
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> ev_pool;  // recorded events are reused, never created per launch
    double prof_ms[MPM_STAGE_COUNT] = {0};
    uint32_t prof_n[MPM_STAGE_COUNT] = {0};
    // graphs
    std::map<GraphKey, hipGraphExec_t> graphs;
    bool capturing = false;
    bool dry = false;  // walk the schedule's host bookkeeping without launching
    // differentiable path
    bool hold_order = false;   // no re-sort: the diff tape keeps one slot order
    std::vector<float*> ckpt;  // checkpoints of the 24 state planes
    uint32_t diff_steps = 0, diff_every = 0;
    float diff_dt = 0.f;
    uint32_t diff_sorts = 0;
    int diff_ps = 0;
    uint64_t diff_gen = 0;     // layout_gen when the forward finished
    float* tape = nullptr;     // pre-states of one checkpoint segment (NCOMP planes each)
    float4* gtape = nullptr;   // per substep of the segment: grid v, then P2G total (m, p)
    float* atape = nullptr;    // per substep of the segment: the P2G affine A (9 planes)
    uint32_t tape_len = 0;
    int tape_seg = -1;         // the segment the tape holds from the forward itself (-1: none)
    float* adj = nullptr;      // NADJ adjoint planes
    float4* gadj = nullptr;    // per-cell grid adjoint
    double* gml = nullptr;     // per-chunk (dJ/dmu, dJ/dlambda)
    bool adj_valid = false;
    // device words of the differentiable path: [0] saved gather count, [4, 8)
    // the summed (dJ/dmu, dJ/dlambda), then the escape/fallback counters at
    // the forward's segment boundaries (2 per boundary) and around each
    // segment's recompute in the backward (4 per segment)
    uint32_t* dsnap = nullptr;
    uint32_t dsnap_len = 0;
    std::vector<uint8_t> seg_fused;  // per segment: the forward saw no escape or fallback scatter
    void* hpin = nullptr;            // pinned host staging of the differentiable path
    size_t hpin_bytes = 0;
    // rigid bodies on the grid (set_rigid_spheres / set_prescribed_solids)
    float4* d_rig = nullptr;     // MPM_MAX_RIGID bodies x 5 rows
    double* d_wrench = nullptr;  // MPM_MAX_RIGID x 8: last substep's reaction
    uint32_t rigid_count = 0;
    bool prescribed = false;
    // soil checkpoint: the archived words the planes do not hold (position and
    // velocity w, C/F column pads, tire compaction; 12 per particle, by id)
    uint32_t* ck_aux = nullptr;
    bool aux_live = false;     // ck_aux holds a restored set (else those words are zero)
    bool w_stepped = false;    // a step ran since: position/velocity w read 0 (mpm.wgsl:1349-1350)
    bool fault_reset = false;  // a latched fault was cleared by mpm_reset_status: no archive
                               // until a step or a restore (a reset does not hide the fault)
    std::string err;
};

namespace {

int fail(mpm_sim* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

int hip_fail(mpm_sim* s, hipError_t e, const char* what) {
    return fail(s, MPM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(s, expr)                               \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hip_fail(s, _e, #expr); \
    } while (0)

uint32_t ceil_log2(uint32_t v) {
    uint32_t b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

SceneDev make_scene(const mpm_scene_config& c, uint32_t tile_base, uint32_t mat_base) {
    SceneDev d{};
    for (int k = 0; k < 3; ++k) {
        d.origin[k] = c.grid_origin[k];
        d.window[k] = c.grid_window_cells[k];
    }
    d.dx = c.cell_size;
    d.inv_dx = 1.0f / c.cell_size;
    d.gravity_y = c.gravity_y;
    d.youngs = c.youngs;
    d.poisson = c.poisson;
    d.boundary_friction = c.boundary_friction;
    d.floor_y = c.floor_y;
    d.floor_friction = c.floor_friction;
    d.mu = c.youngs / (2.0f * (1.0f + c.poisson));
    d.lam = c.youngs * c.poisson / ((1.0f + c.poisson) * (1.0f - 2.0f * c.poisson));
    d.G = c.grid_size;
    d.W = (c.grid_size + 3u) / 4u;
    d.tile_base = tile_base;
    d.model = c.model;
    d.mat_base = mat_base;
    d.n_mat = c.n_materials;
    return d;
}

std::string validate(const mpm_scene_config& c) {
    if (c.grid_size < 5 || c.grid_size > 4096) return "grid_size must be in [5, 4096]";
    if (!(c.cell_size > 0.0f) || !std::isfinite(c.cell_size)) return "cell_size must be finite and > 0";
    if (c.n_materials < 1 || c.n_materials > MPM_MAX_MATERIALS) return "n_materials must be in [1, 8]";
    if (c.model > 1) return "unknown model";
    if (!(c.poisson > -1.0f && c.poisson < 0.5f)) return "poisson must be in (-1, 0.5)";
    if (!std::isfinite(c.youngs) || !std::isfinite(c.gravity_y)) return "non-finite material/gravity";
    return "";
}

void drop_graphs(mpm_sim* s) {
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

void release(mpm_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
    for (auto& r : s->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    void* bufs[] = {s->d_scenes, s->d_mats, s->ps[0], s->ps[1], s->gmp, s->gv, s->tile_flag, s->tile_scene,
                    s->list[0], s->list[1], s->status, s->slot_of_id, s->keys, s->vals, s->sort_scratch,
                    s->d_chunks, s->affine, s->staging, s->tile_lb, s->slist,
                    s->src_slot, s->chunk_box, s->part, s->schunks, s->chunk_of_slot, s->cflags,
                    s->cexcl, s->scan_scratch, s->smark, s->spos, s->sscan, s->tile_slot, s->tile_mark,
                    s->slot_tile, s->defer_list, s->srec};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (float* c : s->ckpt) (void)hipFree(c);
    void* dbufs[] = {s->tape, s->gtape, s->atape, s->adj, s->gadj, s->gml, s->gv2, s->ck_aux, s->dsnap, s->d_rig, s->d_wrench};
    for (void* b : dbufs)
        if (b) (void)hipFree(b);
    if (s->hpin) (void)hipHostFree(s->hpin);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

mpm::StepArgs args(mpm_sim* s, float dt, int list_out) {
    mpm::StepArgs a{};
    a.scenes = s->d_scenes;
    a.mats = s->d_mats;
    a.ps = s->ps[s->ps_cur];
    a.ps_out = a.ps;
    a.cap = s->cap;
    a.gmp = s->gmp;
    a.gv = s->gv;
    a.tile_flag = s->tile_flag;
    a.sorted_keys = s->keys;
    a.chunk_box = s->chunk_box;
    a.part = s->part;
    a.list_out = s->list[list_out];
    a.count_out = s->status + mpm::ST_COUNT0 + list_out;
    a.status = s->status;
    a.chunks = s->d_chunks;
    a.n_chunks = s->n_chunks;
    a.n_chunks_dev = nullptr;
    a.part_capacity = s->part_cap;
    a.dt = dt;
    a.use_lds = s->p2g_mode == MPM_P2G_FUSED ? 1 : 0;
    a.write_vc = s->write_vc;
    a.has_E = s->has_E;
    a.box_gather = s->boxes_valid ? 1 : 0;
    // a sparse pool defers global-atomic scatters, which read v and C back
    if (s->pool_cap) {
        a.write_vc = 1;
        a.tile_slot = s->tile_slot;
        a.tile_mark = s->tile_mark;
        a.defer_list = s->defer_list;
    }
    return a;
}

mpm::GridDev griddev(mpm_sim* s) {
    mpm::GridDev g{};
    g.gmp = s->gmp;
    g.gv = s->gv;
    g.tile_flag = s->tile_flag;
    g.tile_scene = s->tile_scene;
    g.tile_lb = s->tile_lb;
    g.chunk_box = s->chunk_box;
    g.part = s->part;
    g.slist = s->slist;
    g.scount = s->status + mpm::ST_SCOUNT;
    g.srec = s->sorted && s->srec_valid ? s->srec : nullptr;
    g.chunk_of_slot = s->chunk_of_slot;
    g.part_capacity = s->part_cap;
    if (s->pool_cap) {
        g.tile_slot = s->tile_slot;
        g.slot_tile = s->slot_tile;
        g.pool_count = s->status + mpm::ST_POOL;
        g.pool_capacity = s->pool_cap;
    }
    return g;
}

// ---- profiling helpers ----
// Timing-only events: no system-scope fence when recorded, so recording
// writes back and invalidates no caches and does not slow the next launch.
hipError_t timing_event(hipEvent_t* e) { return hipEventCreateWithFlags(e, hipEventDisableSystemFence); }

hipEvent_t take_event(mpm_sim* s) {
    hipEvent_t e = nullptr;
    if (!s->ev_pool.empty()) {
        e = s->ev_pool.back();
        s->ev_pool.pop_back();
    } else {
        (void)timing_event(&e);
    }
    return e;
}

// Events around a launch group (the stages 0-4) or a whole call (5-6) on
// the handle's stream.  mark_end() places the end event early (the last
// call wins); otherwise the destructor places it.
struct Stage {
    mpm_sim* s;
    ProfRec r{};
    bool on, ended = false;
    Stage(mpm_sim* s_, int stage)
        : s(s_),
          on(!s_->capturing &&
             (s_->profiling & (stage >= MPM_STAGE_DIFF_FORWARD ? MPM_PROFILE_CALLS : MPM_PROFILE_STAGES))) {
        if (!on) return;
        r.stage = stage;
        r.a = take_event(s);
        r.b = take_event(s);
        (void)hipEventRecord(r.a, s->stream);
    }
    void mark_end() {
        if (!on) return;
        (void)hipEventRecord(r.b, s->stream);
        ended = true;
    }
    ~Stage() {
        if (!on) return;
        if (!ended) (void)hipEventRecord(r.b, s->stream);
        s->prof_pending.push_back(r);
    }
};

void flush_profile(mpm_sim* s) {
    if (s->prof_pending.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto& r : s->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {  // an unrecorded pair is not counted
            s->prof_ms[r.stage] += ms;
            s->prof_n[r.stage] += 1;
        }
        s->ev_pool.push_back(r.a);
        s->ev_pool.push_back(r.b);
    }
    s->prof_pending.clear();
}

int grid_blocks(mpm_sim* s) {
    if (s->gblocks) return int(s->gblocks);
    return 2048;
}
// Four waves per tile for small scenes, where the active tiles are about as
// many as the grid-update blocks or fewer (64 particles per block: a dense
// scene has a few hundred per tile), one otherwise.  Host values only, so the
// shape, and with it the summation order, is the same every run.
int grid_split(mpm_sim* s) {
    return uint64_t(s->n) <= 64ull * uint64_t(grid_blocks(s)) ? 1 : 0;
}
int particle_blocks(mpm_sim* s) {
    if (s->pblocks) return int(std::min(s->pblocks, s->n_chunks));
    return int(s->n_chunks);
}

// Cells of grid storage: every tile of a dense grid, the pool of a sparse one.
size_t grid_cells(const mpm_sim* s) { return (size_t)(s->pool_cap ? s->pool_cap : s->total_tiles) * 64; }

// The dense tile grid is allocated on first use, so contexts that only order
// particles (or read keys) never pay for it.
int ensure_grid(mpm_sim* s) {
    if (s->gmp) return MPM_OK;
    const size_t cells = grid_cells(s);
    if (s->pool_cap) {
        HIP_TRY(s, dalloc(&s->tile_slot, s->total_tiles));
        HIP_TRY(s, dalloc(&s->tile_mark, s->total_tiles));
        HIP_TRY(s, dalloc(&s->slot_tile, s->pool_cap));
        HIP_TRY(s, dalloc(&s->defer_list, s->cap));
        HIP_TRY(s, hipMemsetAsync(s->tile_slot, 0, (size_t)s->total_tiles * 4, s->stream));
        HIP_TRY(s, hipMemsetAsync(s->tile_mark, 0, (size_t)s->total_tiles * 4, s->stream));
    }
    HIP_TRY(s, dalloc(&s->gmp, cells));
    HIP_TRY(s, dalloc(&s->gv, cells));
    if (s->jacobi_iters > 0 && !s->gv2) HIP_TRY(s, dalloc(&s->gv2, cells));
    HIP_TRY(s, dalloc(&s->tile_flag, s->total_tiles));
    HIP_TRY(s, dalloc(&s->tile_scene, s->total_tiles));
    HIP_TRY(s, dalloc(&s->list[0], s->total_tiles));
    HIP_TRY(s, dalloc(&s->list[1], s->total_tiles));
    HIP_TRY(s, dalloc(&s->tile_lb, (size_t)s->total_tiles + 1));
    HIP_TRY(s, dalloc(&s->slist, s->total_tiles));
    if (!s->pool_cap) HIP_TRY(s, dalloc(&s->srec, (size_t)s->total_tiles * mpm::SREC_WORDS));
    s->part_cap = s->cap / 256 + s->cap / 64 + s->n_scenes + 64;
    s->max_schunks = s->cap + s->cap / 256 + s->n_scenes + 64;
    HIP_TRY(s, dalloc(&s->chunk_box, 2 * (size_t)s->part_cap));
    HIP_TRY(s, dalloc(&s->part, (size_t)s->part_cap * mpm::LDS_CELLS));
    HIP_TRY(s, dalloc(&s->schunks, s->max_schunks));
    HIP_TRY(s, dalloc(&s->chunk_of_slot, s->cap));
    HIP_TRY(s, dalloc(&s->cflags, s->cap));
    HIP_TRY(s, dalloc(&s->cexcl, s->cap));
    HIP_TRY(s, dalloc(&s->scan_scratch, mpm::scan_scratch_words(s->cap)));
    HIP_TRY(s, dalloc(&s->smark, s->total_tiles));
    HIP_TRY(s, dalloc(&s->spos, s->total_tiles));
    HIP_TRY(s, dalloc(&s->sscan, mpm::scan_scratch_words(s->total_tiles)));
    HIP_TRY(s, hipMemsetAsync(s->chunk_box, 0, 2 * (size_t)s->part_cap * sizeof(int4), s->stream));
    HIP_TRY(s, hipMemsetAsync(s->gmp, 0, cells * sizeof(float4), s->stream));
    HIP_TRY(s, hipMemsetAsync(s->gv, 0, cells * sizeof(float4), s->stream));
    HIP_TRY(s, hipMemsetAsync(s->tile_flag, 0, (size_t)s->total_tiles * 4, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    std::vector<uint32_t> ts(s->total_tiles);
    for (uint32_t i = 0; i < s->n_scenes; ++i) {
        uint32_t W = s->sh[i].W;
        std::fill(ts.begin() + s->sh[i].tile_base, ts.begin() + s->sh[i].tile_base + W * W * W, i);
    }
    HIP_TRY(s, hipMemcpy(s->tile_scene, ts.data(), ts.size() * 4, hipMemcpyHostToDevice));
    uint32_t z[4] = {0u, 0u, 0u, 0u};
    HIP_TRY(s, hipMemcpy(s->status + mpm::ST_COUNT0, z, sizeof(z), hipMemcpyHostToDevice));
    s->grid_dirty = false;
    s->sorted = false;
    return MPM_OK;
}

int clear_grid(mpm_sim* s) {
    if (!s->grid_dirty) return MPM_OK;
    if (s->dry) {
        s->grid_dirty = false;
        return MPM_OK;
    }
    HIP_TRY(s, mpm::launch_clear_tiles(griddev(s), s->list[s->list_cur], s->status + mpm::ST_COUNT0 + s->list_cur,
                                       grid_blocks(s), s->stream));
    s->grid_dirty = false;
    return MPM_OK;
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// keys[id] -> stable sort -> reorder particle planes into tile order.
int sort_particles(mpm_sim* s, bool reorder) {
    s->boxes_valid = false;  // the chunk list is rebuilt
    s->layout_gen++;
    if (s->dry) {
        if (reorder) {
            s->ps_cur = 1 - s->ps_cur;
            s->sorted = true;
            s->since_sort = 0;
            s->sorts++;
        }
        return MPM_OK;
    }
    Stage st(s, MPM_STAGE_SORT);
    mpm::StepArgs a = args(s, 0.f, s->list_cur);
    // The fused layout only needs particles grouped by tile: ties keep the
    // current slot order.  The deterministic mode and the key query keep the
    // reference order (ties by original id), which fixes its summation order.
    const bool by_slot = reorder && s->p2g_mode == MPM_P2G_FUSED;
    HIP_TRY(s, mpm::launch_tile_keys(a, s->keys, by_slot ? 1 : 0, s->stream));
    unsigned blocks = std::min<uint32_t>((s->n + 255) / 256, 8192u);
    hipLaunchKernelGGL(k_iota, dim3(std::max(blocks, 1u)), dim3(256), 0, s->stream, s->vals, s->n);
    HIP_TRY(s, hipGetLastError());
    const int bits = int(ceil_log2(std::max(s->total_tiles, 2u)));
    HIP_TRY(s, mpm::radix_sort_pairs(s->keys, s->vals, s->n, bits, s->sort_scratch, s->stream));
    if (reorder) {
        const int other = 1 - s->ps_cur;
        // The active list and the chunk list depend on the sorted keys only.
        HIP_TRY(s, mpm::launch_static_rebuild(s->d_scenes, s->tile_scene, s->tile_flag, s->slist,
                                              s->status + mpm::ST_SCOUNT, s->keys, s->n, s->tile_lb,
                                              s->total_tiles, s->smark, s->spos, s->sscan, s->stream));
        // several particles per thread only for the shader model's fused
        // layout (the differentiable model and its adjoint take one)
        s->chunk_ppt = (by_slot && s->sh[0].model == MPM_MODEL_SHADER) ? s->ppt : 1u;
        HIP_TRY(s, mpm::build_sorted_chunks(s->d_scenes, s->tile_scene, s->keys, s->tile_lb, s->n, s->cflags,
                                            s->cexcl, s->scan_scratch, s->schunks, s->status + mpm::ST_NCHUNKS,
                                            s->chunk_of_slot, CHUNK * s->chunk_ppt, s->seg_tiles, s->total_tiles,
                                            s->stream));
        s->srec_valid = s->srec != nullptr;
        if (s->srec)
            HIP_TRY(s, mpm::launch_static_records(s->d_scenes, s->tile_scene, s->slist, s->status + mpm::ST_SCOUNT,
                                                  s->total_tiles, s->tile_lb, s->chunk_of_slot, s->part_cap, s->srec,
                                                  s->stream));
        // previous slot of every new slot: the sorted slots spread over cells
        // inside each chunk, or the reference order's ids mapped to slots
        if (by_slot)
            HIP_TRY(s, mpm::launch_chunk_spread(s->d_scenes, s->schunks, s->status + mpm::ST_NCHUNKS,
                                                s->max_schunks, s->ps[s->ps_cur], s->cap, s->vals, s->src_slot,
                                                s->keys, CHUNK * s->chunk_ppt, s->stream));
        else
            HIP_TRY(s, mpm::launch_reorder_index(s->vals, s->slot_of_id, s->src_slot, s->n, s->stream));
        HIP_TRY(s, mpm::launch_reorder(s->ps[s->ps_cur], s->ps[other], s->cap, s->src_slot, s->n, s->stream));
        HIP_TRY(s, mpm::launch_index_slots(s->ps[other], s->cap, s->slot_of_id, s->n, s->stream));
        s->ps_cur = other;
        s->sorted = true;
        s->since_sort = 0;
        s->sorts++;
    }
    return MPM_OK;
}

// Sorted layout: the device-built chunk list, walked by a persistent grid.
void use_sorted_chunks(mpm_sim* s, mpm::StepArgs& a, int& blocks) {
    a.chunks = s->schunks;
    a.n_chunks_dev = s->status + mpm::ST_NCHUNKS;
    a.n_chunks = std::min<uint32_t>(s->max_schunks, s->n / 256 + s->n / 64 + s->n_scenes + 64);
    blocks = s->pblocks ? int(s->pblocks) : int(std::min<uint32_t>(a.n_chunks, 4096u));
}

// `ps` (optional) runs the pass on another set of state planes of the same
// slot order (the adjoint's recomputed pre-states); `aff_out` (optional)
// receives the scattered particles' P2G affine (the diff tape).
int particle_pass(mpm_sim* s, int mode, float dt, int list_out, float* ps = nullptr, float* aff_out = nullptr) {
    // a scattering pass over the sorted chunks leaves their boxes for the
    // next pass's gather; a plain G2P moves the particles away from them
    const bool sorted_chunks = s->p2g_mode == MPM_P2G_FUSED && s->sorted;
    const bool boxes_after = sorted_chunks && mode != mpm::MODE_G2P;
    if (s->dry) {
        s->boxes_valid = boxes_after;
        return MPM_OK;
    }
    Stage st(s, mode == mpm::MODE_P2G ? MPM_STAGE_P2G : (mode == mpm::MODE_G2P ? MPM_STAGE_G2P : MPM_STAGE_G2P2G));
    mpm::StepArgs a = args(s, dt, list_out);
    if (ps) a.ps = a.ps_out = ps;
    a.aff_out = aff_out;
    int blocks = particle_blocks(s);
    if (sorted_chunks) use_sorted_chunks(s, a, blocks);
    else a.box_gather = 0;
    s->boxes_valid = boxes_after;
    HIP_TRY(s, mpm::launch_particle_pass(mode, int(s->sh[0].model), a, sorted_chunks ? int(s->chunk_ppt) : 1, blocks,
                                         s->stream));
    return MPM_OK;
}

// `react`: this substep reports the rigid bodies' reaction (the last
// substep of a step: the reference resets the wrench every substep, so only
// the last one is observable).
int grid_update(mpm_sim* s, float dt, bool zero_after, bool react = false) {
    if (s->dry) return MPM_OK;
    const bool rigid = s->rigid_count > 0;
    Stage st(s, MPM_STAGE_GRID);
    const int nxt = 1 - s->list_cur;
    const mpm::GridDev g = griddev(s);
    uint32_t* cur_count = s->status + mpm::ST_COUNT0 + s->list_cur;
    const mpm::RigidDev rg{s->d_rig, rigid && react ? s->d_wrench : nullptr, s->rigid_count, s->prescribed ? 1u : 0u};
    if (rg.wrench) HIP_TRY(s, mpm::launch_zero_f64(s->d_wrench, 8 * s->rigid_count, s->stream));
    const int parts = s->p2g_mode == MPM_P2G_FUSED ? 1 : 0;
    if (s->jacobi_iters > 0) {
        // Semi-implicit step (mpm.rs:2515-2530): the sweeps need this substep's
        // tile flags, so the grid update keeps them and the clear follows.
        HIP_TRY(s, mpm::launch_grid_update(s->d_scenes, g, s->list[s->list_cur], cur_count, nullptr, dt, 0, parts,
                                           grid_blocks(s), s->stream, nullptr, rigid ? &rg : nullptr, grid_split(s)));
        HIP_TRY(s, mpm::launch_implicit_sweeps(s->d_scenes, g, s->gv2, s->jacobi_iters, s->list[s->list_cur],
                                               cur_count, dt, grid_blocks(s), s->stream, rigid ? &rg : nullptr));
        if (rg.wrench && !s->prescribed)
            HIP_TRY(s, mpm::launch_rigid_shell(s->d_scenes, g, s->list[s->list_cur], cur_count, dt, rg,
                                               grid_blocks(s), s->stream));
        if (zero_after) {
            HIP_TRY(s, mpm::launch_clear_tiles(g, s->list[s->list_cur], cur_count, grid_blocks(s), s->stream));
            HIP_TRY(s, mpm::launch_set_u32(s->status + mpm::ST_COUNT0 + nxt, 0u, s->stream));
        }
        return MPM_OK;
    }
    HIP_TRY(s, mpm::launch_grid_update(s->d_scenes, g, s->list[s->list_cur], cur_count,
                                       zero_after ? s->status + mpm::ST_COUNT0 + nxt : nullptr, dt,
                                       zero_after ? 1 : 0, parts, grid_blocks(s), s->stream, nullptr,
                                       rigid ? &rg : nullptr, grid_split(s)));
    // the shell force reads the updated velocities; the current list and its
    // count stay valid until the next particle pass
    if (rg.wrench && !s->prescribed)
        HIP_TRY(s, mpm::launch_rigid_shell(s->d_scenes, g, s->list[s->list_cur], cur_count, dt, rg, grid_blocks(s),
                                           s->stream));
    return MPM_OK;
}

int deterministic_p2g(mpm_sim* s, float dt) {
    if (s->dry) return MPM_OK;
    Stage st(s, MPM_STAGE_P2G);
    if (!s->affine) HIP_TRY(s, dalloc(&s->affine, (size_t)12 * s->cap));
    mpm::StepArgs a = args(s, dt, s->list_cur);
    HIP_TRY(s, mpm::launch_gather_prepare(a, s->affine, particle_blocks(s), s->stream));
    HIP_TRY(s, mpm::launch_gather_p2g(a, s->tile_scene, s->tile_lb, s->affine, s->slist,
                                      s->status + mpm::ST_SCOUNT, 4096, s->stream));
    return MPM_OK;
}

// Sparse pool, after each scattering pass: the legacy spheres' marks, slots
// for the marked tiles, then the global-atomic scatters that waited for them.
int pool_update(mpm_sim* s, float dt, int list_out) {
    if (!s->pool_cap || s->dry) return MPM_OK;
    if (s->rigid_count && !s->prescribed) {
        const mpm::RigidDev rg{s->d_rig, nullptr, s->rigid_count, 0u};
        HIP_TRY(s, mpm::launch_mark_spheres(s->d_scenes, rg, s->tile_mark, s->pool_cap, s->status, s->stream));
    }
    HIP_TRY(s, mpm::launch_pool_compact(s->tile_mark, s->total_tiles, s->tile_slot, s->slot_tile, s->pool_cap,
                                        s->status, s->stream));
    HIP_TRY(s, mpm::launch_deferred_p2g(args(s, dt, list_out), int(s->sh[0].model), s->stream));
    return MPM_OK;
}

// The substep sequence of one step_n call.
int run_step_n(mpm_sim* s, float dt, uint32_t substeps) {
    int rc;
    s->w_stepped = true;
    s->fault_reset = false;
    if (!s->dry && (rc = ensure_grid(s)) != MPM_OK) return rc;
    if ((rc = clear_grid(s)) != MPM_OK) return rc;
    const bool det = s->p2g_mode == MPM_P2G_DETERMINISTIC;
    const bool fused = s->p2g_mode == MPM_P2G_FUSED;
    if (det) {
        for (uint32_t k = 0; k < substeps; ++k) {
            const bool last = k + 1 == substeps;
            if ((rc = sort_particles(s, true)) != MPM_OK) return rc;
            if ((rc = deterministic_p2g(s, dt)) != MPM_OK) return rc;
            if ((rc = grid_update(s, dt, !last, last)) != MPM_OK) return rc;
            if ((rc = particle_pass(s, mpm::MODE_G2P, dt, 1 - s->list_cur)) != MPM_OK) return rc;
            if (!last) s->list_cur = 1 - s->list_cur;
        }
        s->grid_dirty = true;
        return MPM_OK;
    }
    if (fused && !s->hold_order && (!s->sorted || (s->sort_every == 0) || s->since_sort >= s->sort_every)) {
        if ((rc = sort_particles(s, true)) != MPM_OK) return rc;
    }
    // prologue P2G of the current state
    if ((rc = particle_pass(s, mpm::MODE_P2G, dt, s->list_cur)) != MPM_OK) return rc;
    if ((rc = pool_update(s, dt, s->list_cur)) != MPM_OK) return rc;
    for (uint32_t k = 0; k < substeps; ++k) {
        const bool last = k + 1 == substeps;
        if ((rc = grid_update(s, dt, !last, last)) != MPM_OK) return rc;
        if (fused) {
            if (!last && !s->hold_order && s->sort_every > 0 && s->since_sort >= s->sort_every) {
                if ((rc = sort_particles(s, true)) != MPM_OK) return rc;
            }
            if ((rc = particle_pass(s, last ? mpm::MODE_G2P : mpm::MODE_G2P2G, dt, 1 - s->list_cur)) != MPM_OK)
                return rc;
        } else {
            if ((rc = particle_pass(s, mpm::MODE_G2P, dt, 1 - s->list_cur)) != MPM_OK) return rc;
            if (!last && (rc = particle_pass(s, mpm::MODE_P2G, dt, 1 - s->list_cur)) != MPM_OK) return rc;
        }
        if (!last && (rc = pool_update(s, dt, 1 - s->list_cur)) != MPM_OK) return rc;
        s->since_sort++;
        if (!last) s->list_cur = 1 - s->list_cur;
    }
    s->grid_dirty = true;
    return MPM_OK;
}

int field_layout(int field, int* comp0, int* ncomp) {
    switch (field) {
        case MPM_FIELD_POSITION: *comp0 = mpm::CX; *ncomp = 3; return 0;
        case MPM_FIELD_VELOCITY: *comp0 = mpm::CV; *ncomp = 3; return 0;
        case MPM_FIELD_C: *comp0 = mpm::CC; *ncomp = 9; return 0;
        case MPM_FIELD_F: *comp0 = mpm::CF; *ncomp = 9; return 0;
        case MPM_FIELD_JP: *comp0 = mpm::CJP; *ncomp = 1; return 0;
        case MPM_FIELD_MASS: *comp0 = mpm::CMASS; *ncomp = 1; return 0;
        case MPM_FIELD_VOLUME: *comp0 = mpm::CVOL; *ncomp = 1; return 0;
        case MPM_FIELD_YOUNGS: *comp0 = mpm::CE; *ncomp = 1; return 0;
        case MPM_FIELD_MATERIAL: *comp0 = mpm::CMAT; *ncomp = 1; return 0;
        default: return -1;
    }
}

int build_chunks(mpm_sim* s) {
    std::vector<Chunk> ch;
    for (uint32_t sc = 0; sc < s->n_scenes; ++sc) {
        s->sh[sc].cbase = uint32_t(ch.size());
        for (uint32_t off = 0; off < s->counts[sc]; off += CHUNK) {
            Chunk c{};
            c.scene = sc;
            c.start = s->pbase[sc] + off;
            c.count = std::min(CHUNK, s->counts[sc] - off);
            ch.push_back(c);
        }
    }
    s->n_chunks = uint32_t(ch.size());
    if (!ch.empty())
        HIP_TRY(s, hipMemcpy(s->d_chunks, ch.data(), ch.size() * sizeof(Chunk), hipMemcpyHostToDevice));
    return MPM_OK;
}

int upload_scenes(mpm_sim* s) {
    std::vector<float> mats((size_t)s->n_scenes * MPM_MAX_MATERIALS * mpm::MAT_STRIDE, 0.f);
    for (uint32_t i = 0; i < s->n_scenes; ++i) {
        s->sh[i].pbase = s->pbase[i];
        s->sh[i].n = s->counts[i];
        for (uint32_t r = 0; r < s->cfg[i].n_materials; ++r)
            std::memcpy(&mats[((size_t)i * MPM_MAX_MATERIALS + r) * mpm::MAT_STRIDE], &s->cfg[i].materials[r],
                        sizeof(mpm_material));
    }
    HIP_TRY(s, hipMemcpy(s->d_scenes, s->sh.data(), s->sh.size() * sizeof(SceneDev), hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(s->d_mats, mats.data(), mats.size() * sizeof(float), hipMemcpyHostToDevice));
    return MPM_OK;
}

int create(const mpm_scene_config* cfgs, uint32_t n_scenes, mpm_sim** out) {
    if (!out) return MPM_ERR_INVALID;
    *out = nullptr;
    if (!cfgs || n_scenes == 0) return MPM_ERR_INVALID;
    mpm_sim* s = new mpm_sim();
    s->n_scenes = n_scenes;
    s->cfg.assign(cfgs, cfgs + n_scenes);
    uint64_t tiles = 0, cap = 0;
    for (uint32_t i = 0; i < n_scenes; ++i) {
        std::string v = validate(cfgs[i]);
        if (!v.empty()) {
            *out = s;
            return fail(s, MPM_ERR_INVALID, "scene " + std::to_string(i) + ": " + v);
        }
        if (cfgs[i].model != cfgs[0].model) {
            *out = s;
            return fail(s, MPM_ERR_INVALID, "all scenes of a batch must use the same model");
        }
        uint64_t W = (cfgs[i].grid_size + 3u) / 4u;
        s->sh.push_back(make_scene(cfgs[i], uint32_t(tiles), i * MPM_MAX_MATERIALS));
        tiles += W * W * W;
        s->pbase.push_back(uint32_t(cap));
        cap += cfgs[i].max_particles;
    }
    // kernels address a particle plane by a 32-bit byte offset (4 slot)
    if (tiles >= (1ull << 26) || cap >= (1ull << 30) || cap == 0) {
        *out = s;
        return fail(s, MPM_ERR_INVALID, "grid or particle capacity too large");
    }
    s->total_tiles = uint32_t(tiles);
    s->cap = uint32_t(cap);
    s->counts.assign(n_scenes, 0u);
    *out = s;
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) return hip_fail(s, e, "hipGetDevice");
    HIP_TRY(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    const size_t cells = (size_t)s->total_tiles * 64;
    HIP_TRY(s, dalloc(&s->d_scenes, n_scenes));
    HIP_TRY(s, dalloc(&s->d_mats, (size_t)n_scenes * MPM_MAX_MATERIALS * mpm::MAT_STRIDE));
    HIP_TRY(s, dalloc(&s->ps[0], (size_t)mpm::NCOMP * s->cap));
    HIP_TRY(s, dalloc(&s->ps[1], (size_t)mpm::NCOMP * s->cap));
    HIP_TRY(s, dalloc(&s->status, mpm::ST_WORDS));
    HIP_TRY(s, dalloc(&s->slot_of_id, s->cap));
    HIP_TRY(s, dalloc(&s->keys, s->cap));
    HIP_TRY(s, dalloc(&s->vals, s->cap));
    HIP_TRY(s, dalloc(&s->src_slot, s->cap));
    HIP_TRY(s, dalloc(&s->sort_scratch, mpm::radix_scratch_words(s->cap)));
    s->max_chunks = s->cap / CHUNK + n_scenes + 1;
    HIP_TRY(s, dalloc(&s->d_chunks, (size_t)s->max_chunks));
    HIP_TRY(s, dalloc(&s->staging, (size_t)9 * s->cap));
    HIP_TRY(s, hipMemsetAsync(s->status, 0, mpm::ST_WORDS * 4, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->ps[0], 0, (size_t)mpm::NCOMP * s->cap * 4, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    (void)cells;
    int rc = upload_scenes(s);
    if (rc != MPM_OK) return rc;
    return MPM_OK;
}

int set_particles(mpm_sim* s, const uint32_t* counts, const float* pos, const float* vel) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < s->n_scenes; ++i) {
        if (counts[i] > s->cfg[i].max_particles)
            return fail(s, MPM_ERR_INVALID, "set_particles: n=" + std::to_string(counts[i]) +
                                                " > max_particles=" + std::to_string(s->cfg[i].max_particles));
        total += counts[i];
    }
    if (total > 0 && !pos) return fail(s, MPM_ERR_INVALID, "positions must not be NULL");
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->n = uint32_t(total);
    s->counts.assign(counts, counts + s->n_scenes);
    s->pbase.clear();
    uint32_t acc = 0;
    for (uint32_t i = 0; i < s->n_scenes; ++i) {
        s->pbase.push_back(acc);
        acc += counts[i];
    }
    // host-built planes [0, n) of every component
    std::vector<float> plane(std::max<size_t>(s->n, 1));
    float* ps = s->ps[0];
    auto put = [&](int comp) -> hipError_t {
        return hipMemcpy(ps + (size_t)comp * s->cap, plane.data(), (size_t)s->n * 4, hipMemcpyHostToDevice);
    };
    for (int k = 0; k < 3; ++k) {
        for (uint32_t p = 0; p < s->n; ++p) plane[p] = pos[3 * (size_t)p + k];
        HIP_TRY(s, put(mpm::CX + k));
        for (uint32_t p = 0; p < s->n; ++p) plane[p] = vel ? vel[3 * (size_t)p + k] : 0.0f;
        HIP_TRY(s, put(mpm::CV + k));
    }
    std::fill(plane.begin(), plane.end(), 0.0f);
    for (int k = 0; k < 9; ++k) HIP_TRY(s, put(mpm::CC + k));
    HIP_TRY(s, put(mpm::CJP));
    HIP_TRY(s, put(mpm::CMAT));
    for (int k = 0; k < 9; ++k) {
        std::fill(plane.begin(), plane.end(), (k % 4 == 0) ? 1.0f : 0.0f);
        HIP_TRY(s, put(mpm::CF + k));
    }
    auto per_scene = [&](auto getter) {
        for (uint32_t sc = 0; sc < s->n_scenes; ++sc)
            for (uint32_t p = 0; p < counts[sc]; ++p) plane[s->pbase[sc] + p] = getter(s->cfg[sc]);
    };
    per_scene([](const mpm_scene_config& c) { return c.particle_mass; });
    HIP_TRY(s, put(mpm::CMASS));
    per_scene([](const mpm_scene_config& c) { return c.particle_volume; });
    HIP_TRY(s, put(mpm::CVOL));
    per_scene([](const mpm_scene_config& c) { return c.youngs; });
    HIP_TRY(s, put(mpm::CE));
    for (uint32_t p = 0; p < s->n; ++p) {
        uint32_t id = p;
        std::memcpy(&plane[p], &id, 4);
    }
    HIP_TRY(s, put(mpm::CID));
    s->ps_cur = 0;
    HIP_TRY(s, mpm::launch_index_slots(s->ps[0], s->cap, s->slot_of_id, s->n, s->stream));
    int rc = build_chunks(s);
    if (rc != MPM_OK) return rc;
    rc = upload_scenes(s);
    if (rc != MPM_OK) return rc;
    s->has_E = 0;
    s->sorted = false;
    s->diff_steps = 0;
    s->aux_live = false;
    drop_graphs(s);  // captured steps bake in the particle count and chunk table
    if (s->gmp) {
        rc = clear_grid(s);
        if (rc != MPM_OK) return rc;
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return MPM_OK;
}

// ---------------------------------------------------------------------------
// Differentiable path (mpm_cpu_diff.rs forward_with_tape :397-418,
// mpm_adjoint.rs backward :121-137).  The forward sorts once and then keeps
// that slot order for the whole run, storing the 24 state planes every
// `every` substeps.  The backward walks the segments last to first: it
// recomputes the segment's pre-states into the tape from its checkpoint and
// then reverses each substep over its recomputed grid.
int diff_check_models(mpm_sim* s) {
    for (uint32_t i = 0; i < s->n_scenes; ++i)
        if (s->cfg[i].model != MPM_MODEL_CPUDIFF)
            return fail(s, MPM_ERR_INVALID, "the differentiable path is defined for MPM_MODEL_CPUDIFF scenes");
    return MPM_OK;
}

// Room for `every` pre-states and grids.
int ensure_tape(mpm_sim* s, uint32_t every) {
    if (s->tape_len >= every) return MPM_OK;
    const size_t cells = (size_t)s->total_tiles * 64;
    s->tape_seg = -1;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->tape) HIP_TRY(s, hipFree(s->tape));
    if (s->gtape) HIP_TRY(s, hipFree(s->gtape));
    if (s->atape) HIP_TRY(s, hipFree(s->atape));
    s->tape = nullptr;
    s->gtape = nullptr;
    s->atape = nullptr;
    s->tape_len = 0;
    HIP_TRY(s, dalloc(&s->tape, (size_t)every * mpm::NCOMP * s->cap));
    HIP_TRY(s, dalloc(&s->gtape, (size_t)every * 2 * cells));
    HIP_TRY(s, dalloc(&s->atape, (size_t)every * 9 * s->cap));
    s->tape_len = every;
    return MPM_OK;
}

// The last segment of a forward, stepped through the tape: the launches of
// run_step_n in the fused mode with the order held, except that substep k
// reads its pre-state from tape slot k and writes the next one to slot k+1
// (the last G2P writes the live state), and its grid velocity and P2G total
// go to the grid tape.  The backward then reverses this segment without
// recomputing it; the live state, the counters and the grid left for
// read_grid_mass_momentum are those of run_step_n.
int record_segment(mpm_sim* s, float dt, uint32_t len) {
    s->w_stepped = true;
    s->fault_reset = false;
    int rc = clear_grid(s);
    if (rc != MPM_OK) return rc;
    const size_t st_words = (size_t)mpm::NCOMP * s->cap;
    const size_t cells = (size_t)s->total_tiles * 64;
    float* live = s->ps[s->ps_cur];
    HIP_TRY(s, hipMemcpyAsync(s->tape, live, (size_t)24 * s->cap * 4, hipMemcpyDeviceToDevice, s->stream));
    if ((rc = particle_pass(s, mpm::MODE_P2G, dt, s->list_cur, s->tape, s->atape)) != MPM_OK) return rc;
    const mpm::GridDev g = griddev(s);
    const int gb = grid_blocks(s);
    for (uint32_t k = 0; k < len; ++k) {
        const bool last = k + 1 == len;
        const int L = s->list_cur;
        mpm::GridDev gk = g;
        gk.gv = s->gtape + (size_t)2 * k * cells;
        HIP_TRY(s, mpm::launch_grid_update(s->d_scenes, gk, s->list[L], s->status + mpm::ST_COUNT0 + L,
                                           last ? nullptr : s->status + mpm::ST_COUNT0 + (1 - L), dt, last ? 0 : 1,
                                           1, gb, s->stream, gk.gv + cells, nullptr, grid_split(s)));
        mpm::StepArgs a = args(s, dt, 1 - L);
        a.ps = s->tape + k * st_words;
        a.ps_out = last ? live : a.ps + st_words;
        a.gv = gk.gv;
        if (!last) a.aff_out = s->atape + (size_t)(k + 1) * 9 * s->cap;  // substep k+1's P2G
        int blocks = 0;
        use_sorted_chunks(s, a, blocks);
        HIP_TRY(s, mpm::launch_particle_pass(last ? mpm::MODE_G2P : mpm::MODE_G2P2G, int(s->sh[0].model), a, 1,
                                             blocks, s->stream));
        s->since_sort++;
        if (!last) s->list_cur = 1 - L;
    }
    s->boxes_valid = false;
    s->grid_dirty = true;
    return MPM_OK;
}

int diff_forward(mpm_sim* s, float dt, uint32_t steps, uint32_t every) {
    if (!std::isfinite(dt) || !(dt > 0.0f)) return fail(s, MPM_ERR_INVALID, "dt must be finite and > 0");
    if (steps == 0 || every == 0) return fail(s, MPM_ERR_INVALID, "steps and checkpoint_every must be >= 1");
    if (s->n == 0) return fail(s, MPM_ERR_INVALID, "no particles");
    int rc = diff_check_models(s);
    if (rc != MPM_OK) return rc;
    if ((rc = ensure_grid(s)) != MPM_OK) return rc;
    const uint32_t nseg = (steps + every - 1) / every;
    const size_t ck = (size_t)24 * s->cap;
    while (s->ckpt.size() < nseg) {
        float* c = nullptr;
        HIP_TRY(s, dalloc(&c, ck));
        s->ckpt.push_back(c);
    }
    const uint32_t snap_words = 8 + 2 * (nseg + 1) + 4 * nseg;
    if (s->dsnap_len < snap_words) {
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        if (s->dsnap) HIP_TRY(s, hipFree(s->dsnap));
        s->dsnap = nullptr;
        s->dsnap_len = 0;
        HIP_TRY(s, dalloc(&s->dsnap, snap_words));
        s->dsnap_len = snap_words;
    }
    if ((rc = ensure_tape(s, every)) != MPM_OK) return rc;
    uint32_t* fsnap = s->dsnap + 8;
    const int mode = s->p2g_mode, wvc = s->write_vc;
    s->p2g_mode = MPM_P2G_FUSED;
    s->write_vc = 1;
    s->diff_steps = 0;
    s->tape_seg = -1;
    s->adj_valid = false;
    Stage span(s, MPM_STAGE_DIFF_FORWARD);
    hipError_t e = hipSuccess;
    rc = clear_grid(s);
    if (rc == MPM_OK) rc = sort_particles(s, true);
    s->hold_order = true;
    // planes 24.. (jp, mass, volume, E, material, id) do not change in the
    // CPU-diff model: every tape slot gets them once
    const size_t st_words = (size_t)mpm::NCOMP * s->cap;
    for (uint32_t k = 0; rc == MPM_OK && k < every; ++k) {
        e = hipMemcpyAsync(s->tape + k * st_words + (size_t)24 * s->cap, s->ps[s->ps_cur] + (size_t)24 * s->cap,
                           (size_t)(mpm::NCOMP - 24) * s->cap * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e != hipSuccess) rc = hip_fail(s, e, "tape copy");
    }
    // The escape and fallback counters at every segment boundary tell the
    // backward which segments it may reverse with the fused kernel, so it
    // needs no host round trip per segment (the recompute repeats them).
    if (rc == MPM_OK && (e = mpm::launch_snap_counters(s->status, fsnap, s->stream)) != hipSuccess)
        rc = hip_fail(s, e, "counter snapshot");
    for (uint32_t seg = 0; rc == MPM_OK && seg < nseg; ++seg) {
        const uint32_t len = std::min(every, steps - seg * every);
        e = hipMemcpyAsync(s->ckpt[seg], s->ps[s->ps_cur], ck * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e != hipSuccess) rc = hip_fail(s, e, "checkpoint copy");
        if (rc == MPM_OK) rc = seg + 1 < nseg ? run_step_n(s, dt, len) : record_segment(s, dt, len);
        if (rc == MPM_OK && (e = mpm::launch_snap_counters(s->status, fsnap + 2 * (seg + 1), s->stream)) != hipSuccess)
            rc = hip_fail(s, e, "counter snapshot");
    }
    span.mark_end();
    s->hold_order = false;
    s->p2g_mode = mode;
    s->write_vc = wvc;
    if (rc != MPM_OK) return rc;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    std::vector<uint32_t> snap(2 * (nseg + 1));
    HIP_TRY(s, hipMemcpy(snap.data(), fsnap, snap.size() * 4, hipMemcpyDeviceToHost));
    s->seg_fused.assign(nseg, 0);
    for (uint32_t seg = 0; seg < nseg; ++seg)
        s->seg_fused[seg] = snap[2 * seg] == snap[2 * seg + 2] && snap[2 * seg + 1] == snap[2 * seg + 3];
    s->diff_steps = steps;
    s->tape_seg = int(nseg) - 1;
    s->diff_every = every;
    s->diff_dt = dt;
    s->diff_sorts = s->sorts;
    s->diff_ps = s->ps_cur;
    s->diff_gen = s->layout_gen;
    return MPM_OK;
}

mpm::AdjArgs adj_args(mpm_sim* s, float dt, int list, float* ps, int& blocks) {
    mpm::AdjArgs q{};
    q.a = args(s, dt, list);
    q.a.ps = q.a.ps_out = ps;
    q.a.gmp = s->gadj;  // the reverse G2P's escaped particles accumulate here
    q.a.use_lds = 1;
    use_sorted_chunks(s, q.a, blocks);
    q.adj = s->adj;
    q.gadj = s->gadj;
    q.gml = s->gml;
    return q;
}

int ensure_pinned(mpm_sim* s, size_t bytes) {
    if (s->hpin_bytes >= bytes) return MPM_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->hpin) HIP_TRY(s, hipHostFree(s->hpin));
    s->hpin = nullptr;
    s->hpin_bytes = 0;
    HIP_TRY(s, hipHostMalloc(&s->hpin, bytes, hipHostMallocDefault));
    s->hpin_bytes = bytes;
    return MPM_OK;
}

constexpr int BACKWARD_REPLAN = -1;  // a segment's recompute escaped where its forward did not

// Recomputes segment `seg` from its checkpoint with the fused forward:
// substep k's pre-state goes to tape slot k and its grid velocity and P2G
// total to the grid tape.  L: the list the segment's first P2G fills; on
// return, the list its last grid update consumed.  `fused`: whether no
// particle escaped to the global atomics (read back when `checked`, else the
// forward's plan, with the counters snapshotted for the check at the end).
int recompute_segment(mpm_sim* s, uint32_t seg, uint32_t len, bool checked, uint32_t* bsnap, int& L, bool& fused) {
    const float dt = s->diff_dt;
    const size_t st_words = (size_t)mpm::NCOMP * s->cap;
    const size_t cells = (size_t)s->total_tiles * 64;
    const mpm::GridDev g = griddev(s);
    const int gb = grid_blocks(s);
    auto count = [&](int l) { return s->status + mpm::ST_COUNT0 + l; };
    uint32_t esc0 = 0, fb0 = 0;
    if (checked) {
        uint32_t w0[mpm::ST_WORDS];
        HIP_TRY(s, hipMemcpyAsync(w0, s->status, sizeof(w0), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        esc0 = w0[mpm::ST_ESCAPES];
        fb0 = w0[mpm::ST_FALLBACK];
    } else {
        HIP_TRY(s, mpm::launch_snap_counters(s->status, bsnap + 4 * seg, s->stream));
    }
    float* t0 = s->tape;
    HIP_TRY(s, hipMemcpyAsync(t0, s->ckpt[seg], (size_t)24 * s->cap * 4, hipMemcpyDeviceToDevice, s->stream));
    int rc = particle_pass(s, mpm::MODE_P2G, dt, L, t0, s->atape);
    if (rc != MPM_OK) return rc;
    for (uint32_t k = 0; k < len; ++k) {
        float* tk = s->tape + k * st_words;
        mpm::GridDev gk = g;
        gk.gv = s->gtape + (size_t)2 * k * cells;
        HIP_TRY(s, mpm::launch_grid_update(s->d_scenes, gk, s->list[L], count(L), count(1 - L), dt, 1, 1, gb,
                                           s->stream, gk.gv + cells, nullptr, grid_split(s)));
        if (k + 1 == len) break;
        mpm::StepArgs a = args(s, dt, 1 - L);
        a.ps = tk;
        a.ps_out = tk + st_words;  // state k+1 goes straight to the next tape slot
        a.gv = gk.gv;
        a.write_vc = 1;
        a.aff_out = s->atape + (size_t)(k + 1) * 9 * s->cap;  // and its P2G affine
        int blocks = 0;
        use_sorted_chunks(s, a, blocks);
        HIP_TRY(s, mpm::launch_particle_pass(mpm::MODE_G2P2G, int(s->sh[0].model), a, 1, blocks, s->stream));
        L = 1 - L;
    }
    if (checked) {
        uint32_t w1[mpm::ST_WORDS];
        HIP_TRY(s, hipMemcpyAsync(w1, s->status, sizeof(w1), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        fused = w1[mpm::ST_ESCAPES] == esc0 && w1[mpm::ST_FALLBACK] == fb0;
    } else {
        HIP_TRY(s, mpm::launch_snap_counters(s->status, bsnap + 4 * seg + 2, s->stream));
        fused = s->seg_fused[seg] != 0;
    }
    return MPM_OK;
}

// One reverse sweep.  `checked`: read the escape counters around every
// segment's recompute on the host (two synchronisations per segment) instead
// of using the plan the forward recorded.
int diff_backward_pass(mpm_sim* s, bool has_gv, bool checked, Stage& span) {
    const uint32_t cap = s->cap, every = s->diff_every, steps = s->diff_steps;
    const size_t plane = cap;
    const size_t cells = (size_t)s->total_tiles * 64;
    const uint32_t nseg = (steps + every - 1) / every;
    uint32_t* bsnap = s->dsnap + 8 + 2 * (nseg + 1);
    const int mode = s->p2g_mode;
    s->p2g_mode = MPM_P2G_FUSED;
    int rc = clear_grid(s);
    if (rc != MPM_OK) {
        s->p2g_mode = mode;
        return rc;
    }
    // the recompute's G2P passes must not count as completed gathers
    HIP_TRY(s, mpm::launch_copy_u32(s->status + mpm::ST_GATHERS, s->dsnap, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->adj, 0, (size_t)mpm::NADJ * plane * 4, s->stream));
    HIP_TRY(s, hipMemsetAsync(s->gml, 0, 2 * (size_t)s->max_schunks * sizeof(double), s->stream));
    HIP_TRY(s, mpm::launch_permute_in(s->adj, cap, mpm::AX, 3, s->slot_of_id, s->staging, s->n, s->stream));
    if (has_gv)
        HIP_TRY(s, mpm::launch_permute_in(s->adj, cap, mpm::AV, 3, s->slot_of_id, s->staging + 3 * (size_t)s->n,
                                          s->n, s->stream));
    const float dt = s->diff_dt;
    const size_t st_words = (size_t)mpm::NCOMP * plane;
    const mpm::GridDev g = griddev(s);
    const int gb = grid_blocks(s);
    int L = s->list_cur;
    auto count = [&](int l) { return s->status + mpm::ST_COUNT0 + l; };
    // the tape's planes 24.. hold the forward's constant planes
    int kept = -1;
    for (int seg = int(nseg) - 1; seg >= 0 && rc == MPM_OK; --seg) {
        const uint32_t len = std::min(every, steps - uint32_t(seg) * every);
        // No particle of the segment scattered with global atomics (its reverse
        // scatters use the same stencils): the reverse P2G of substep k and the
        // reverse G2P of substep k-1 run as one pass.
        bool fused;
        if (seg == s->tape_seg) {
            // the forward stepped this segment through the tape: its last grid
            // update consumed list L and left the other list's count behind
            kept = seg;
            HIP_TRY(s, mpm::launch_set_u32(count(1 - L), 0u, s->stream));
            fused = s->seg_fused[seg] != 0;
        } else {
            s->tape_seg = -1;  // the recompute overwrites the tape
            if ((rc = recompute_segment(s, uint32_t(seg), len, checked, bsnap, L, fused)) != MPM_OK) break;
        }
        // The last grid update reset the other list; the reverse builds its
        // own list of escaped tiles there.
        const int R = 1 - L;
        auto tape_at = [&](int k) { return s->tape + (size_t)k * st_words; };
        auto gv_at = [&](int k) { return s->gtape + (size_t)2 * k * cells; };
        auto aff_at = [&](int k) { return s->atape + (size_t)k * 9 * cap; };
        if (fused) {
            int blocks = 0;
            mpm::AdjArgs q = adj_args(s, dt, R, tape_at(int(len) - 1), blocks);
            q.a.gv = gv_at(int(len) - 1);
            HIP_TRY(s, mpm::launch_adj_g2p(q, blocks, s->stream));
            for (int k = int(len) - 1; k >= 0; --k) {
                mpm::GridDev gk = g;
                gk.gmp = gv_at(k) + cells;
                HIP_TRY(s, mpm::launch_adj_grid(s->d_scenes, gk, s->gadj, s->list[R], count(R), dt, 0, gb,
                                                s->stream, grid_split(s)));
                mpm::AdjArgs qk = adj_args(s, dt, R, tape_at(k), blocks);
                qk.aff = aff_at(k);
                if (k > 0)
                    HIP_TRY(s, mpm::launch_adj_fused(qk, tape_at(k - 1), gv_at(k - 1), blocks, s->stream));
                else
                    HIP_TRY(s, mpm::launch_adj_p2g(qk, blocks, s->stream));
            }
            mpm::GridDev gc = g;
            gc.gmp = s->gadj;
            HIP_TRY(s, mpm::launch_clear_tiles(gc, s->list[R], count(R), gb, s->stream));
            L = R;
            continue;
        }
        for (int k = int(len) - 1; k >= 0 && rc == MPM_OK; --k) {
            float* tk = tape_at(k);
            float4* gvk = gv_at(k);
            int blocks = 0;
            mpm::AdjArgs q = adj_args(s, dt, R, tk, blocks);
            q.a.gv = gvk;
            q.aff = aff_at(k);
            HIP_TRY(s, mpm::launch_adj_g2p(q, blocks, s->stream));
            mpm::GridDev gk = g;
            gk.gmp = gvk + cells;
            HIP_TRY(s, mpm::launch_adj_grid(s->d_scenes, gk, s->gadj, s->list[R], count(R), dt, 1, gb, s->stream,
                                            grid_split(s)));
            HIP_TRY(s, mpm::launch_adj_p2g(q, blocks, s->stream));
            mpm::GridDev gc = g;
            gc.gmp = s->gadj;
            HIP_TRY(s, mpm::launch_clear_tiles(gc, s->list[R], count(R), gb, s->stream));
        }
        L = R;
    }
    span.mark_end();
    s->list_cur = L;
    s->grid_dirty = false;
    s->p2g_mode = mode;
    s->boxes_valid = false;  // the reverse passes left their own boxes
    if (rc != MPM_OK) return rc;
    HIP_TRY(s, mpm::launch_copy_u32(s->dsnap, s->status + mpm::ST_GATHERS, s->stream));
    if (!checked)
        HIP_TRY(s, hipMemcpyAsync(s->hpin, bsnap, (size_t)16 * nseg, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (!checked) {
        const uint32_t* w = static_cast<const uint32_t*>(s->hpin);
        for (uint32_t seg = 0; seg < nseg; ++seg)
            if (int(seg) != kept && s->seg_fused[seg] &&
                (w[4 * seg] != w[4 * seg + 2] || w[4 * seg + 1] != w[4 * seg + 3]))
                return BACKWARD_REPLAN;
    }
    return MPM_OK;
}

// `device`: gx and gv are device pointers (the caller's HBM), else host arrays.
int diff_backward(mpm_sim* s, const float* gx, const float* gv, bool device = false) {
    if (s->diff_steps == 0) return fail(s, MPM_ERR_INVALID, "mpm_diff_forward must run first");
    if (s->layout_gen != s->diff_gen || s->sorts != s->diff_sorts || s->ps_cur != s->diff_ps)
        return fail(s, MPM_ERR_INVALID,
                    "the particle order or sort state changed since mpm_diff_forward (a step, sort, key query "
                    "or autotune ran in between)");
    if (!gx) return fail(s, MPM_ERR_INVALID, "grad_x_final must not be NULL");
    const uint32_t every = s->diff_every, steps = s->diff_steps;
    const uint32_t nseg = (steps + every - 1) / every;
    const size_t cells = (size_t)s->total_tiles * 64;
    if (!s->adj) HIP_TRY(s, dalloc(&s->adj, (size_t)mpm::NADJ * s->cap));
    if (!s->gml) HIP_TRY(s, dalloc(&s->gml, 2 * (size_t)s->max_schunks));
    if (!s->gadj) {
        HIP_TRY(s, dalloc(&s->gadj, cells));
        HIP_TRY(s, hipMemsetAsync(s->gadj, 0, cells * sizeof(float4), s->stream));
    }
    // pinned staging: the final-state gradients in, the initial-state ones
    // (x, v, E and the (mu, lambda) pair) out, the counter snapshots
    const size_t pin = std::max<size_t>({(size_t)s->n * 24, (size_t)s->n * 28 + 64, (size_t)16 * nseg});
    int rc = ensure_pinned(s, pin);
    if (rc != MPM_OK) return rc;
    s->adj_valid = false;
    // the final-state gradients to the staging buffer, in the caller's order
    const size_t vec = (size_t)s->n * 12;
    if (device) {
        HIP_TRY(s, hipMemcpyAsync(s->staging, gx, vec, hipMemcpyDeviceToDevice, s->stream));
        if (gv)
            HIP_TRY(s, hipMemcpyAsync(s->staging + 3 * (size_t)s->n, gv, vec, hipMemcpyDeviceToDevice, s->stream));
    } else {
        std::memcpy(s->hpin, gx, vec);
        if (gv) std::memcpy(static_cast<char*>(s->hpin) + vec, gv, vec);
        HIP_TRY(s, hipMemcpyAsync(s->staging, s->hpin, gv ? 2 * vec : vec, hipMemcpyHostToDevice, s->stream));
    }
    Stage span(s, MPM_STAGE_DIFF_BACKWARD);
    rc = diff_backward_pass(s, gv != nullptr, false, span);
    if (rc == BACKWARD_REPLAN) rc = diff_backward_pass(s, gv != nullptr, true, span);
    if (rc != MPM_OK) return rc;
    s->adj_valid = true;
    return MPM_OK;
}

int adj_layout(int field, int* comp0, int* ncomp) {
    switch (field) {
        case MPM_FIELD_POSITION: *comp0 = mpm::AX; *ncomp = 3; return 0;
        case MPM_FIELD_VELOCITY: *comp0 = mpm::AV; *ncomp = 3; return 0;
        case MPM_FIELD_C: *comp0 = mpm::AC; *ncomp = 9; return 0;
        case MPM_FIELD_F: *comp0 = mpm::AF; *ncomp = 9; return 0;
        case MPM_FIELD_YOUNGS: *comp0 = mpm::AE; *ncomp = 1; return 0;
        default: return -1;
    }
}

int read_adjoint(mpm_sim* s, int field, float* out) {
    int c0, nc;
    if (adj_layout(field, &c0, &nc) != 0) return fail(s, MPM_ERR_INVALID, "no adjoint for this field");
    HIP_TRY(s, mpm::launch_unpermute(s->adj, s->cap, c0, nc, s->slot_of_id, s->staging, s->n, s->stream));
    HIP_TRY(s, hipMemcpyAsync(out, s->staging, (size_t)s->n * nc * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return MPM_OK;
}

// The outputs of a backward in one transfer: the requested adjoint planes in
// original order, then the fixed-order sum of the per-chunk (mu, lambda).
struct Out3 {
    float* dst;
    int comp, nc;
};

int read_backward_outputs(mpm_sim* s, float* gx0, float* gv0, float* gE, double* gml) {
    const size_t n = s->n;
    const Out3 outs[3] = {{gx0, mpm::AX, 3}, {gv0, mpm::AV, 3}, {gE, mpm::AE, 1}};
    size_t off = 0;
    for (const Out3& o : outs) {
        if (!o.dst) continue;
        HIP_TRY(s, mpm::launch_unpermute(s->adj, s->cap, o.comp, o.nc, s->slot_of_id, s->staging + off, s->n,
                                         s->stream));
        off += n * o.nc;
    }
    const size_t dbl = (off * 4 + 15) / 16 * 16;  // byte offset of the pair in the pinned buffer
    char* pin = static_cast<char*>(s->hpin);
    if (off) HIP_TRY(s, hipMemcpyAsync(pin, s->staging, off * 4, hipMemcpyDeviceToHost, s->stream));
    if (gml) {
        double* pair = reinterpret_cast<double*>(s->dsnap + 4);
        HIP_TRY(s, mpm::launch_sum_pairs(s->gml, s->status + mpm::ST_NCHUNKS, s->max_schunks, pair, s->stream));
        HIP_TRY(s, hipMemcpyAsync(pin + dbl, pair, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    off = 0;
    for (const Out3& o : outs) {
        if (!o.dst) continue;
        std::memcpy(o.dst, pin + off * 4, n * o.nc * 4);
        off += n * o.nc;
    }
    if (gml) std::memcpy(gml, pin + dbl, 2 * sizeof(double));
    return MPM_OK;
}

// The outputs of a backward straight into the caller's device buffers.
int read_backward_outputs_device(mpm_sim* s, float* gx0, float* gv0, float* gE, double* gml) {
    const Out3 outs[3] = {{gx0, mpm::AX, 3}, {gv0, mpm::AV, 3}, {gE, mpm::AE, 1}};
    for (const Out3& o : outs)
        if (o.dst)
            HIP_TRY(s, mpm::launch_unpermute(s->adj, s->cap, o.comp, o.nc, s->slot_of_id, o.dst, s->n, s->stream));
    if (gml) HIP_TRY(s, mpm::launch_sum_pairs(s->gml, s->status + mpm::ST_NCHUNKS, s->max_schunks, gml, s->stream));
    return MPM_OK;
}

// ---------------------------------------------------------------------------
// Launch autotuner: coordinate search over particles per thread x segment
// length (shader model), the persistent particle-pass grid, the grid-update
// grid and the sort interval, timed with HIP events on the context's stream.
// The particle state is backed up first and restored afterwards (unsorted, so
// the next step re-sorts).
int autotune(mpm_sim* s, float dt, uint32_t trial, uint32_t* out_pb, uint32_t* out_gb, uint32_t* out_se,
             uint32_t* out_ppt = nullptr, uint32_t* out_seg = nullptr) {
    if (!std::isfinite(dt) || dt < 0.0f) return fail(s, MPM_ERR_INVALID, "dt must be finite and >= 0");
    if (trial == 0) return fail(s, MPM_ERR_INVALID, "trial_substeps must be >= 1");
    if (s->n == 0) return MPM_OK;
    int rc = ensure_grid(s);
    if (rc != MPM_OK) return rc;
    const size_t words = (size_t)mpm::NCOMP * s->cap;
    float* backup = nullptr;
    HIP_TRY(s, dalloc(&backup, words));
    HIP_TRY(s, hipMemcpyAsync(backup, s->ps[s->ps_cur], words * 4, hipMemcpyDeviceToDevice, s->stream));
    uint32_t st[mpm::ST_WORDS];
    HIP_TRY(s, hipMemcpyAsync(st, s->status, sizeof(st), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    const uint32_t sorts0 = s->sorts;
    // host flags the trials' steps set that belong to the caller's state: a
    // stepped handle reads position/velocity w as 0, and a reset of a latched
    // fault blocks an archive until the caller steps or restores
    const bool w_stepped0 = s->w_stepped, fault_reset0 = s->fault_reset;
    const int prof = s->profiling;
    s->profiling = 0;
    hipEvent_t ea = nullptr, eb = nullptr;
    auto restore = [&]() -> int {
        int r = clear_grid(s);
        if (r != MPM_OK) return r;
        HIP_TRY(s, hipMemcpyAsync(s->ps[0], backup, words * 4, hipMemcpyDeviceToDevice, s->stream));
        s->ps_cur = 0;
        HIP_TRY(s, mpm::launch_index_slots(s->ps[0], s->cap, s->slot_of_id, s->n, s->stream));
        s->sorted = false;
        return MPM_OK;
    };
    // ms: per substep over `len` timed substeps after a 2-substep warm-up
    uint32_t try_ppt = s->ppt, try_seg = s->seg_tiles;
    auto run_trial = [&](uint32_t pb, uint32_t gb, uint32_t se, uint32_t len, float& ms) -> int {
        int r = restore();
        if (r != MPM_OK) return r;
        s->pblocks = pb;
        s->gblocks = gb;
        s->sort_every = se;
        s->ppt = try_ppt;
        s->seg_tiles = try_seg;
        if ((r = run_step_n(s, dt, 2)) != MPM_OK) return r;
        HIP_TRY(s, hipEventRecord(ea, s->stream));
        if ((r = run_step_n(s, dt, len)) != MPM_OK) return r;
        HIP_TRY(s, hipEventRecord(eb, s->stream));
        HIP_TRY(s, hipEventSynchronize(eb));
        HIP_TRY(s, hipEventElapsedTime(&ms, ea, eb));
        ms /= float(len);
        return MPM_OK;
    };
    uint32_t best_pb = s->pblocks, best_gb = s->gblocks, best_se = s->sort_every;
    uint32_t best_ppt = s->ppt, best_seg = s->seg_tiles;
    if (timing_event(&ea) != hipSuccess || timing_event(&eb) != hipSuccess) rc = fail(s, MPM_ERR_DEVICE, "event");
    float best = 1e30f, ms = 0.f;
    const uint32_t pbs[] = {0u, 1024u, 2048u, 4096u, 8192u, 16384u};
    const uint32_t gbs[] = {512u, 1024u, 2048u, 4096u, 8192u};
    const uint32_t ses[] = {8u, 16u, 32u, 64u, 128u};
    const uint32_t gb0 = best_gb, se0 = best_se ? best_se : 16u;
    if (s->sh[0].model == MPM_MODEL_SHADER && s->p2g_mode == MPM_P2G_FUSED) {
        for (uint32_t pp : {1u, 2u})
            for (uint32_t sg : {4u, 6u, 8u}) {
                if (rc != MPM_OK) break;
                try_ppt = pp;
                try_seg = sg;
                if ((rc = run_trial(best_pb, gb0, se0, trial, ms)) == MPM_OK && ms < best) {
                    best = ms;
                    best_ppt = pp;
                    best_seg = sg;
                }
            }
        try_ppt = best_ppt;
        try_seg = best_seg;
    }
    best = 1e30f;
    for (uint32_t pb : pbs) {
        if (rc != MPM_OK) break;
        if ((rc = run_trial(pb, gb0, se0, trial, ms)) == MPM_OK && ms < best) {
            best = ms;
            best_pb = pb;
        }
    }
    best = 1e30f;
    for (uint32_t gb : gbs) {
        if (rc != MPM_OK) break;
        if ((rc = run_trial(best_pb, gb, se0, trial, ms)) == MPM_OK && ms < best) {
            best = ms;
            best_gb = gb;
        }
    }
    best = 1e30f;
    for (uint32_t se : ses) {
        if (rc != MPM_OK) break;
        // two whole intervals, so each candidate pays its sorts at its own rate
        if ((rc = run_trial(best_pb, best_gb, se, std::max(trial, 2 * se), ms)) == MPM_OK && ms < best) {
            best = ms;
            best_se = se;
        }
    }
    int rr = restore();
    if (rc == MPM_OK) rc = rr;
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
    s->profiling = prof;
    s->pblocks = best_pb;
    s->gblocks = best_gb;
    s->sort_every = best_se;
    s->ppt = best_ppt;
    s->seg_tiles = best_seg;
    s->sorts = sorts0;
    s->w_stepped = w_stepped0;
    s->fault_reset = fault_reset0;
    // the trials rebuilt keys, chunks and the slot order: a forward taped
    // before this call cannot be reversed (layout_gen stays advanced)
    s->layout_gen++;
    // counters of the trials are not the caller's: put back the fault latch,
    // gathers, escapes and fallback counters
    const int keep[] = {mpm::ST_FAULT, mpm::ST_GATHERS, mpm::ST_ESCAPES, mpm::ST_FALLBACK};
    for (int w : keep) HIP_TRY(s, hipMemcpyAsync(s->status + w, &st[w], 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipFree(backup));
    if (rc != MPM_OK) return rc;
    if (out_pb) *out_pb = best_pb;
    if (out_gb) *out_gb = best_gb;
    if (out_se) *out_se = best_se;
    if (out_ppt) *out_ppt = best_ppt;
    if (out_seg) *out_seg = best_seg;
    return MPM_OK;
}

}  // namespace

// ---- key query without allocation (mpm_sort_keys_noalloc) ------------------
//
// The radix sort of mpm_sort_keys_noalloc works in `scratch`, a static device
// array sized by radix_scratch_words(MPM_NOALLOC_MAX_PARTICLES): ping-pong
// keys and ids (2 n words), the per-block digit histogram and its scan
// (2 x 256 x n/4096 words) and the scan's block sums.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code `scratch`
// names a host-side stand-in; its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch), and only that
// address is ever handed to a kernel or a copy.
constexpr size_t NOALLOC_HIST = 256u * ((MPM_NOALLOC_MAX_PARTICLES + 4095u) / 4096u);
constexpr size_t NOALLOC_WORDS =
    2u * size_t(MPM_NOALLOC_MAX_PARTICLES) + 2u * NOALLOC_HIST + NOALLOC_HIST / 1024u + 1024u + 256u;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {

constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
uint32_t* noalloc_ptr[NOALLOC_DEVICES] = {};

// Device address of `scratch` on the handle's device (current when called).
int noalloc_scratch(mpm_sim* s, uint32_t** out) {
    if (s->device < 0 || s->device >= NOALLOC_DEVICES)
        return fail(s, MPM_ERR_INVALID, "mpm_sort_keys_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[s->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(s, e, "hipGetSymbolAddress(scratch)");
        noalloc_ptr[s->device] = static_cast<uint32_t*>(p);
    }
    *out = noalloc_ptr[s->device];
    return MPM_OK;
}

// Scene-local tile key and id of every particle of one scene (the scene's
// slots are [pbase, pbase + cnt) in any layout), written at the local id so a
// stable sort over ids 0..cnt-1 gives the reference order.  Same key formula
// as k_tile_keys (mpm_particle_order.wgsl:13-25), without the tile base.
__global__ void k_scene_keys(const float* __restrict__ ps, uint32_t cap, const SceneDev* __restrict__ scenes,
                             uint32_t scene, uint32_t pbase, uint32_t cnt, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ ids) {
    const SceneDev& sc = scenes[scene];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
        const uint32_t slot = pbase + j;
        uint32_t t[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = floorf((ps[(size_t)(mpm::CX + k) * cap + slot] - sc.origin[k]) * sc.inv_dx);
            const int ci = min(max(int(c) - sc.window[k], 0), int(sc.G) - 1);
            t[k] = uint32_t(ci) >> 2;
        }
        const uint32_t id = __float_as_uint(ps[(size_t)mpm::CID * cap + slot]) - pbase;
        keys[id] = t[0] + sc.W * (t[1] + sc.W * t[2]);
        ids[id] = id;
    }
}

// Padding of the key query beyond the scene's particles (keys 0xffffffff,
// identity ids), as the ParticleOrder encoding pads to its capacity.
__global__ void k_pad_keys(uint32_t* __restrict__ keys, uint32_t* __restrict__ ids, uint32_t from, uint32_t to) {
    for (uint32_t i = from + blockIdx.x * blockDim.x + threadIdx.x; i < to; i += gridDim.x * blockDim.x) {
        keys[i] = 0xffffffffu;
        ids[i] = i;
    }
}

}  // namespace

extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

int mpm_create(const mpm_scene_config* cfg, mpm_sim** out) { return create(cfg, 1, out); }

int mpm_create_batch(const mpm_scene_config* cfgs, uint32_t n_scenes, mpm_sim** out) {
    return create(cfgs, n_scenes, out);
}

int mpm_destroy(mpm_sim* sim) {
    if (!sim) return MPM_ERR_INVALID;
    release(sim);
    delete sim;
    return MPM_OK;
}

const char* mpm_last_error(const mpm_sim* sim) {
    if (!sim) return "null handle";
    return sim->err.c_str();
}

int mpm_set_particles(mpm_sim* sim, const float* positions, const float* velocities, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    if (sim->n_scenes != 1) return fail(sim, MPM_ERR_INVALID, "use mpm_set_particles_batch for batches");
    return set_particles(sim, &n, positions, velocities);
}

int mpm_set_particles_batch(mpm_sim* sim, const uint32_t* counts, const float* positions,
                            const float* velocities) {
    if (!sim || !counts) return MPM_ERR_INVALID;
    return set_particles(sim, counts, positions, velocities);
}

int mpm_write_field(mpm_sim* sim, int field, const void* data, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int c0, nc;
    if (field_layout(field, &c0, &nc) != 0) return fail(sim, MPM_ERR_INVALID, "unknown field");
    if (n != sim->n) return fail(sim, MPM_ERR_INVALID, "field length must equal the particle count");
    if (n == 0) return MPM_OK;
    if (!data) return fail(sim, MPM_ERR_INVALID, "data must not be NULL");
    HIP_TRY(sim, hipMemcpyAsync(sim->staging, data, (size_t)n * nc * 4, hipMemcpyHostToDevice, sim->stream));
    HIP_TRY(sim, mpm::launch_permute_in(sim->ps[sim->ps_cur], sim->cap, c0, nc, sim->slot_of_id, sim->staging, n,
                                        sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    if (field == MPM_FIELD_YOUNGS) sim->has_E = 1;
    if (field == MPM_FIELD_POSITION) sim->sorted = false;
    sim->diff_steps = 0;
    return MPM_OK;
}

int mpm_read_field(mpm_sim* sim, int field, void* out, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int c0, nc;
    if (field_layout(field, &c0, &nc) != 0) return fail(sim, MPM_ERR_INVALID, "unknown field");
    if (n != sim->n) return fail(sim, MPM_ERR_INVALID, "output length must equal the particle count");
    if (n == 0) return MPM_OK;
    if (!out) return fail(sim, MPM_ERR_INVALID, "out must not be NULL");
    HIP_TRY(sim, mpm::launch_unpermute(sim->ps[sim->ps_cur], sim->cap, c0, nc, sim->slot_of_id, sim->staging, n,
                                       sim->stream));
    HIP_TRY(sim, hipMemcpyAsync(out, sim->staging, (size_t)n * nc * 4, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    flush_profile(sim);
    return MPM_OK;
}

int mpm_read_field_device(mpm_sim* sim, int field, void* out_dev, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int c0, nc;
    if (field_layout(field, &c0, &nc) != 0) return fail(sim, MPM_ERR_INVALID, "unknown field");
    if (n != sim->n) return fail(sim, MPM_ERR_INVALID, "output length must equal the particle count");
    if (n == 0) return MPM_OK;
    if (!out_dev) return fail(sim, MPM_ERR_INVALID, "out must not be NULL");
    HIP_TRY(sim, mpm::launch_unpermute(sim->ps[sim->ps_cur], sim->cap, c0, nc, sim->slot_of_id,
                                       static_cast<float*>(out_dev), n, sim->stream));
    return MPM_OK;
}

int mpm_read_positions(mpm_sim* sim, float* out, uint32_t n) {
    return mpm_read_field(sim, MPM_FIELD_POSITION, out, n);
}
int mpm_read_velocities(mpm_sim* sim, float* out, uint32_t n) {
    return mpm_read_field(sim, MPM_FIELD_VELOCITY, out, n);
}

int mpm_set_materials(mpm_sim* sim, uint32_t scene, const mpm_material* rows, uint32_t n_rows) {
    if (!sim) return MPM_ERR_INVALID;
    if (scene >= sim->n_scenes) return fail(sim, MPM_ERR_INVALID, "scene out of range");
    if (!rows || n_rows < 1 || n_rows > MPM_MAX_MATERIALS)
        return fail(sim, MPM_ERR_INVALID, "materials: 1..8 rows required");
    for (uint32_t r = 0; r < n_rows; ++r) sim->cfg[scene].materials[r] = rows[r];
    sim->cfg[scene].n_materials = n_rows;
    sim->sh[scene].n_mat = n_rows;
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    return upload_scenes(sim);
}

int mpm_set_scene_params(mpm_sim* sim, uint32_t scene, const mpm_scene_config* cfg) {
    if (!sim || !cfg) return MPM_ERR_INVALID;
    if (scene >= sim->n_scenes) return fail(sim, MPM_ERR_INVALID, "scene out of range");
    std::string v = validate(*cfg);
    if (!v.empty()) return fail(sim, MPM_ERR_INVALID, v);
    const mpm_scene_config& old = sim->cfg[scene];
    if (cfg->grid_size != old.grid_size || cfg->max_particles != old.max_particles || cfg->model != old.model)
        return fail(sim, MPM_ERR_INVALID, "grid_size, max_particles and model are fixed at creation");
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    SceneDev d = make_scene(*cfg, sim->sh[scene].tile_base, sim->sh[scene].mat_base);
    sim->cfg[scene] = *cfg;
    sim->sh[scene] = d;
    for (auto& kv : sim->graphs) (void)hipGraphExecDestroy(kv.second);
    sim->graphs.clear();
    sim->sorted = false;
    return upload_scenes(sim);
}

int mpm_set_p2g_mode(mpm_sim* sim, int mode) {
    if (!sim) return MPM_ERR_INVALID;
    if (mode < 0 || mode > 2) return fail(sim, MPM_ERR_INVALID, "unknown p2g mode");
    if (mode == MPM_P2G_DETERMINISTIC && sim->sh[0].model != MPM_MODEL_SHADER)
        return fail(sim, MPM_ERR_INVALID, "deterministic P2G is defined for the shader model");
    if (mode == MPM_P2G_DETERMINISTIC && sim->pool_cap)  // mpm.rs:1265-1268
        return fail(sim, MPM_ERR_INVALID, "deterministic P2G needs the dense grid; this handle has a sparse tile pool");
    sim->p2g_mode = mode;
    return MPM_OK;
}

int mpm_set_time_step(mpm_sim* sim, uint32_t jacobi_iters) {
    if (!sim) return MPM_ERR_INVALID;
    if (jacobi_iters > 0)
        for (uint32_t i = 0; i < sim->n_scenes; ++i)
            if (sim->cfg[i].model != MPM_MODEL_SHADER)
                return fail(sim, MPM_ERR_INVALID, "the semi-implicit step is defined for MPM_MODEL_SHADER scenes");
    if (jacobi_iters > 1024) return fail(sim, MPM_ERR_INVALID, "jacobi_iters must be <= 1024");
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    if (jacobi_iters > 0 && sim->gmp && !sim->gv2) HIP_TRY(sim, dalloc(&sim->gv2, grid_cells(sim)));
    sim->jacobi_iters = jacobi_iters;
    for (auto& kv : sim->graphs) (void)hipGraphExecDestroy(kv.second);
    sim->graphs.clear();
    return MPM_OK;
}

int mpm_set_sort_interval(mpm_sim* sim, uint32_t every) {
    if (!sim) return MPM_ERR_INVALID;
    sim->sort_every = every;
    return MPM_OK;
}

int mpm_set_write_vc(mpm_sim* sim, int every_substep) {
    if (!sim) return MPM_ERR_INVALID;
    sim->write_vc = every_substep ? 1 : 0;
    return MPM_OK;
}

int mpm_set_launch(mpm_sim* sim, uint32_t particle_blocks, uint32_t grid_blocks_) {
    if (!sim) return MPM_ERR_INVALID;
    sim->pblocks = particle_blocks;
    sim->gblocks = grid_blocks_;
    return MPM_OK;
}

int mpm_set_launch_config(mpm_sim* sim, const mpm_launch_config* cfg) {
    if (!sim) return MPM_ERR_INVALID;
    if (!cfg) return fail(sim, MPM_ERR_INVALID, "cfg must not be NULL");
    if (cfg->particles_per_thread < 1 || cfg->particles_per_thread > uint32_t(mpm::MAX_PPT))
        return fail(sim, MPM_ERR_INVALID, "particles_per_thread must be 1 or 2");
    if (cfg->segment_tiles != 4 && cfg->segment_tiles != 6 && cfg->segment_tiles != 8)
        return fail(sim, MPM_ERR_INVALID, "segment_tiles must be 4, 6 or 8");
    sim->pblocks = cfg->particle_blocks;
    sim->gblocks = cfg->grid_blocks;
    sim->sort_every = cfg->sort_every;
    if (cfg->particles_per_thread != sim->ppt || cfg->segment_tiles != sim->seg_tiles) {
        // a new chunk list at the next step (a taped forward is not reversible after it)
        sim->ppt = cfg->particles_per_thread;
        sim->seg_tiles = cfg->segment_tiles;
        if (sim->sorted) {
            sim->sorted = false;
            sim->boxes_valid = false;
            sim->layout_gen++;
        }
    }
    return MPM_OK;
}

int mpm_get_launch_config(mpm_sim* sim, mpm_launch_config* cfg) {
    if (!sim) return MPM_ERR_INVALID;
    if (!cfg) return fail(sim, MPM_ERR_INVALID, "cfg must not be NULL");
    cfg->particle_blocks = sim->pblocks;
    cfg->grid_blocks = sim->gblocks;
    cfg->sort_every = sim->sort_every;
    cfg->particles_per_thread = sim->ppt;
    cfg->segment_tiles = sim->seg_tiles;
    return MPM_OK;
}

// validate_prescribed_step (mpm_prescribed_solids.rs:78-89)
int check_prescribed_step(mpm_sim* s, float dt) {
    if (s->rigid_count && s->prescribed &&
        (s->jacobi_iters > 0 || !(dt > 0.0f) || !std::isfinite(dt) || !std::isfinite(s->cfg[0].gravity_y)))
        return fail(s, MPM_ERR_INVALID,
                    "MPM prescribed solids require explicit stepping, finite gravity and positive finite dt");
    return MPM_OK;
}

int mpm_step_n(mpm_sim* sim, float dt, uint32_t substeps) {
    if (!sim) return MPM_ERR_INVALID;
    if (!std::isfinite(dt) || dt < 0.0f) return fail(sim, MPM_ERR_INVALID, "dt must be finite and >= 0");
    if (check_prescribed_step(sim, dt) != MPM_OK) return MPM_ERR_INVALID;
    if (sim->n == 0 || substeps == 0) return MPM_OK;
    sim->diff_steps = 0;
    return run_step_n(sim, dt, substeps);
}

int mpm_step_n_graph(mpm_sim* sim, float dt, uint32_t substeps) {
    if (!sim) return MPM_ERR_INVALID;
    if (!std::isfinite(dt) || dt < 0.0f) return fail(sim, MPM_ERR_INVALID, "dt must be finite and >= 0");
    if (check_prescribed_step(sim, dt) != MPM_OK) return MPM_ERR_INVALID;
    if (sim->n == 0 || substeps == 0) return MPM_OK;
    sim->diff_steps = 0;
    uint32_t dtb;
    std::memcpy(&dtb, &dt, 4);
    const bool sort_start = sim->p2g_mode == MPM_P2G_FUSED &&
                            (!sim->sorted || sim->sort_every == 0 || sim->since_sort >= sim->sort_every);
    GraphKey key{dtb, substeps, uint32_t(sim->p2g_mode), uint32_t(sim->ps_cur), uint32_t(sim->list_cur),
                 uint32_t(sim->grid_dirty), uint32_t(sort_start) | (sim->since_sort << 1), uint32_t(sim->write_vc),
                 sim->pblocks, sim->gblocks, sim->sort_every, uint32_t(sim->has_E),
                 sim->ppt | (sim->seg_tiles << 4) | (sim->chunk_ppt << 8)};
    auto it = sim->graphs.find(key);
    if (it == sim->graphs.end()) {
        // Lazily allocated tables must exist before capture.
        {
            int rc0 = ensure_grid(sim);
            if (rc0 != MPM_OK) return rc0;
        }
        if (sim->p2g_mode == MPM_P2G_DETERMINISTIC && !sim->affine)
            HIP_TRY(sim, dalloc(&sim->affine, (size_t)12 * sim->cap));
        hipGraph_t g;
        HIP_TRY(sim, hipStreamBeginCapture(sim->stream, hipStreamCaptureModeThreadLocal));
        sim->capturing = true;
        int rc = run_step_n(sim, dt, substeps);
        sim->capturing = false;
        hipError_t e = hipStreamEndCapture(sim->stream, &g);
        if (rc != MPM_OK) return rc;
        if (e != hipSuccess) return hip_fail(sim, e, "hipStreamEndCapture");
        hipGraphExec_t ex;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(sim, e, "hipGraphInstantiate");
        it = sim->graphs.emplace(key, ex).first;
    } else {
        sim->dry = true;
        int rc = run_step_n(sim, dt, substeps);
        sim->dry = false;
        if (rc != MPM_OK) return rc;
    }
    HIP_TRY(sim, hipGraphLaunch(it->second, sim->stream));
    return MPM_OK;
}

int mpm_synchronize(mpm_sim* sim) {
    if (!sim) return MPM_ERR_INVALID;
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    flush_profile(sim);
    return MPM_OK;
}

int mpm_read_grid_mass_momentum(mpm_sim* sim, uint32_t scene, float* mass, float* momentum) {
    if (!sim || !mass || !momentum) return MPM_ERR_INVALID;
    if (scene >= sim->n_scenes) return fail(sim, MPM_ERR_INVALID, "scene out of range");
    if (sim->pool_cap)  // mpm.rs:2826-2832
        return fail(sim, MPM_ERR_INVALID,
                    "sparse tile pool: use mpm_read_grid_storage_mass_momentum for the physical slots");
    const uint32_t G = sim->cfg[scene].grid_size;
    const size_t cells = (size_t)G * G * G;
    {
        int rc0 = ensure_grid(sim);
        if (rc0 != MPM_OK) return rc0;
    }
    float* buf = nullptr;
    HIP_TRY(sim, hipMalloc((void**)&buf, cells * 4 * sizeof(float)));
    hipError_t e = mpm::launch_grid_to_dense(sim->gmp, sim->d_scenes, scene, G, buf, buf + cells, sim->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(mass, buf, cells * 4, hipMemcpyDeviceToHost, sim->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(momentum, buf + cells, cells * 12, hipMemcpyDeviceToHost, sim->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(sim, e, "read_grid_mass_momentum");
    return MPM_OK;
}

int mpm_read_status(mpm_sim* sim, mpm_status_info* out) {
    if (!sim || !out) return MPM_ERR_INVALID;
    uint32_t w[mpm::ST_WORDS];
    HIP_TRY(sim, hipMemcpyAsync(w, sim->status, sizeof(w), hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    out->fault_bits = w[mpm::ST_FAULT];
    out->completed_gathers = w[mpm::ST_GATHERS];
    out->active_tiles = sim->pool_cap ? w[mpm::ST_POOL] : w[mpm::ST_COUNT0 + sim->list_cur];
    out->sorts = sim->sorts;
    out->escapes = w[mpm::ST_ESCAPES];
    out->fallback_chunks = w[mpm::ST_FALLBACK];
    out->chunks = w[mpm::ST_NCHUNKS];
    return MPM_OK;
}

int mpm_reset_status(mpm_sim* sim) {
    if (!sim) return MPM_ERR_INVALID;
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    uint32_t fault = 0;
    HIP_TRY(sim, hipMemcpy(&fault, sim->status + mpm::ST_FAULT, 4, hipMemcpyDeviceToHost));
    if (fault) sim->fault_reset = true;
    uint32_t z[2] = {0u, 0u};
    HIP_TRY(sim, hipMemcpy(sim->status, z, sizeof(z), hipMemcpyHostToDevice));
    HIP_TRY(sim, hipMemcpy(sim->status + mpm::ST_ESCAPES, z, 4, hipMemcpyHostToDevice));
    HIP_TRY(sim, hipMemcpy(sim->status + mpm::ST_FALLBACK, z, 4, hipMemcpyHostToDevice));
    HIP_TRY(sim, hipMemcpy(sim->status + mpm::ST_POOL, z, 8, hipMemcpyHostToDevice));
    return MPM_OK;
}

// ---- sparse tile pool (MpmSimulation::new_sparse / reserve_sparse_tiles,
//      mpm.rs:848-853,1280-1390) ----------------------------------------------

int mpm_create_sparse(const mpm_scene_config* cfg, uint32_t tile_capacity, mpm_sim** out) {
    if (!cfg || !out) return MPM_ERR_INVALID;
    int rc = create(cfg, 1, out);
    if (rc != MPM_OK) return rc;
    mpm_sim* s = *out;
    const uint64_t W = s->sh[0].W;
    if (tile_capacity == 0 || tile_capacity > W * W * W)
        return fail(s, MPM_ERR_INVALID, "sparse tile capacity must be in [1, ceil(grid_size / 4)^3]");
    if (cfg->model != MPM_MODEL_SHADER)
        return fail(s, MPM_ERR_INVALID, "the sparse tile pool is defined for MPM_MODEL_SHADER scenes");
    s->pool_cap = tile_capacity;
    return MPM_OK;
}

int mpm_reserve_sparse_tiles(mpm_sim* sim, uint32_t tile_capacity) {
    if (!sim) return MPM_ERR_INVALID;
    const uint64_t W = sim->sh[0].W;
    if (!sim->pool_cap || tile_capacity == 0 || tile_capacity > W * W * W)
        return fail(sim, MPM_ERR_INVALID, "MPM compact tile reservation is invalid");
    if (tile_capacity <= sim->pool_cap) return MPM_OK;
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    if (sim->gmp) {
        // new storage; grid state starts over on the next step (particles keep their history)
        const size_t cells = (size_t)tile_capacity * 64;
        float4 *gmp = nullptr, *gv = nullptr, *gv2 = nullptr;
        uint32_t* slot_tile = nullptr;
        HIP_TRY(sim, dalloc(&gmp, cells));
        HIP_TRY(sim, dalloc(&gv, cells));
        if (sim->gv2) HIP_TRY(sim, dalloc(&gv2, cells));
        HIP_TRY(sim, dalloc(&slot_tile, tile_capacity));
        HIP_TRY(sim, hipMemsetAsync(gmp, 0, cells * sizeof(float4), sim->stream));
        HIP_TRY(sim, hipMemsetAsync(gv, 0, cells * sizeof(float4), sim->stream));
        HIP_TRY(sim, hipMemsetAsync(sim->tile_slot, 0, (size_t)sim->total_tiles * 4, sim->stream));
        HIP_TRY(sim, hipMemsetAsync(sim->tile_mark, 0, (size_t)sim->total_tiles * 4, sim->stream));
        HIP_TRY(sim, hipMemsetAsync(sim->tile_flag, 0, (size_t)sim->total_tiles * 4, sim->stream));
        const uint32_t z[4] = {0u, 0u, 0u, 0u};
        HIP_TRY(sim, hipMemcpyAsync(sim->status + mpm::ST_COUNT0, z, 8, hipMemcpyHostToDevice, sim->stream));
        HIP_TRY(sim, hipMemcpyAsync(sim->status + mpm::ST_POOL, z, 8, hipMemcpyHostToDevice, sim->stream));
        HIP_TRY(sim, hipStreamSynchronize(sim->stream));
        (void)hipFree(sim->gmp);
        (void)hipFree(sim->gv);
        if (sim->gv2) (void)hipFree(sim->gv2);
        (void)hipFree(sim->slot_tile);
        sim->gmp = gmp;
        sim->gv = gv;
        sim->gv2 = gv2;
        sim->slot_tile = slot_tile;
        sim->grid_dirty = false;
        sim->sorted = false;  // the static list and flags start over with the next sort
    }
    sim->pool_cap = tile_capacity;
    drop_graphs(sim);
    return MPM_OK;
}

int mpm_grid_storage_bytes(mpm_sim* sim, uint64_t* bytes) {
    if (!sim || !bytes) return MPM_ERR_INVALID;
    const uint64_t cells = grid_cells(sim);
    uint64_t b = cells * 32u + (sim->jacobi_iters > 0 ? cells * 16u : 0u);  // gmp + gv (+ the sweep buffer)
    b += sim->pool_cap ? (uint64_t)sim->total_tiles * 8u + (uint64_t)sim->pool_cap * 4u : 0u;  // slot table, marks
    *bytes = b;
    return MPM_OK;
}

int mpm_read_grid_storage_mass_momentum(mpm_sim* sim, float* mass, float* momentum, uint32_t n_cells) {
    if (!sim || !mass || !momentum) return MPM_ERR_INVALID;
    const size_t cells = grid_cells(sim);
    if (n_cells != cells)
        return fail(sim, MPM_ERR_INVALID, "n_cells must equal the grid storage (" + std::to_string(cells) + " cells)");
    int rc = ensure_grid(sim);
    if (rc != MPM_OK) return rc;
    float* buf = nullptr;
    HIP_TRY(sim, hipMalloc((void**)&buf, cells * 4 * sizeof(float)));
    hipError_t e = mpm::launch_storage_to_host_layout(sim->gmp, uint32_t(cells), buf, buf + cells, sim->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(mass, buf, cells * 4, hipMemcpyDeviceToHost, sim->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(momentum, buf + cells, cells * 12, hipMemcpyDeviceToHost, sim->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(sim, e, "read_grid_storage_mass_momentum");
    return MPM_OK;
}

int mpm_read_active_tile_ids(mpm_sim* sim, uint32_t* ids, uint32_t capacity, uint32_t* count) {
    if (!sim || !count) return MPM_ERR_INVALID;
    if (!sim->pool_cap) return fail(sim, MPM_ERR_INVALID, "tile ids are read from a sparse tile pool");
    uint32_t w[2] = {0u, 0u};
    HIP_TRY(sim, hipMemcpyAsync(w, sim->status, 8, hipMemcpyDeviceToHost, sim->stream));
    uint32_t pool = 0;
    HIP_TRY(sim, hipMemcpyAsync(&pool, sim->status + mpm::ST_POOL, 4, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    if (w[mpm::ST_FAULT]) return fail(sim, MPM_ERR_FAULT, "MPM tile snapshot has a latched grid fault");
    const uint32_t n = std::min(pool, sim->pool_cap);
    *count = n;
    if (!sim->slot_tile || n == 0) return MPM_OK;
    if (!ids || capacity < n) return fail(sim, MPM_ERR_INVALID, "ids capacity below the active tile count");
    HIP_TRY(sim, hipMemcpy(ids, sim->slot_tile, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) ids[i] -= sim->sh[0].tile_base;
    return MPM_OK;
}

int mpm_sort_keys_debug(mpm_sim* sim, uint32_t scene, uint32_t capacity, uint32_t* keys, uint32_t* ids) {
    if (!sim || !keys || !ids) return MPM_ERR_INVALID;
    if (scene >= sim->n_scenes) return fail(sim, MPM_ERR_INVALID, "scene out of range");
    const uint32_t cnt = sim->counts[scene];
    if (capacity < cnt) return fail(sim, MPM_ERR_INVALID, "capacity below the particle count");
    if (sim->n > 0) {
        int rc = sort_particles(sim, false);
        if (rc != MPM_OK) return rc;
        sim->sorted = false;  // keys now describe unsorted positions: re-sort before stepping
        HIP_TRY(sim, hipMemcpyAsync(keys, sim->keys + sim->pbase[scene], (size_t)cnt * 4, hipMemcpyDeviceToHost,
                                    sim->stream));
        HIP_TRY(sim, hipMemcpyAsync(ids, sim->vals + sim->pbase[scene], (size_t)cnt * 4, hipMemcpyDeviceToHost,
                                    sim->stream));
        HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    }
    const uint32_t tb = sim->sh[scene].tile_base, pb = sim->pbase[scene];
    for (uint32_t i = 0; i < cnt; ++i) {
        keys[i] -= tb;
        ids[i] -= pb;
    }
    for (uint32_t i = cnt; i < capacity; ++i) {
        keys[i] = 0xffffffffu;
        ids[i] = i;
    }
    return MPM_OK;
}

int mpm_sort_keys_noalloc(mpm_sim* sim, uint32_t scene, uint32_t capacity, uint32_t* keys_dev, uint32_t* ids_dev) {
    if (!sim) return MPM_ERR_INVALID;
    if (!keys_dev || !ids_dev) return fail(sim, MPM_ERR_INVALID, "mpm_sort_keys_noalloc: null output buffer");
    if (scene >= sim->n_scenes) return fail(sim, MPM_ERR_INVALID, "scene out of range");
    const uint32_t cnt = sim->counts[scene];
    if (capacity < cnt) return fail(sim, MPM_ERR_INVALID, "capacity below the particle count");
    if (cnt > MPM_NOALLOC_MAX_PARTICLES)
        return fail(sim, MPM_ERR_INVALID,
                    "mpm_sort_keys_noalloc: the scene holds " + std::to_string(cnt) +
                        " particles, above the no-allocation limit of " +
                        std::to_string(MPM_NOALLOC_MAX_PARTICLES));
    const uint32_t W = sim->sh[scene].W;
    const int bits = int(ceil_log2(std::max(W * W * W, 2u)));
    if (mpm::radix_scratch_words(cnt) > NOALLOC_WORDS)
        return fail(sim, MPM_ERR_INVALID, "mpm_sort_keys_noalloc: scratch too small for the sort");
    HIP_TRY(sim, hipSetDevice(sim->device));
    uint32_t* scr = nullptr;
    int rc = noalloc_scratch(sim, &scr);
    if (rc != MPM_OK) return rc;
    if (cnt > 0) {
        const unsigned blocks = std::min<uint32_t>((cnt + 255) / 256, 8192u);
        hipLaunchKernelGGL(k_scene_keys, dim3(blocks), dim3(256), 0, sim->stream, sim->ps[sim->ps_cur], sim->cap,
                           sim->d_scenes, scene, sim->pbase[scene], cnt, keys_dev, ids_dev);
        HIP_TRY(sim, hipGetLastError());
        HIP_TRY(sim, mpm::radix_sort_pairs(keys_dev, ids_dev, cnt, bits, scr, sim->stream));
    }
    if (capacity > cnt) {
        const unsigned blocks = std::min<uint32_t>((capacity - cnt + 255) / 256, 8192u);
        hipLaunchKernelGGL(k_pad_keys, dim3(blocks), dim3(256), 0, sim->stream, keys_dev, ids_dev, cnt, capacity);
        HIP_TRY(sim, hipGetLastError());
    }
    return MPM_OK;
}

int mpm_set_profiling(mpm_sim* sim, int enabled) {
    if (!sim) return MPM_ERR_INVALID;
    const int all = MPM_PROFILE_STAGES | MPM_PROFILE_CALLS;
    sim->profiling = (enabled & ~all) ? all : enabled;
    // events for a few hundred profiled substeps up front: creating one per
    // launch inside a timed region costs more host time than the launch
    if (sim->profiling & MPM_PROFILE_STAGES) {
        while (sim->ev_pool.size() < 4096) {
            hipEvent_t e = nullptr;
            HIP_TRY(sim, timing_event(&e));
            sim->ev_pool.push_back(e);
        }
    }
    return MPM_OK;
}

int mpm_read_profile(mpm_sim* sim, float* ms, uint32_t* launches) {
    if (!sim) return MPM_ERR_INVALID;
    flush_profile(sim);
    for (int i = 0; i < MPM_STAGE_COUNT; ++i) {
        if (ms) ms[i] = float(sim->prof_ms[i]);
        if (launches) launches[i] = sim->prof_n[i];
    }
    return MPM_OK;
}

int mpm_reset_profile(mpm_sim* sim) {
    if (!sim) return MPM_ERR_INVALID;
    flush_profile(sim);
    for (int i = 0; i < MPM_STAGE_COUNT; ++i) {
        sim->prof_ms[i] = 0.0;
        sim->prof_n[i] = 0;
    }
    return MPM_OK;
}

int mpm_diff_forward(mpm_sim* sim, float dt, uint32_t steps, uint32_t checkpoint_every) {
    if (!sim) return MPM_ERR_INVALID;
    return diff_forward(sim, dt, steps, checkpoint_every);
}

int mpm_diff_backward(mpm_sim* sim, const float* grad_x_final, const float* grad_v_final, float* grad_x0,
                      float* grad_v0, float* grad_E, double* grad_mu_lambda) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = diff_backward(sim, grad_x_final, grad_v_final);
    if (rc != MPM_OK) return rc;
    return read_backward_outputs(sim, grad_x0, grad_v0, grad_E, grad_mu_lambda);
}

int mpm_diff_backward_device(mpm_sim* sim, const float* grad_x_final_dev, const float* grad_v_final_dev,
                             float* grad_x0_dev, float* grad_v0_dev, float* grad_E_dev, double* grad_mu_lambda_dev) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = diff_backward(sim, grad_x_final_dev, grad_v_final_dev, true);
    if (rc != MPM_OK) return rc;
    return read_backward_outputs_device(sim, grad_x0_dev, grad_v0_dev, grad_E_dev, grad_mu_lambda_dev);
}

int mpm_autotune(mpm_sim* sim, float dt, uint32_t trial_substeps, uint32_t* particle_blocks,
                 uint32_t* grid_blocks_, uint32_t* sort_every) {
    if (!sim) return MPM_ERR_INVALID;
    return autotune(sim, dt, trial_substeps, particle_blocks, grid_blocks_, sort_every);
}

int mpm_autotune_config(mpm_sim* sim, float dt, uint32_t trial_substeps, mpm_launch_config* out) {
    if (!sim) return MPM_ERR_INVALID;
    mpm_launch_config c{};
    const int rc = autotune(sim, dt, trial_substeps, &c.particle_blocks, &c.grid_blocks, &c.sort_every,
                            &c.particles_per_thread, &c.segment_tiles);
    if (rc == MPM_OK && out) *out = c;
    return rc;
}

int mpm_diff_read_adjoint(mpm_sim* sim, int field, float* out, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    if (!sim->adj_valid) return fail(sim, MPM_ERR_INVALID, "mpm_diff_backward must run first");
    if (n != sim->n) return fail(sim, MPM_ERR_INVALID, "output length must equal the particle count");
    if (!out) return fail(sim, MPM_ERR_INVALID, "out must not be NULL");
    return read_adjoint(sim, field, out);
}

// ---- rigid bodies on the grid ----------------------------------------------

namespace {

int rigid_handle(mpm_sim* s) {
    if (s->n_scenes != 1 || s->cfg[0].model != MPM_MODEL_SHADER)
        return fail(s, MPM_ERR_INVALID, "rigid bodies are defined for single-scene MPM_MODEL_SHADER handles");
    return MPM_OK;
}

// Replaces the body table (5 float4 rows per body) and clears the reaction.
int upload_rigid(mpm_sim* s, const std::vector<float4>& rows, uint32_t n, bool prescribed) {
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (!s->d_rig) {
        HIP_TRY(s, dalloc(&s->d_rig, (size_t)mpm::RIGID_ROWS * MPM_MAX_RIGID));
        HIP_TRY(s, dalloc(&s->d_wrench, (size_t)8 * MPM_MAX_RIGID));
    }
    if (n) HIP_TRY(s, hipMemcpy(s->d_rig, rows.data(), rows.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemset(s->d_wrench, 0, (size_t)8 * MPM_MAX_RIGID * sizeof(double)));
    const bool pres = prescribed && n > 0;
    if (n != s->rigid_count || pres != s->prescribed) drop_graphs(s);  // launches carry the count and kind
    s->rigid_count = n;
    s->prescribed = pres;
    return MPM_OK;
}

}  // namespace

int mpm_set_rigid_spheres(mpm_sim* sim, const float* spheres, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = rigid_handle(sim);
    if (rc != MPM_OK) return rc;
    if (n > MPM_MAX_RIGID)
        return fail(sim, MPM_ERR_INVALID, "set_rigid_spheres: n=" + std::to_string(n) + " > max_rigid=" +
                                              std::to_string(MPM_MAX_RIGID));
    if (n && !spheres) return fail(sim, MPM_ERR_INVALID, "spheres must not be NULL");
    std::vector<float4> rows;
    for (uint32_t r = 0; r < n; ++r) {
        const float* q = spheres + 7 * (size_t)r;
        rows.push_back(make_float4(q[0], q[1], q[2], q[3]));
        rows.push_back(make_float4(q[4], q[5], q[6], 0.f));
        rows.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
        rows.push_back(make_float4(0.f, 0.f, 0.f, 1.f));
        rows.push_back(make_float4(0.f, 0.f, 0.f, 0.f));  // kind 0: legacy sphere
    }
    return upload_rigid(sim, rows, n, false);
}

int mpm_set_prescribed_solids(mpm_sim* sim, const mpm_prescribed_solid* solids, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = rigid_handle(sim);
    if (rc != MPM_OK) return rc;
    if (n && (sim->jacobi_iters > 0 || !std::isfinite(sim->cfg[0].gravity_y)))
        return fail(sim, MPM_ERR_INVALID, "MPM prescribed solids require explicit stepping and finite gravity");
    if (n > MPM_MAX_RIGID) return fail(sim, MPM_ERR_INVALID, "MPM prescribed solid capacity exceeded");
    if (n && !solids) return fail(sim, MPM_ERR_INVALID, "solids must not be NULL");
    // MpmPrescribedSolid::rows (mpm_prescribed_solids.rs:34-75)
    std::vector<float4> rows;
    for (uint32_t r = 0; r < n; ++r) {
        const mpm_prescribed_solid& b = solids[r];
        bool ok = true;
        for (int k = 0; k < 3; ++k)
            ok = ok && std::isfinite(b.center[k]) && std::isfinite(b.linear_velocity[k]) &&
                 std::isfinite(b.angular_velocity[k]);
        float n2 = 0.f;
        for (int k = 0; k < 4; ++k) {
            ok = ok && std::isfinite(b.orientation[k]);
            n2 += b.orientation[k] * b.orientation[k];
        }
        const float norm = std::sqrt(n2);
        ok = ok && std::isfinite(norm) && std::fabs(norm - 1.0f) <= 1e-4f;
        float dims[4], radius = 0.f;
        switch (b.shape) {
            case MPM_SOLID_SPHERE:
                dims[0] = dims[1] = dims[2] = b.dims[0];
                dims[3] = 1.f;
                radius = b.dims[0];
                break;
            case MPM_SOLID_BOX:
                dims[0] = b.dims[0], dims[1] = b.dims[1], dims[2] = b.dims[2];
                dims[3] = 2.f;
                radius = std::sqrt(b.dims[0] * b.dims[0] + b.dims[1] * b.dims[1] + b.dims[2] * b.dims[2]);
                break;
            case MPM_SOLID_CYLINDER:
                dims[0] = b.dims[0], dims[1] = b.dims[1], dims[2] = b.dims[0];
                dims[3] = 3.f;
                radius = std::hypot(b.dims[0], b.dims[1]);
                break;
            default: ok = false;
        }
        if (ok) {
            ok = std::isfinite(radius);
            for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(dims[k]) && dims[k] > 0.f;
        }
        if (!ok) return fail(sim, MPM_ERR_INVALID, "MPM prescribed solid: invalid geometry or pose");
        rows.push_back(make_float4(b.center[0], b.center[1], b.center[2], radius));
        rows.push_back(make_float4(b.linear_velocity[0], b.linear_velocity[1], b.linear_velocity[2], 0.f));
        rows.push_back(make_float4(b.angular_velocity[0], b.angular_velocity[1], b.angular_velocity[2], 0.f));
        rows.push_back(make_float4(b.orientation[0] / norm, b.orientation[1] / norm, b.orientation[2] / norm,
                                   b.orientation[3] / norm));
        rows.push_back(make_float4(dims[0], dims[1], dims[2], dims[3]));
    }
    return upload_rigid(sim, rows, n, true);
}

int mpm_read_rigid_wrenches(mpm_sim* sim, float* out, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    if (n != sim->rigid_count) return fail(sim, MPM_ERR_INVALID, "n must equal the rigid body count");
    if (n == 0) return MPM_OK;
    if (!out) return fail(sim, MPM_ERR_INVALID, "out must not be NULL");
    std::vector<double> w((size_t)8 * n);
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    HIP_TRY(sim, hipMemcpy(w.data(), sim->d_wrench, w.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < n; ++r)
        for (int k = 0; k < 3; ++k) {
            out[6 * r + k] = float(w[8 * r + k]);
            out[6 * r + 3 + k] = float(w[8 * r + 4 + k]);
        }
    return MPM_OK;
}

// ---- soil checkpoint (mpm_checkpoint.rs) ----------------------------------

namespace {

int soil_handle(mpm_sim* s) {
    if (s->n_scenes != 1 || s->cfg[0].model != MPM_MODEL_SHADER)
        return fail(s, MPM_ERR_INVALID, "the soil checkpoint is defined for single-scene MPM_MODEL_SHADER handles");
    return MPM_OK;
}

int soil_fail(mpm_sim* s) { return fail(s, MPM_ERR_INVALID, mpm_soil_last_error()); }

// A device buffer that lives for one archive operation.
struct TempWords {
    uint32_t* p = nullptr;
    ~TempWords() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int mpm_checkpoint_capture(mpm_sim* sim, uint32_t* words, uint32_t n) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = soil_handle(sim);
    if (rc != MPM_OK) return rc;
    if (n != sim->n) return fail(sim, MPM_ERR_INVALID, "n must equal the particle count");
    if (n && !words) return fail(sim, MPM_ERR_INVALID, "words must not be NULL");
    HIP_TRY(sim, hipSetDevice(sim->device));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    uint32_t fault = 0;
    HIP_TRY(sim, hipMemcpy(&fault, sim->status + mpm::ST_FAULT, 4, hipMemcpyDeviceToHost));
    if (fault || sim->fault_reset)
        return fail(sim, MPM_ERR_FAULT,
                    "MPM soil checkpoint: latched grid fault; incomplete state cannot be archived");
    if (n == 0) return MPM_OK;
    TempWords d;
    const size_t nw = (size_t)MPM_SOIL_WORDS_PER_PARTICLE * n;
    HIP_TRY(sim, dalloc(&d.p, nw));
    HIP_TRY(sim, mpm::launch_ckpt_pack(sim->ps[sim->ps_cur], sim->cap, sim->slot_of_id,
                                       sim->aux_live ? sim->ck_aux : nullptr, sim->w_stepped ? 1 : 0, n, d.p,
                                       sim->stream));
    HIP_TRY(sim, hipMemcpyAsync(words, d.p, nw * 4, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_checkpoint_restore(mpm_sim* sim, const mpm_soil_meta* meta, const uint32_t* words,
                           const mpm_material* rows, uint32_t n_rows) {
    if (!sim) return MPM_ERR_INVALID;
    int rc = soil_handle(sim);
    if (rc != MPM_OK) return rc;
    if (!meta || !rows) return fail(sim, MPM_ERR_INVALID, "meta and rows must not be NULL");
    if (n_rows != std::max<uint32_t>(meta->n_materials, 1) || n_rows > MPM_MAX_MATERIALS)
        return fail(sim, MPM_ERR_INVALID, "rows must pack the checkpoint's material table (max(n_materials, 1) rows)");
    if (mpm_soil_validate(meta, words) != MPM_OK) return soil_fail(sim);
    const int32_t* win = meta->grid_window_cells;
    if ((win[0] || win[1] || win[2]) && mpm_soil_validate_window(meta, words) != MPM_OK) return soil_fail(sim);
    const mpm_scene_config& c = sim->cfg[0];
    auto bits = [](float f) {
        uint32_t w;
        std::memcpy(&w, &f, 4);
        return w;
    };
    const uint32_t* k = meta->config;
    const uint32_t iters = k[19] == 1 ? k[20] : 0;
    const bool same = k[0] == c.grid_size && k[3] == bits(c.cell_size) && k[4] == bits(c.grid_origin[0]) &&
                      k[5] == bits(c.grid_origin[1]) && k[6] == bits(c.grid_origin[2]) &&
                      k[7] == bits(c.gravity_y) && k[8] == bits(c.youngs) && k[9] == bits(c.poisson) &&
                      k[12] == bits(c.boundary_friction) && k[13] == bits(c.particle_mass) &&
                      k[14] == bits(c.particle_volume) && iters == sim->jacobi_iters;
    if (!same || meta->count > c.max_particles || meta->count > sim->cap || sim->rigid_count != 0)
        return fail(sim, MPM_ERR_INVALID,
                    "MPM soil checkpoint: incompatible configuration, capacity or attached rigid bodies");
    // validated: from here on the handle is replaced
    HIP_TRY(sim, hipSetDevice(sim->device));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    const uint32_t n = meta->count;
    if (!sim->ck_aux) HIP_TRY(sim, dalloc(&sim->ck_aux, (size_t)12 * sim->cap));
    TempWords d;
    if (n) {
        const size_t nw = (size_t)MPM_SOIL_WORDS_PER_PARTICLE * n;
        HIP_TRY(sim, dalloc(&d.p, nw));
        HIP_TRY(sim, hipMemcpyAsync(d.p, words, nw * 4, hipMemcpyHostToDevice, sim->stream));
        HIP_TRY(sim, mpm::launch_ckpt_unpack(d.p, n, sim->ps[0], sim->cap, c.youngs, sim->ck_aux, sim->stream));
    }
    mpm_scene_config nc = c;
    for (int a = 0; a < 3; ++a) nc.grid_window_cells[a] = win[a];
    for (uint32_t r = 0; r < n_rows; ++r) nc.materials[r] = rows[r];
    nc.n_materials = n_rows;
    sim->cfg[0] = nc;
    sim->sh[0] = make_scene(nc, sim->sh[0].tile_base, sim->sh[0].mat_base);
    sim->n = n;
    sim->counts.assign(1, n);
    sim->pbase.assign(1, 0);
    sim->ps_cur = 0;
    HIP_TRY(sim, mpm::launch_index_slots(sim->ps[0], sim->cap, sim->slot_of_id, n, sim->stream));
    if ((rc = build_chunks(sim)) != MPM_OK) return rc;
    if ((rc = upload_scenes(sim)) != MPM_OK) return rc;
    if (sim->gmp && (rc = clear_grid(sim)) != MPM_OK) return rc;
    const uint32_t z[2] = {0u, 0u};
    HIP_TRY(sim, hipMemcpyAsync(sim->status + mpm::ST_FAULT, z, sizeof(z), hipMemcpyHostToDevice, sim->stream));
    HIP_TRY(sim, hipMemcpyAsync(sim->status + mpm::ST_ESCAPES, z, 4, hipMemcpyHostToDevice, sim->stream));
    HIP_TRY(sim, hipMemcpyAsync(sim->status + mpm::ST_FALLBACK, z, 4, hipMemcpyHostToDevice, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
    sim->has_E = 0;
    sim->sorted = false;
    sim->diff_steps = 0;
    sim->adj_valid = false;
    sim->aux_live = true;
    sim->w_stepped = false;
    sim->fault_reset = false;
    drop_graphs(sim);
    return MPM_OK;
}

}  // extern "C"
