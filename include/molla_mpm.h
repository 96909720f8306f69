/*
 * molla_mpm.h -- C ABI of the MI355X-native MPM substep (libmolla_mpm.so).
 *
 * Drop-in boundary for molla's `MpmSimulation` stepping API
 * (crates/molla-solvers/src/mpm.rs).  Every entry point names the reference
 * item it replaces.  Plain pointers and sizes only; all host arrays are
 * little-endian f32/u32, particles in their original (caller) order:
 *   positions / velocities : n * 3 floats  ([f32; 3] per particle)
 *   C, F                   : n * 9 floats  (row-major 3x3 per particle)
 *   grid mass              : G^3 floats, momentum 3 * G^3, cell = x + G(y + G z)
 *
 * Errors: every call returns an mpm_status; the message of the last failure
 * on a handle is available from mpm_last_error() (reference:
 * molla_core::Error::{Solver, Gpu}, crates/molla-core/src/error.rs:10-39).
 * Device faults are latched, not returned, and are read with
 * mpm_read_status() (reference: read_active_grid_status_with, mpm.rs:1418).
 *
 * Threading: a handle is externally synchronised (one host thread at a time,
 * like the reference's `&mut self`); distinct handles are independent.  All
 * work is ordered on the handle's HIP stream; reads block until it drains.
 */
#ifndef MOLLA_MPM_H
#define MOLLA_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 1
#define MPM_MAX_MATERIALS 8 /* MAX_MPM_MATERIALS, mpm.rs:661 */

typedef enum {
    MPM_OK = 0,
    MPM_ERR_INVALID = 1, /* Error::Solver: validation failure          */
    MPM_ERR_DEVICE = 2,  /* Error::Gpu: HIP runtime / allocation error */
    MPM_ERR_FAULT = 3    /* a latched device fault blocks the request  */
} mpm_status;

/* Semantics: the GPU step of mpm.wgsl, or the differentiable
 * CPU step of molla-autograd/src/mpm_cpu_diff.rs. */
typedef enum { MPM_MODEL_SHADER = 0, MPM_MODEL_CPUDIFF = 1 } mpm_model;

/* Rheology ids of the material table (material_row, mpm.rs:666-686);
 * MPM_RHEO_FLUID is this library's weakly-compressible fluid. */
typedef enum {
    MPM_RHEO_SAND = 0,
    MPM_RHEO_SNOW = 1,
    MPM_RHEO_VISCOUS = 2,
    MPM_RHEO_FLUID = 3
} mpm_rheology;

/* One material-table row, in the reference's packing order
 * (p0 = rheology, dp_alpha, theta_c, theta_s; p1 = hardening,
 * velocity_damping, affine_damping, <bulk>). */
typedef struct {
    float rheology;
    float dp_alpha;
    float theta_c;
    float theta_s;
    float hardening;
    float velocity_damping;
    float affine_damping;
    float bulk;
} mpm_material;

/* Mirrors MpmConfig (mpm.rs:394-433) for one scene. */
typedef struct {
    uint32_t grid_size;
    float cell_size;
    float grid_origin[3];
    int32_t grid_window_cells[3]; /* MpmSimulation::grid_window_cells, mpm.rs:400-402 */
    float gravity_y;
    float youngs;
    float poisson;
    float boundary_friction;
    float particle_mass;
    float particle_volume;
    uint32_t max_particles;
    uint32_t model;   /* mpm_model */
    float floor_y;    /* MPM_MODEL_CPUDIFF: MpmCpuConfig::floor_y        */
    float floor_friction; /* MPM_MODEL_CPUDIFF: 1 = sticky (mpm_cpu_diff.rs:96-99) */
    uint32_t n_materials; /* rows used in `materials` (>= 1)             */
    mpm_material materials[MPM_MAX_MATERIALS];
} mpm_scene_config;

typedef struct mpm_sim mpm_sim;

/* Stepping options (set_p2g_mode, mpm.rs:1642; set_particle_tile_ordering,
 * mpm_particle_order.rs:244; set_deterministic_p2g,
 * mpm_deterministic_p2g.rs:550). */
typedef enum {
    MPM_P2G_FUSED = 0,         /* fused G2P->P2G kernel, LDS-staged (default)      */
    MPM_P2G_GLOBAL_ATOMIC = 1, /* unfused, one global atomic per node (legacy p2g) */
    MPM_P2G_DETERMINISTIC = 2  /* sorted gather-order P2G, bitwise repeatable      */
} mpm_p2g_mode;

typedef struct {
    uint32_t fault_bits;        /* 1 overflow, 2 numerical (mpm.wgsl:175-231) */
    uint32_t completed_gathers; /* G2P passes run (g_kdiag metadata word 3)   */
    uint32_t active_tiles;      /* 4^3 tiles touched by the last P2G          */
    uint32_t sorts;             /* particle re-sorts performed                */
    uint32_t escapes;           /* particle-substeps scattered by global atomics because the
                                   particle left its sorted tile neighbourhood */
    uint32_t fallback_chunks;   /* block passes that scattered with global atomics: P2G box
                                   above the LDS budget, or chunk beyond the partial buffer */
    uint32_t chunks;            /* blocks of the sorted layout's fused pass   */
} mpm_status_info;

/* ---- lifetime (MpmSimulation::new / new_with_device, mpm.rs:834-853) ---- */
int mpm_create(const mpm_scene_config* cfg, mpm_sim** out);
/* Batched: `n_scenes` independent scenes in one context (no reference
 * counterpart: the reference runs one MpmSimulation per scene). */
int mpm_create_batch(const mpm_scene_config* cfgs, uint32_t n_scenes, mpm_sim** out);
int mpm_destroy(mpm_sim* sim);
const char* mpm_last_error(const mpm_sim* sim);
int mpm_abi_version(void);

/* ---- sparse grid: a pool of 4^3 tiles (MpmSimulation::new_sparse,
 *      mpm.rs:848-853; reserve_sparse_tiles :1308; grid_storage_bytes :1280;
 *      read_active_tile_ids_with :1430; read_grid_storage_mass_momentum
 *      :2837) ----
 * One MPM_MODEL_SHADER scene whose grid storage holds `tile_capacity` tiles
 * (1 .. ceil(grid_size/4)^3).  Each substep gives slots to the tiles within
 * one cell of the particles' stencils (cells base-1 .. base+3, the
 * reference's tile marks) and of legacy rigid spheres (radius + 2 cells).
 * More marked tiles than the capacity latch the overflow fault: the
 * substep leaves particle state unchanged and completed_gathers stops, as
 * for any latched fault, until mpm_reset_status.  mpm_read_status reports
 * the marked tiles as active_tiles.  The deterministic P2G mode and
 * mpm_read_grid_mass_momentum need the dense grid (mpm.rs:1265-1268,
 * :2826-2832). */
int mpm_create_sparse(const mpm_scene_config* cfg, uint32_t tile_capacity, mpm_sim** out);
/* Grows the pool (a smaller capacity is a no-op); the grid state starts
 * over on the next step, the particles keep their history. */
int mpm_reserve_sparse_tiles(mpm_sim* sim, uint32_t tile_capacity);
/* Bytes of grid storage and tile metadata (particle state excluded). */
int mpm_grid_storage_bytes(mpm_sim* sim, uint64_t* bytes);
/* The physical storage of the last P2G: n_cells = 64 x tiles of storage
 * (the pool's slots in order, or every tile of a dense grid, tile-major);
 * mass n_cells floats, momentum 3 n_cells. */
int mpm_read_grid_storage_mass_momentum(mpm_sim* sim, float* mass, float* momentum, uint32_t n_cells);
/* Tile ids (tx + W (ty + W tz)) of the pool's slots in slot order; *count
 * receives the number in use.  Fails with MPM_ERR_FAULT while a fault is
 * latched. */
int mpm_read_active_tile_ids(mpm_sim* sim, uint32_t* ids, uint32_t capacity, uint32_t* count);

/* ---- particles (set_particles / set_particles_with_velocity, mpm.rs:1451-1519) ----
 * F = I, C = 0, jp = 0, mass/volume from the config; `velocities` may be NULL. */
int mpm_set_particles(mpm_sim* sim, const float* positions, const float* velocities, uint32_t n);
/* Batched variant: counts[s] particles per scene, concatenated scene-major. */
int mpm_set_particles_batch(mpm_sim* sim, const uint32_t* counts, const float* positions,
                            const float* velocities);

/* Per-particle fields (original order).  Field ids: */
typedef enum {
    MPM_FIELD_POSITION = 0, /* 3 */
    MPM_FIELD_VELOCITY = 1, /* 3 */
    MPM_FIELD_C = 2,        /* 9 */
    MPM_FIELD_F = 3,        /* 9 */
    MPM_FIELD_JP = 4,       /* 1  p_jp                                  */
    MPM_FIELD_MASS = 5,     /* 1  p_mass                                */
    MPM_FIELD_VOLUME = 6,   /* 1  p_volume                              */
    MPM_FIELD_YOUNGS = 7,   /* 1  per-particle E (extension)            */
    MPM_FIELD_MATERIAL = 8  /* 1  u32 row of the material table, p_material */
} mpm_field;
int mpm_write_field(mpm_sim* sim, int field, const void* data, uint32_t n);
int mpm_read_field(mpm_sim* sim, int field, void* out, uint32_t n);
/* No counterpart in the reference (its reads return host Vecs).  The same
 * field into a caller-owned DEVICE buffer of n x components words, in
 * original particle order, ordered on the handle's stream with no host
 * synchronisation: for callers that keep the state in HBM (a loss and its
 * gradient computed on the GPU, as for mpm_diff_backward_device). */
int mpm_read_field_device(mpm_sim* sim, int field, void* out_dev, uint32_t n);
/* read_positions_with / read_velocities_with (mpm.rs:2718-2790). */
int mpm_read_positions(mpm_sim* sim, float* out, uint32_t n);
int mpm_read_velocities(mpm_sim* sim, float* out, uint32_t n);

/* ---- materials (set_materials / set_rheology, mpm.rs:1570-1635) ---- */
int mpm_set_materials(mpm_sim* sim, uint32_t scene, const mpm_material* rows, uint32_t n_rows);
int mpm_set_scene_params(mpm_sim* sim, uint32_t scene, const mpm_scene_config* cfg);

/* ---- modes ---- */
int mpm_set_p2g_mode(mpm_sim* sim, int mode);
/* Re-sort particles by 4^3 tile every `every` substeps (0 = only when
 * stepping starts).  Sorting never changes results, only locality. */
int mpm_set_sort_interval(mpm_sim* sim, uint32_t every);
/* set_time_step (mpm.rs:1632): 0 = MpmTimeStep::Explicit, k > 0 =
 * Implicit { jacobi_iters: k } -- k semi-implicit Jacobi sweeps of the grid
 * velocity after each grid update (mpm.wgsl:1060-1153).  Shader model only. */
int mpm_set_time_step(mpm_sim* sim, uint32_t jacobi_iters);
/* Store v and C at every substep (1, default, reference semantics) or only
 * at the last substep of a step_n call (0). */
int mpm_set_write_vc(mpm_sim* sim, int every_substep);
/* Launch tuning: blocks for the particle pass (0 = one per chunk). */
int mpm_set_launch(mpm_sim* sim, uint32_t particle_blocks, uint32_t grid_blocks);

/* The whole launch configuration (no counterpart in the reference, whose
 * workgroups are fixed at 64 threads, mpm.rs:314):
 *   particle_blocks, grid_blocks  as mpm_set_launch;
 *   sort_every                    as mpm_set_sort_interval;
 *   particles_per_thread          1 or 2: particles each thread of the fused
 *                                 pass takes per chunk of the sorted layout
 *                                 (a chunk holds 256 x this many; shader
 *                                 model only, MPM_MODEL_CPUDIFF handles run 1);
 *   segment_tiles                 the chunk rule of the sorted layout; all
 *                                 keep a chunk inside one tile row, which
 *                                 bounds the LDS box of its scatter.  4 or
 *                                 6 (default): greedy rows, a chunk runs for
 *                                 256 x particles_per_thread particles or
 *                                 until it would span more than that many
 *                                 tiles; 8: aligned groups of eight tiles
 *                                 where they hold two chunks' worth, else of
 *                                 four, cut every 256 x particles_per_thread.
 * The block is 256 threads (4 waves): with the 168-VGPR cap of three waves
 * per SIMD and three ~49 KB LDS boxes per CU, it is the size that fills a
 * CU.  Changing particles_per_thread or segment_tiles re-sorts at the next
 * step.  Results agree to f32 rounding across configurations (the fixed-point
 * scatter's scale is per chunk). */
typedef struct {
    uint32_t particle_blocks;
    uint32_t grid_blocks;
    uint32_t sort_every;
    uint32_t particles_per_thread;
    uint32_t segment_tiles;
} mpm_launch_config;
int mpm_set_launch_config(mpm_sim* sim, const mpm_launch_config* cfg);
int mpm_get_launch_config(mpm_sim* sim, mpm_launch_config* cfg);

/* Launch autotuner.  No counterpart in the reference, whose dispatch sizes
 * are fixed; it serves the batched mode, where scene mixes differ widely.
 * Times `trial_substeps` substeps (after a 2-substep warm-up) for candidate
 * persistent particle-pass grids and grid-update grids, and two whole sort
 * intervals (at least `trial_substeps`) for each candidate interval, so
 * sorting is paid at its own rate (coordinate search, per-substep time);
 * applies the fastest as mpm_set_launch /
 * mpm_set_sort_interval would, and restores the particle state it started
 * from (stepping resumes from the same state).  Outputs may be NULL. */
int mpm_autotune(mpm_sim* sim, float dt, uint32_t trial_substeps, uint32_t* particle_blocks,
                 uint32_t* grid_blocks, uint32_t* sort_every);
/* The same search with particles_per_thread x segment_tiles first (all six
 * pairs, shader model), then the three above; `out` (may be NULL) receives
 * the configuration applied.  mpm_autotune runs this search too. */
int mpm_autotune_config(mpm_sim* sim, float dt, uint32_t trial_substeps, mpm_launch_config* out);

/* ---- stepping (step / step_n, mpm.rs:2232-2272) ---- */
int mpm_step_n(mpm_sim* sim, float dt, uint32_t substeps);
/* Capture step_n(dt, substeps) into a hipGraph once and replay it
 * (PreparedMpmStep / MpmCommandBatch, mpm_command_batch.rs:10-235). */
int mpm_step_n_graph(mpm_sim* sim, float dt, uint32_t substeps);
int mpm_synchronize(mpm_sim* sim);

/* ---- grid / status / ordering ---- */
/* read_grid_mass_momentum (mpm.rs:2792-2906): the last substep's P2G. */
int mpm_read_grid_mass_momentum(mpm_sim* sim, uint32_t scene, float* mass, float* momentum);
int mpm_read_status(mpm_sim* sim, mpm_status_info* out);
int mpm_reset_status(mpm_sim* sim); /* reset_active_grid, mpm.rs:1388 */
/* Tile keys and stable sorted ids of the current positions, exactly as the
 * ParticleOrder::encode path (mpm_particle_order.rs:100-146): keys of one
 * scene, padded to `capacity` with 0xffffffff / identity ids. */
int mpm_sort_keys_debug(mpm_sim* sim, uint32_t scene, uint32_t capacity, uint32_t* keys,
                        uint32_t* ids);
/* No counterpart in the reference.  mpm_sort_keys_debug for callers that
 * cannot allocate device memory: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc.  The same keys and ids, written to
 * the caller's DEVICE buffers keys_dev / ids_dev (`capacity` u32 each),
 * ordered on the handle's stream (mpm_synchronize waits for them).  Nothing
 * is allocated and the handle's sort state is left as it was.  The radix
 * sort works in a static device scratch array of the library, sized for
 * MPM_NOALLOC_MAX_PARTICLES particles in the scene (about 18 MB, held by
 * every GPU the library runs on until the process ends); a larger scene
 * fails with MPM_ERR_INVALID.  All calls on one GPU share that array, so
 * calls that overlap on one GPU (handles driven from several host threads
 * or streams at once) race. */
#define MPM_NOALLOC_MAX_PARTICLES 2097152u
int mpm_sort_keys_noalloc(mpm_sim* sim, uint32_t scene, uint32_t capacity, uint32_t* keys_dev,
                          uint32_t* ids_dev);

/* ---- profiling (step_n_profiled_with / read_step_profile, mpm_profile.rs:649-666) ---- */
typedef enum {
    MPM_STAGE_SORT = 0,
    MPM_STAGE_P2G = 1,
    MPM_STAGE_GRID = 2,
    MPM_STAGE_G2P2G = 3,
    MPM_STAGE_G2P = 4,
    /* whole calls: the device work of an mpm_diff_forward, and of an
     * mpm_diff_backward from the uploaded final-state gradient to the last
     * reverse kernel (the gradients' host transfers excluded) */
    MPM_STAGE_DIFF_FORWARD = 5,
    MPM_STAGE_DIFF_BACKWARD = 6,
    MPM_STAGE_COUNT = 7
} mpm_stage;
/* flags: MPM_PROFILE_STAGES times every launch group of the stages 0-4,
 * MPM_PROFILE_CALLS the whole-call stages 5-6 (two events per call);
 * any other nonzero value: both. */
#define MPM_PROFILE_STAGES 1
#define MPM_PROFILE_CALLS 2
int mpm_set_profiling(mpm_sim* sim, int flags);
/* Accumulated milliseconds and launch counts per stage since the last reset,
 * MPM_STAGE_COUNT entries each (HIP events on the handle's stream). */
int mpm_read_profile(mpm_sim* sim, float* ms, uint32_t* launches);
int mpm_reset_profile(mpm_sim* sim);

/* ---- differentiable path (mpm_cpu_diff.rs forward_with_tape :397-418 and
 *      mpm_adjoint.rs backward :121-137), MPM_MODEL_CPUDIFF only ----
 * Forward `steps` substeps keeping a checkpoint of the full particle state
 * every `checkpoint_every` substeps. */
int mpm_diff_forward(mpm_sim* sim, float dt, uint32_t steps, uint32_t checkpoint_every);
/* Reverse pass.  grad_x_final: n*3 dJ/dx_T (required); grad_v_final may be
 * NULL.  Outputs (each may be NULL): dJ/dx_0, dJ/dv_0 (n*3), dJ/dE_p (n),
 * and the global (dJ/dmu, dJ/dlambda) accumulators of StateAdjoint
 * (mpm_adjoint.rs:76-113). */
int mpm_diff_backward(mpm_sim* sim, const float* grad_x_final, const float* grad_v_final,
                      float* grad_x0, float* grad_v0, float* grad_E, double* grad_mu_lambda);
/* No counterpart in the reference, whose adjoint is host-only
 * (mpm_adjoint.rs:121-137).  mpm_diff_backward with every array a DEVICE
 * pointer (same shapes, original particle order): the gradients stay in HBM
 * and nothing crosses PCIe; the work is ordered on the handle's stream
 * (mpm_synchronize waits for the outputs). */
int mpm_diff_backward_device(mpm_sim* sim, const float* grad_x_final_dev, const float* grad_v_final_dev,
                             float* grad_x0_dev, float* grad_v0_dev, float* grad_E_dev,
                             double* grad_mu_lambda_dev);
/* The rest of the StateAdjoint of the last backward (mpm_adjoint.rs:76-113):
 * dJ/d(initial field) in original particle order for field POSITION,
 * VELOCITY, C, F (row-major 3x3) or YOUNGS. */
int mpm_diff_read_adjoint(mpm_sim* sim, int field, float* out, uint32_t n);

/* ---- rigid bodies on the grid: set_rigid_spheres (mpm.rs:1521-1566),
 *      set_prescribed_solids (mpm_prescribed_solids.rs:91-140),
 *      read_rigid_wrenches_with (mpm.rs:2908-2955); shader model, one scene ----
 * Every substep, after the P2G, the first body containing a grid node gives
 * the node its velocity (rigid_imprint, mpm.wgsl:937-1006) and the node skips
 * gravity, walls and the implicit sweeps.  Legacy spheres stamp a heavy cell
 * (mass 1e6) and also feel the sand in a one-cell shell outside
 * (mpm_rigid_force_integrate, mpm.wgsl:1174-1226); prescribed solids keep
 * the sand mass and need explicit stepping with dt > 0. */
#define MPM_MAX_RIGID 64 /* bodies per handle; the caller's max_rigid may be lower */
typedef enum { MPM_SOLID_SPHERE = 1, MPM_SOLID_BOX = 2, MPM_SOLID_CYLINDER = 3 } mpm_solid_shape;
/* MpmPrescribedSolid (mpm_prescribed_solids.rs:21-32). */
typedef struct {
    uint32_t shape;            /* mpm_solid_shape                                       */
    float dims[3];             /* sphere: radius; box: half extents; cylinder: radius,
                                  half width along local y                              */
    float center[3];
    float orientation[4];      /* unit quaternion, xyzw                                 */
    float linear_velocity[3];
    float angular_velocity[3]; /* rad/s: a node moves with v + w x (node - center)      */
} mpm_prescribed_solid;
/* Replaces the bodies with n legacy spheres: 7 floats each (center xyz,
 * radius, velocity xyz). */
int mpm_set_rigid_spheres(mpm_sim* sim, const float* spheres, uint32_t n);
/* Replaces the bodies with n prescribed solids (validated first: finite pose
 * and velocities, unit quaternion to 1e-4, positive finite dimensions;
 * nothing changes on failure). */
int mpm_set_prescribed_solids(mpm_sim* sim, const mpm_prescribed_solid* solids, uint32_t n);
/* Reaction of the last substep stepped: n = the body count, 6 floats each
 * (force xyz, torque xyz about the body center), bodies in set order. */
int mpm_read_rigid_wrenches(mpm_sim* sim, float* out, uint32_t n);

/* ---- soil particle checkpoint, "MOLLSOIL" v1/v2 (mpm_checkpoint.rs) ----
 * Byte layout (little-endian, MpmSoilCheckpoint::to_bytes, :185-214):
 *   [0, 8)   "MOLLSOIL"                [8, 12)  version: 1, or 2 when the
 *   [12, 44) BLAKE3 of the payload              grid window is non-zero
 *            (v2: of version bytes || payload, checkpoint_digest :87-94)
 *   payload  u32 words: [window x3 (v2 only)] config (21) n_materials
 *            materials (4 each) count, then 40 words per particle in nine
 *            streams of STRIDES = 4,4,12,12,1,1,1,1,4 words (:78-83):
 *            position+w, velocity+w, C and F as three padded columns
 *            (column-major, word 4j+i = M_ij), volume, mass, log-Jp,
 *            material id (u32), tire compaction.
 * The codec below is host-only and needs no handle; its messages are read
 * with mpm_soil_last_error(). */
#define MPM_SOIL_HEADER_BYTES 44        /* HEADER_BYTES, :81       */
#define MPM_SOIL_WORDS_PER_PARTICLE 40  /* WORDS_PER_PARTICLE, :83 */
#define MPM_SOIL_CONFIG_WORDS 21        /* config_words, :594-619  */

/* MpmSoilCheckpoint (:137-147) minus its particle words.  `config` is
 * config_words(MpmConfig): grid_size, max_particles, max_rigid, then the
 * f32 bits of cell_size, grid_origin[3], gravity_y, youngs, poisson,
 * friction_angle, cohesion, boundary_friction, particle_mass,
 * particle_volume, then the rheology (4 words, rheology_words :544-557:
 * kind 0 sand, 1 snow, 2 viscous, then its f32 parameters) and the time
 * step (0,0 explicit; 1,iters implicit). */
typedef struct {
    int32_t grid_window_cells[3];
    uint32_t config[MPM_SOIL_CONFIG_WORDS];
    uint32_t n_materials; /* 0: the config's rheology is authoritative */
    uint32_t materials[MPM_MAX_MATERIALS][4];
    uint32_t count;       /* particles; words hold 40 * count u32 */
} mpm_soil_meta;

/* Encoded size of a checkpoint with this metadata. */
uint64_t mpm_soil_encoded_bytes(const mpm_soil_meta* meta);
/* MpmSoilCheckpoint::to_bytes (:185-214): no semantic validation, like the
 * reference; `cap` must hold mpm_soil_encoded_bytes(meta). */
int mpm_soil_encode(const mpm_soil_meta* meta, const uint32_t* words, uint8_t* out, uint64_t cap);
/* MpmSoilCheckpoint::from_bytes (:216-267) up to, not including, the copy
 * of the particle words and validate(): format, version, allocation limit
 * (`max_particles`), magic, digest, canonical metadata, count and payload
 * length.  Fills `meta` so the caller can size the word buffer. */
int mpm_soil_decode_meta(const uint8_t* bytes, uint64_t len, uint32_t max_particles, mpm_soil_meta* meta);
/* The whole of from_bytes: as above, then validate() and copy the particle
 * words (words_cap >= 40 * count). */
int mpm_soil_decode(const uint8_t* bytes, uint64_t len, uint32_t max_particles, mpm_soil_meta* meta,
                    uint32_t* words, uint64_t words_cap);
/* MpmSoilCheckpoint::validate (:281-330): configuration ranges, rheology
 * parameters, finite particle history, positive mass and volume, material
 * ids inside the table. */
int mpm_soil_validate(const mpm_soil_meta* meta, const uint32_t* words);
/* validate_lattice_window (mpm_soil_lattice_window.rs:6-39, with
 * validate_node_window, soil_candidates_gpu.rs:68-94): every particle's
 * stencil lies inside the integer grid window; checked on restore of a
 * non-zero window. */
int mpm_soil_validate_window(const mpm_soil_meta* meta, const uint32_t* words);
/* The digest primitive (blake3::hash, 32 bytes). */
int mpm_blake3(const void* data, uint64_t len, uint8_t out[32]);
/* Message of the calling thread's last codec failure. */
const char* mpm_soil_last_error(void);

/* Device side of checkpoint_soil_with (:388-431, with encode_checkpoint_copy
 * :352-369 and CheckpointMetadata::finish :111-133): the nine streams of all
 * `n` particles (n = the particle count) in slot order, one gather kernel and
 * one device-to-host copy.  Single-scene MPM_MODEL_SHADER handles only.
 * Fails with MPM_ERR_FAULT while a grid fault is latched, and after
 * mpm_reset_status() of a latched fault until a step or restore has run
 * (an incomplete state cannot be archived). */
int mpm_checkpoint_capture(mpm_sim* sim, uint32_t* words, uint32_t n);
/* Device side of restore_soil_with (:436-541: validate_soil_restore_origin,
 * encode_soil_restore and finish_soil_restore).  `meta` and `words` are a
 * decoded checkpoint; `rows` is its material table packed as material_row
 * (mpm.rs:666-686) does, n_rows = max(meta->n_materials, 1).  The call
 * validates the checkpoint (and its window when non-zero), checks that it
 * fits the handle's capacity and that every configuration field the handle
 * holds matches (grid size, cell size, origin, gravity, E, nu, wall
 * friction, particle mass and volume, time step), then replaces the particle
 * set, the material table and the grid window in one upload, clears the grid
 * and the fault latch and drops captured step graphs.  Nothing changes when
 * it fails.  Fields the handle does not hold (max_rigid, friction_angle,
 * cohesion, the config rheology) are compared by the caller, which owns the
 * MpmConfig.  Per-particle E, which the format does not carry, is reset to
 * the scene's E. */
int mpm_checkpoint_restore(mpm_sim* sim, const mpm_soil_meta* meta, const uint32_t* words,
                           const mpm_material* rows, uint32_t n_rows);

#ifdef __cplusplus
}
#endif
#endif /* MOLLA_MPM_H */
