/*
 * mpm_oracle.h -- CPU restatement of molla's MPM substep, used ONLY as the
 * checker for the HIP path (tests/, __graft_entry__.smoke(), bench.py's
 * cpu_baseline leg).  Nothing in molla_amd/ links, loads or calls this.
 *
 * Every function restates a reference function; the citation is on the
 * definition in mpm_oracle.cpp.  Matrices are row-major 3x3 (M[3*i+j] = M_ij).
 * Particle arrays are interleaved per particle: x[3n], v[3n], C[9n], F[9n].
 * Grid arrays are dense G^3, x fastest: cell = ix + G*(iy + G*iz).
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MATERIALS 8

/* Rheology ids (reference material_row, mpm.rs:666-686) plus a
 * build-defined weakly-compressible fluid (id 3, no reference counterpart). */
enum { ORC_SAND = 0, ORC_SNOW = 1, ORC_VISCOUS = 2, ORC_FLUID = 3 };

/* Constitutive/boundary semantics.
 * ORC_MODEL_SHADER : shaders/mpm.wgsl (fixed-corotated, 6 walls, plasticity)
 * ORC_MODEL_CPUDIFF: molla-autograd mpm_cpu_diff.rs (corotated-SVK, floor) */
enum { ORC_MODEL_SHADER = 0, ORC_MODEL_CPUDIFF = 1 };

typedef struct {
    float rheology;         /* 0 sand, 1 snow, 2 viscous, 3 fluid      */
    float dp_alpha;         /* Drucker-Prager alpha (sand)              */
    float theta_c, theta_s; /* snow envelope                            */
    float hardening;        /* snow hardening exponent                  */
    float velocity_damping; /* viscous, 1/s                             */
    float affine_damping;   /* viscous, 1/s                             */
    float bulk;             /* fluid bulk modulus (Pa)                  */
} orc_material;

typedef struct {
    uint32_t grid_size;
    float cell_size;
    float origin[3];
    int32_t window[3];
    float gravity_y;
    float youngs, poisson;
    float boundary_friction;
    uint32_t model;
    float floor_y, floor_friction; /* CPUDIFF model only */
    uint32_t n_materials;
    orc_material materials[ORC_MAX_MATERIALS];
    uint32_t jacobi_iters; /* MpmTimeStep::Implicit sweeps (0 = explicit), shader model */
} orc_scene;

/* Fault bits (reference g_kdiag tail metadata, mpm.wgsl:175-231). */
enum { ORC_FAULT_OVERFLOW = 1, ORC_FAULT_NUMERICAL = 2 };

/* ---- ordering (bit-exact integer work, f32 inputs) ---- */
void orc_tile_keys(const orc_scene* s, uint32_t n, uint32_t capacity, const float* x,
                   uint32_t* keys, uint32_t* ids);
void orc_sort_pairs(uint32_t n, uint32_t* keys, uint32_t* ids);
void orc_base_cells(const orc_scene* s, uint32_t n, const float* x, int32_t* base);

/* ---- small-matrix math (f64) ---- */
void orc_svd3_shader_f64(const double* F, double* U, double* sigma, double* V);
void orc_svd3_shader_f32(const float* F, float* U, float* sigma, float* V);
void orc_svd3_cpu(const double* F, double* U, double* sigma, double* V);
void orc_adj_svd3(const double* U, const double* sigma, const double* V, const double* adjU,
                  const double* adjS, const double* adjV, double* adjF);
void orc_pk1_fixed_corotated(const double* F, double mu, double lam, double* P);
void orc_pk1_corotated_svk(const double* F, double mu, double lam, double* P);
void orc_pk1_corotated_svk_adjoint(const double* F, double mu, double lam, const double* adjP,
                                   double* adjF, double* adj_mu, double* adj_lam);
void orc_quad_weights(double fx, double* w);
void orc_quad_weights_dfx(double fx, double* dw);
/* Plasticity return maps on singular values (mpm.wgsl:1355-1479). */
void orc_plastic_project(const orc_scene* s, const orc_material* m, const double* sigma,
                         double jp_in, double* sigma_out, double* jp_out);

/* Rigid bodies on the grid (mpm.wgsl:927-1006, 1174-1226): rows are 20
 * floats per body -- center + bounding radius, linear velocity, angular
 * velocity, unit quaternion xyzw, shape dims + kind (0 legacy sphere,
 * 1 sphere, 2 box, 3 cylinder along local y).  wrench: 6 doubles per body
 * (force, torque) of the last substep. */
int orc_step_rigid_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                       double* jp, const double* mass, const double* vol, const uint32_t* mat,
                       const double* E, double dt, uint32_t substeps, uint32_t* status, double* grid_m,
                       double* grid_p, const float* rows, uint32_t n_bodies, double* wrench);
int orc_step_rigid_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F,
                       float* jp, const float* mass, const float* vol, const uint32_t* mat,
                       const float* E, float dt, uint32_t substeps, uint32_t* status, float* grid_m,
                       float* grid_p, const float* rows, uint32_t n_bodies, double* wrench);
void orc_rigid_grid_f64(const orc_scene* s, double dt, const float* rows, uint32_t n_bodies, double* m,
                        double* p, double* v, uint8_t* mask, double* wrench);

/* Semi-implicit Jacobi sweeps (scene jacobi_iters) on a dense grid. */
void orc_implicit_sweeps(const orc_scene* s, double dt, const double* grid_m, double* grid_v);

/* ---- full substeps ---- */
/* Runs `substeps` substeps in place.  E may be NULL (scene youngs used).
 * grid_m / grid_p (G^3 and 3*G^3) receive the last substep's P2G result if
 * non-NULL.  status[0] = fault bits, status[1] = completed gathers. */
int orc_step_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                 double* jp, const double* mass, const double* vol, const uint32_t* mat,
                 const double* E, double dt, uint32_t substeps, uint32_t* status,
                 double* grid_m, double* grid_p);
int orc_step_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F,
                 float* jp, const float* mass, const float* vol, const uint32_t* mat,
                 const float* E, float dt, uint32_t substeps, uint32_t* status, float* grid_m,
                 float* grid_p);
/* The SHADER step on all host threads (OpenMP), f32: a timing port for the
 * CPU baseline whose sums run in another order than orc_step_f32's.
 * *threads receives the thread count. */
int orc_step_threads_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F,
                         float* jp, const float* mass, const float* vol, const uint32_t* mat,
                         const float* E, float dt, uint32_t substeps, uint32_t* status,
                         uint32_t* threads);
/* Plants one deliberate mistake in the rule it names (the Mutation enum in
 * mpm_oracle.cpp; 0 = none), so a test can show that its closed-form check
 * fails when that rule is perturbed.  Process-global; tests reset it to 0. */
void orc_set_mutation(int m);
/* The same threaded SHADER step in f64: a fast FP64 reference for long runs
 * of large scenes (sums in another order than orc_step_f64's). */
int orc_step_threads_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                         double* jp, const double* mass, const double* vol, const uint32_t* mat,
                         const double* E, double dt, uint32_t substeps, uint32_t* status,
                         uint32_t* threads);
/* P2G only (shader or cpudiff model per s->model), scatter order = particle order. */
int orc_p2g_f64(const orc_scene* s, uint32_t n, const double* x, const double* v, const double* C,
                const double* F, const double* jp, const double* mass, const double* vol,
                const uint32_t* mat, const double* E, double dt, double* grid_m, double* grid_p);
/* Deterministic gather-order P2G (mpm_deterministic_p2g.wgsl / mpm_gather_masks.wgsl):
 * per cell, 27 neighbour tiles in dx-fastest order, then ascending sorted index. */
int orc_p2g_gather_f32(const orc_scene* s, uint32_t n, const float* x, const float* v,
                       const float* C, const float* F, const float* jp, const float* mass,
                       const float* vol, const uint32_t* mat, const float* E, float dt,
                       float* grid_m, float* grid_p);

/* ---- differentiable CPU path (mpm_cpu_diff.rs / mpm_adjoint.rs) ---- */
/* Forward of `steps` substeps (CPUDIFF model).  E per particle (may be NULL).
 * If tape != NULL it must hold steps*n*(3+3+9+9) doubles: the pre-step
 * x, v, C, F of every step (row-major per particle). */
int orc_diff_forward(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                     const double* mass, const double* vol, const double* E, double dt,
                     uint32_t steps, double* tape);
/* Reverse pass over a tape from orc_diff_forward.  On entry gx, gv, gC, gF
 * hold dJ/d(final state); on return dJ/d(initial state).  gE (per particle,
 * may be NULL) accumulates dJ/dE_p; g_mu_lam[2] accumulates the global
 * (mu, lambda) gradients of the reference's StateAdjoint. */
int orc_diff_backward(const orc_scene* s, uint32_t n, const double* mass, const double* vol,
                      const double* E, double dt, uint32_t steps, const double* tape, double* gx,
                      double* gv, double* gC, double* gF, double* gE, double* g_mu_lam);

#ifdef __cplusplus
}
#endif
#endif
