// mpm_kernels.hip -- HIP kernels of the MPM substep for gfx950 (MI355X).
//
// The per-substep hot path is two launches:
//   k_grid_update  over the active-tile list (mpm.wgsl grid_update :1010-1058)
//   k_particles    G2P -> x/v/C/F update -> plasticity -> stress -> P2G, fused
//                  (mpm.wgsl g2p :1241-1353, plasticity :1355-1479,
//                  p2g/p2g_two_pass :494-925) so a particle's state is read
//                  once and written once per substep and the polar/SVD of F is
//                  computed once per substep (FP64).
// The P2G half reduces the block's scatter in LDS (fixed point: 32-bit for the
// shader model, 64-bit for the differentiable model) over
// the bounding box of its particles' stencils -- particles are kept sorted by
// 4^3 tile and cut into chunks that stay inside 4 or 8 tiles of one tile row,
// so that box is small -- and stores the box as the chunk's partial with
// plain stores.  The next grid update sums the partials of the chunks around
// each active tile in ascending chunk order.  Particles that moved beyond the
// 27 tiles around their sorted tile since the last sort scatter with global
// f32 atomics instead.
#include <algorithm>

#include "mpm_rigid.hpp"
#include "mpm_scatter.hpp"

namespace mpm {

// Kirchhoff-form stress of a particle from its SVD; returns the affine
// matrix A = m C + 4/dx^2 * (-dt vol U diag(tau) U^T).
template <int MODEL>
__device__ __forceinline__ void affine_from_svd(const SceneDev& sc, const float* mrow, const Svd32& d,
                                                float mu, float lam, float jp, float dt, float vol,
                                                float m, const float* C, float* A) {
    float tau[3];
    if (MODEL == MODEL_SHADER) {
        int rheo = int(mrow[0] + 0.5f);
        float h = 1.0f;
        if (rheo == RHEO_SNOW) h = __expf(clampf(-mrow[4] * jp, -3.0f, 3.0f));
        else if (rheo == RHEO_SAND) h = __expf(clampf(jp, -3.0f, 3.0f));
        float mu_h = mu * h, lam_h = lam * h;
        if (rheo == RHEO_FLUID) {
            mu_h = 0.0f;
            lam_h = mrow[7];
        }
        float s0 = clampf(d.s[0], 0.1f, 10.0f), s1 = clampf(d.s[1], 0.1f, 10.0f),
              s2 = clampf(d.s[2], 0.1f, 10.0f);
        float J = s0 * s1 * s2;
        float vj = lam_h * (J - 1.0f) * J;
        tau[0] = 2.0f * mu_h * (s0 - 1.0f) * s0 + vj;
        tau[1] = 2.0f * mu_h * (s1 - 1.0f) * s1 + vj;
        tau[2] = 2.0f * mu_h * (s2 - 1.0f) * s2 + vj;
    } else {
        float tr = d.s[0] + d.s[1] + d.s[2] - 3.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i) tau[i] = (2.0f * mu * (d.s[i] - 1.0f) + lam * tr) * d.s[i];
    }
    float S[9];
    udut(d.U, tau, S);
    const float k = -dt * vol * 4.0f * sc.inv_dx * sc.inv_dx;
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = m * C[i] + k * S[i];
}

// LDS for the scatter inputs of a thread's earlier particles of a chunk.
template <int N>
struct StashSpace {
    ScatterStash s[N - 1];
    __device__ __forceinline__ ScatterStash* ptr() { return s; }
};
template <>
struct StashSpace<1> {
    __device__ __forceinline__ ScatterStash* ptr() { return nullptr; }
};

// Three waves per SIMD (168 VGPRs) with three blocks per CU (LDS_CELLS):
// the pass is latency-bound, and this beats the unconstrained 2 waves/SIMD.
// SPARSE: the grid is a tile pool (one shader-model scene; a.tile_slot set).
// PPT: particles per thread (shader model, sorted chunks of up to 256 PPT
// particles): a thread takes positions t, t + 256, ... of the chunk in turn,
// and one block scatter reduces all of them, so the chunk's box, its partial
// and the grid update's work are shared by PPT times the particles.
template <int MODE, int MODEL, bool SPARSE, int PPT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_particles(StepArgs a) {
    // The differentiable model scatters in int64: its reverse pass divides by
    // grid masses, so edge cells need each contribution kept to f32 precision.
    constexpr bool WIDE = MODEL == MODEL_CPUDIFF;
    static_assert(PPT == 1 || MODEL == MODEL_SHADER, "several particles per thread: shader model only");
    __shared__ ScatterLdsT<WIDE> L;
    __shared__ StashSpace<MODE == MODE_G2P ? 1 : PPT> Lst;
    if (threadIdx.x == 0) L.box[15] = a.status[ST_FAULT] != 0u ? 1 : 0;
    if (SPARSE && threadIdx.x < 64) L.tmask[threadIdx.x] = 0u;
    __syncthreads();
    if (L.box[15]) return;  // block-uniform: a latched fault stops the pass
    if (MODE != MODE_P2G && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.status[ST_GATHERS], 1u);
    const uint32_t cap = a.cap;
    const float* ps = a.ps;
    float* po = a.ps_out;  // the updated x, v, C, F, jp (may alias ps)
    const float dt = a.dt;
    const uint32_t nch = a.n_chunks_dev ? *a.n_chunks_dev : a.n_chunks;
    for (ChunkWalk w = chunk_walk(nch); w.ci < w.end; w.ci += w.step) {
        const uint32_t ci = w.ci;
        const Chunk ch = a.chunks[ci];
        const SceneDev sc = a.scenes[ch.scene];
        // Gather staging: the chunk's P2G box of the previous pass holds every
        // stencil this gather reads (same particles, same positions), so the
        // block loads that box of grid velocities into LDS once and the 27
        // node reads of each particle come from LDS.  Particles outside the
        // box (escaped ones) read global memory.
        int4 gbox = make_int4(0, 0, 0, 0), gdim = make_int4(0, 0, 0, 0);
        const bool staged = MODE != MODE_P2G && stage_gather_box<WIDE, SPARSE>(a, sc, ci, L, gbox, gdim);
        WaveBounds pre;  // the wave's box and bounds over its earlier particles
        Scatter sp;
#pragma unroll 1
        for (int pq = 0; pq < PPT; ++pq) {
        const uint32_t slot = ch.start + threadIdx.x + uint32_t(pq) * BLOCK;
        const uint32_t boff = 4u * slot;
        const bool active = threadIdx.x + uint32_t(pq) * BLOCK < ch.count;
        float xs[3] = {0.f, 0.f, 0.f};  // the scattered position
        // every field written here, so nothing of the previous particle stays live
        sp.on = false;
        sp.esc = false;
        sp.slot = slot;
        sp.m = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sp.b[k] = 0;
            sp.fx[k] = 0.f;
            sp.mv[k] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) sp.A[k] = 0.f;
        if (active) {
            float x[3], F[9];
#pragma unroll
            for (int k = 0; k < 3; ++k) x[k] = plane_ld(ps, cap, CX + k, boff);
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = plane_ld(ps, cap, CF + k, boff);
            float jp = plane_ld(ps, cap, CJP, boff);
            const float m = plane_ld(ps, cap, CMASS, boff);
            const float vol = plane_ld(ps, cap, CVOL, boff);
            const uint32_t matid = __float_as_uint(plane_ld(ps, cap, CMAT, boff));
            const float* mrow = mat_row(a, sc, matid);
            float mu = sc.mu, lam = sc.lam;
            if (a.has_E) {
                float E = plane_ld(ps, cap, CE, boff);
                mu = E / (2.0f * (1.0f + sc.poisson));
                lam = E * sc.poisson / ((1.0f + sc.poisson) * (1.0f - 2.0f * sc.poisson));
            }
            bool ok = true;
            float v[3], C[9];
            // the slot's sorted tile key, for the scatter's escape test; read
            // before this particle's state stores, which the compiler cannot
            // move it past
            uint32_t skey = 0u;
            if (MODE == MODE_P2G) {
                if (a.use_lds) skey = a.sorted_keys[slot];
#pragma unroll
                for (int k = 0; k < 3; ++k) v[k] = plane_ld(ps, cap, CV + k, boff);
#pragma unroll
                for (int k = 0; k < 9; ++k) C[k] = plane_ld(ps, cap, CC + k, boff);
                if (MODEL == MODEL_SHADER) {
                    float fa = finite_acc(0.0f, jp);
#pragma unroll
                    for (int k = 0; k < 3; ++k) fa = finite_acc(finite_acc(fa, x[k]), v[k]);
#pragma unroll
                    for (int k = 0; k < 9; ++k) fa = finite_acc(finite_acc(fa, C[k]), F[k]);
                    if (!finite_acc_ok(fa)) {
                        latch(a.status, FAULT_NUMERICAL);
                        ok = false;
                    }
                }
                if (ok) {
                    Svd32 d;
                    svd3<MODEL>(F, d);
                    affine_from_svd<MODEL>(sc, mrow, d, mu, lam, jp, dt, vol, m, C, sp.A);
                    if (a.aff_out)
#pragma unroll
                        for (int k = 0; k < 9; ++k) plane_st(a.aff_out, cap, k, boff, sp.A[k]);
                }
            } else {
                // ---- G2P (mpm.wgsl:1241-1353 / mpm_cpu_diff.rs:544-580) ----
                float xp[3], fx[3], w[3][3];
                int b[3];
                float fa = finite_acc(0.0f, jp);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    xp[k] = (x[k] - sc.origin[k]) * sc.inv_dx;
                    fa = finite_acc(finite_acc(fa, x[k]), xp[k]);
                }
#pragma unroll
                for (int k = 0; k < 9; ++k) fa = finite_acc(fa, F[k]);
                if (MODEL == MODEL_SHADER && !finite_acc_ok(fa)) {
                    latch(a.status, FAULT_NUMERICAL);
                    ok = false;
                }
                if (ok) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        float bf = floorf(xp[k] - 0.5f);
                        b[k] = int(bf) - (MODEL == MODEL_SHADER ? sc.window[k] : 0);
                        fx[k] = xp[k] - bf;
                        quad_weights(fx[k], w[k]);
                    }
                    // v = sum w v_i and, with d = ((i, j, k) - fx) dx,
                    // C = 4/dx^2 sum w v_i d^T = 4/dx (M - v fx^T),
                    // M = sum w v_i (i, j, k)^T, accumulated per x-row.
                    float vn[3] = {0.f, 0.f, 0.f};
                    float M[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) M[k] = 0.f;
                    // fetch(i, j, k, g): node (b + (i, j, k)), false when outside the grid
                    // Row sums over i weighted by w0 (and by i w0), then one
                    // multiply by w1 w2 per row: 3-6 FMAs per node.
                    const float wi1 = w[0][1], wi2 = 2.0f * w[0][2];
                    auto gather = [&](auto fetch) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
#pragma unroll
                            for (int j = 0; j < 3; ++j) {
                                const float wjk = w[1][j] * w[2][k];
                                float rs[3] = {0.f, 0.f, 0.f}, ri[3] = {0.f, 0.f, 0.f};
#pragma unroll
                                for (int i = 0; i < 3; ++i) {
                                    float4 g;
                                    if (!fetch(i, j, k, g)) continue;
                                    rs[0] += w[0][i] * g.x;
                                    rs[1] += w[0][i] * g.y;
                                    rs[2] += w[0][i] * g.z;
                                    if (i > 0) {
                                        const float wi = i == 1 ? wi1 : wi2;
                                        ri[0] += wi * g.x;
                                        ri[1] += wi * g.y;
                                        ri[2] += wi * g.z;
                                    }
                                }
                                // j and k are unrolled constants: rows with j = 0 or
                                // k = 0 add nothing to those moments
                                const float wj = float(j) * wjk, wk = float(k) * wjk;
#pragma unroll
                                for (int r = 0; r < 3; ++r) {
                                    vn[r] += wjk * rs[r];
                                    M[3 * r] += wjk * ri[r];
                                    if (j > 0) M[3 * r + 1] += wj * rs[r];
                                    if (k > 0) M[3 * r + 2] += wk * rs[r];
                                }
                            }
                        }
                    };
                    const bool in_box = staged && b[0] >= gbox.x && b[1] >= gbox.y && b[2] >= gbox.z &&
                                        b[0] + 2 < gbox.x + gdim.x && b[1] + 2 < gbox.y + gdim.y &&
                                        b[2] + 2 < gbox.z + gdim.z;
                    if (in_box) {
                        const float4* stage = reinterpret_cast<const float4*>(&L.acc[0][0]);
                        const int li0 = (b[0] - gbox.x) + gdim.x * ((b[1] - gbox.y) + gdim.y * (b[2] - gbox.z));
                        gather([&](int i, int j, int k, float4& g) {
                            g = stage[li0 + i + gdim.x * (j + gdim.y * k)];
                            return true;
                        });
                    } else {
                        const int G = int(sc.G);
                        gather([&](int i, int j, int k, float4& g) {
                            const int ix = b[0] + i, iy = b[1] + j, iz = b[2] + k;
                            if (ix < 0 || ix >= G || iy < 0 || iy >= G || iz < 0 || iz >= G) return false;
                            uint32_t c;
                            if (!cell_storage(sc, SPARSE ? a.tile_slot : nullptr, ix, iy, iz, c)) return false;
                            g = a.gv[c];
                            return true;
                        });
                    }
                    const float dx = sc.dx, s4dx = 4.0f * sc.inv_dx;
                    float Cn[9];
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) Cn[3 * r + c] = s4dx * (M[3 * r + c] - vn[r] * fx[c]);
                    const int rheo = int(mrow[0] + 0.5f);
                    if (MODEL == MODEL_SHADER && rheo == RHEO_VISCOUS) {
                        float vd = __expf(-fmaxf(mrow[5], 0.0f) * dt);
                        float ad = __expf(-fmaxf(mrow[6], 0.0f) * dt);
#pragma unroll
                        for (int k = 0; k < 3; ++k) vn[k] *= vd;
#pragma unroll
                        for (int k = 0; k < 9; ++k) Cn[k] *= ad;
                    }
                    float nx[3];
                    float fa2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        nx[k] = x[k] + dt * vn[k];
                        fa2 = finite_acc(finite_acc(fa2, nx[k]), vn[k]);
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) fa2 = finite_acc(fa2, Cn[k]);
                    if (!finite_acc_ok(fa2)) {
                        latch(a.status, FAULT_NUMERICAL);
                        ok = false;
                    }
                    if (ok && MODEL == MODEL_SHADER) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            float lo = sc.origin[k] + float(sc.window[k] + 2) * dx;
                            float hi = sc.origin[k] + float(sc.window[k] + int(sc.G) - 3) * dx;
                            if (nx[k] < lo || nx[k] > hi) vn[k] = 0.0f;
                            nx[k] = fminf(fmaxf(nx[k], lo), hi);
                        }
                    }
                    float Fn[9];
                    if (ok) {
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int j = 0; j < 3; ++j)
                                Fn[3 * i + j] = F[3 * i + j] + dt * (Cn[3 * i] * F[j] + Cn[3 * i + 1] * F[3 + j] +
                                                                     Cn[3 * i + 2] * F[6 + j]);
                        float fa3 = 0.0f;
#pragma unroll
                        for (int k = 0; k < 9; ++k) fa3 = finite_acc(fa3, Fn[k]);
                        if (!finite_acc_ok(fa3)) ok = false;
                        if (!ok) latch(a.status, FAULT_NUMERICAL);
                    }
                    if (ok) {
                        // ---- plasticity (mpm.wgsl:1355-1479) + stress from the same SVD ----
                        Svd32 d;
                        bool have_svd = false;
                        bool plast_ok = true;
                        float jp_new = jp;
                        if (MODEL == MODEL_SHADER) {
                            if (rheo == RHEO_FLUID) {
                                float J = Fn[0] * (Fn[4] * Fn[8] - Fn[5] * Fn[7]) -
                                          Fn[1] * (Fn[3] * Fn[8] - Fn[5] * Fn[6]) +
                                          Fn[2] * (Fn[3] * Fn[7] - Fn[4] * Fn[6]);
                                float s = clampf(cbrtf(J), 0.05f, 20.0f);
                                if (!finitef(s)) {
                                    plast_ok = false;
                                } else {
#pragma unroll
                                    for (int k = 0; k < 9; ++k) Fn[k] = (k % 4 == 0) ? s : 0.0f;
                                    jp_new = 0.0f;
                                    d.s[0] = d.s[1] = d.s[2] = s;
#pragma unroll
                                    for (int k = 0; k < 9; ++k) d.U[k] = d.V[k] = (k % 4 == 0) ? 1.0f : 0.0f;
                                    have_svd = true;
                                }
                            } else {
                                svd3<MODEL_SHADER>(Fn, d);
                                float ns[3] = {d.s[0], d.s[1], d.s[2]};
                                float dq = 0.0f;
                                bool jp_replace = false;
                                bool fin3 = finitef(ns[0]) && finitef(ns[1]) && finitef(ns[2]);
                                if (!fin3) plast_ok = false;
                                if (plast_ok && rheo == RHEO_SNOW) {
                                    const float tc = mrow[2], tsv = mrow[3];
#pragma unroll
                                    for (int k = 0; k < 3; ++k) ns[k] = clampf(d.s[k], 1.0f - tc, 1.0f + tsv);
                                    float det_old = fmaxf(d.s[0] * d.s[1] * d.s[2], 1.0e-6f);
                                    float det_new = fmaxf(ns[0] * ns[1] * ns[2], 1.0e-6f);
                                    jp_replace = true;
                                    jp_new = jp + (__logf(det_old) - __logf(det_new));
                                    if (!finitef(jp_new)) plast_ok = false;
                                    jp_new = clampf(jp_new, -5.0f, 5.0f);
                                } else if (plast_ok && rheo == RHEO_VISCOUS) {
                                    jp_replace = true;
                                    jp_new = 0.0f;
                                } else if (plast_ok) {
                                    // sand (Drucker-Prager), mpm.wgsl:1421-1457
                                    const float e0 = __logf(fmaxf(d.s[0], 1.0e-6f)),
                                                e1 = __logf(fmaxf(d.s[1], 1.0e-6f)),
                                                e2 = __logf(fmaxf(d.s[2], 1.0e-6f));
                                    float tre = e0 + e1 + e2, mean = tre / 3.0f;
                                    float d0 = e0 - mean, d1 = e1 - mean, d2 = e2 - mean;
                                    float dn = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
                                    float kappa = (3.0f * lam + 2.0f * mu) / (2.0f * mu);
                                    if (tre > 0.0f) {
                                        ns[0] = ns[1] = ns[2] = 1.0f;
                                        dq = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
                                    } else if (dn != 0.0f) {
                                        float fy = dn + kappa * tre * mrow[1];
                                        if (fy > 0.0f) {
                                            float rc = fmaxf(0.0f, -mrow[1] * kappa * tre);
                                            float sc2 = rc / dn;
                                            ns[0] = __expf(d0 * sc2 + mean);
                                            ns[1] = __expf(d1 * sc2 + mean);
                                            ns[2] = __expf(d2 * sc2 + mean);
                                            dq = fy;
                                        }
                                    }
                                }
                                if (plast_ok && (!finitef(ns[0]) || !finitef(ns[1]) || !finitef(ns[2]) ||
                                                 !finitef(dq) || !finitef(jp + dq)))
                                    plast_ok = false;
                                if (plast_ok) {
                                    float ss[3];
                                    bool same = true;
#pragma unroll
                                    for (int k = 0; k < 3; ++k) {
                                        ss[k] = clampf(ns[k], 0.05f, 20.0f);
                                        same = same && (ss[k] == d.s[k]);
                                    }
                                    if (!same) {
                                        udvt(d.U, ss, d.V, Fn);
                                        d.s[0] = ss[0];
                                        d.s[1] = ss[1];
                                        d.s[2] = ss[2];
                                    }
                                    jp_new = jp_replace ? jp_new : clampf(jp + dq, 0.0f, 5.0f);
                                    have_svd = true;
                                }
                            }
                            if (!plast_ok) latch(a.status, FAULT_NUMERICAL);
                        }
                        if (MODE != MODE_G2P && a.use_lds) skey = a.sorted_keys[slot];
                        // ---- write the particle state ----
#pragma unroll
                        for (int k = 0; k < 3; ++k) plane_st(po, cap, CX + k, boff, nx[k]);
                        if (a.write_vc || MODE == MODE_G2P) {
#pragma unroll
                            for (int k = 0; k < 3; ++k) plane_st(po, cap, CV + k, boff, vn[k]);
#pragma unroll
                            for (int k = 0; k < 9; ++k) plane_st(po, cap, CC + k, boff, Cn[k]);
                        }
#pragma unroll
                        for (int k = 0; k < 9; ++k) plane_st(po, cap, CF + k, boff, Fn[k]);
                        if (MODEL == MODEL_SHADER && plast_ok) plane_st(po, cap, CJP, boff, jp_new);
                        if (MODE == MODE_G2P2G && plast_ok) {
                            if (MODEL == MODEL_CPUDIFF) svd3<MODEL_CPUDIFF>(Fn, d);
                            (void)have_svd;
                            affine_from_svd<MODEL>(sc, mrow, d, mu, lam, jp_new, dt, vol, m, Cn, sp.A);
                            if (a.aff_out)
#pragma unroll
                                for (int k = 0; k < 9; ++k) plane_st(a.aff_out, cap, k, boff, sp.A[k]);
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                x[k] = nx[k];
                                v[k] = vn[k];
                            }
                        } else {
                            ok = false;
                        }
                    }
                }
            }
            if (ok && MODE != MODE_G2P) {
                stencil_of<MODEL>(sc, x, sp.b, sp.fx);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    sp.mv[k] = m * v[k];
                    xs[k] = x[k];
                }
                sp.m = m;
                sp.on = true;
                sp.esc = !a.use_lds;
                if (a.use_lds && stencil_escapes(sc, ch, skey, sp.b)) sp.esc = true;
            }
        }
        // the earlier particles wait in LDS; the last one scatters them all
        if (MODE != MODE_G2P && PPT > 1 && pq < PPT - 1) {
            Lst.ptr()[pq].put(sp, xs);
            wave_bounds_add(sp, sc.dx, pre);
        }
        }
        if (MODE != MODE_G2P) block_scatter<true, WIDE, SPARSE, PPT>(a, sc, sp, ci, L, Lst.ptr(), &pre);
        else if (staged) __syncthreads();  // the next chunk's staging overwrites the box
    }
}

hipError_t launch_particle_pass(int mode, int model, const StepArgs& a, int ppt, int blocks,
                                hipStream_t s) {
    if (ppt < 1 || ppt > MAX_PPT || (ppt > 1 && model != MODEL_SHADER)) return hipErrorInvalidValue;
    if (a.n_chunks == 0) return hipSuccess;
    if (blocks <= 0 || uint32_t(blocks) > a.n_chunks) blocks = int(a.n_chunks);
    if (blocks > int(NUM_XCD)) blocks &= ~int(NUM_XCD - 1);  // XCD-aware chunk walk
    blocks = std::max(blocks, 1);
    dim3 g(blocks), b(BLOCK);
#define MPM_LAUNCH(M, D, S) hipLaunchKernelGGL((k_particles<M, D, S, 1>), g, b, 0, s, a)
#define MPM_LAUNCH2(M, S) hipLaunchKernelGGL((k_particles<M, MODEL_SHADER, S, 2>), g, b, 0, s, a)
    if (ppt == 2) {
        if (mode == MODE_P2G) {
            if (a.tile_slot) MPM_LAUNCH2(MODE_P2G, true);
            else MPM_LAUNCH2(MODE_P2G, false);
        } else if (mode == MODE_G2P) {
            if (a.tile_slot) MPM_LAUNCH2(MODE_G2P, true);
            else MPM_LAUNCH2(MODE_G2P, false);
        } else {
            if (a.tile_slot) MPM_LAUNCH2(MODE_G2P2G, true);
            else MPM_LAUNCH2(MODE_G2P2G, false);
        }
    } else if (model == MODEL_SHADER && a.tile_slot) {
        if (mode == MODE_P2G) MPM_LAUNCH(MODE_P2G, MODEL_SHADER, true);
        else if (mode == MODE_G2P) MPM_LAUNCH(MODE_G2P, MODEL_SHADER, true);
        else MPM_LAUNCH(MODE_G2P2G, MODEL_SHADER, true);
    } else if (model == MODEL_SHADER) {
        if (mode == MODE_P2G) MPM_LAUNCH(MODE_P2G, MODEL_SHADER, false);
        else if (mode == MODE_G2P) MPM_LAUNCH(MODE_G2P, MODEL_SHADER, false);
        else MPM_LAUNCH(MODE_G2P2G, MODEL_SHADER, false);
    } else {
        if (a.tile_slot) return hipErrorInvalidValue;  // pools hold shader-model scenes
        if (mode == MODE_P2G) MPM_LAUNCH(MODE_P2G, MODEL_CPUDIFF, false);
        else if (mode == MODE_G2P) MPM_LAUNCH(MODE_G2P, MODEL_CPUDIFF, false);
        else MPM_LAUNCH(MODE_G2P2G, MODEL_CPUDIFF, false);
    }
#undef MPM_LAUNCH
#undef MPM_LAUNCH2
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grid update over the active tiles (mpm.wgsl:1010-1058; CPUDIFF model:
// mpm_cpu_diff.rs:513-542), fused with the reduction of the P2G partials:
// one wave per tile, one lane per cell.  The tiles are the static list of the
// last sort followed by the dynamic list of tiles reached by global atomics.
// The chunks that may hold partials for tile T are those of the particles
// sorted into T's 3x3x3 neighbourhood: nine x-rows of consecutive tile keys,
// i.e. nine ascending slot ranges read from the tile lower-bound table.
__device__ __forceinline__ void cell_update(const SceneDev& sc, float4 q, int ix, int iy, int iz, float dt,
                                            float4& out) {
    const float m = q.x;
    out = make_float4(0.f, 0.f, 0.f, m);
    const int n_ = int(sc.G);
    if (sc.model == MODEL_SHADER) {
        if (m > 1.0e-9f && m == m) {
            float vx = q.y / m, vy = q.z / m, vz = q.w / m;
            vy = vy + dt * sc.gravity_y;
            const float bf = sc.boundary_friction;
            const int pad = 2;
            if (ix < pad) { vx = fmaxf(vx, 0.f); vy *= bf; vz *= bf; }
            if (ix >= n_ - pad) { vx = fminf(vx, 0.f); vy *= bf; vz *= bf; }
            if (iy < pad) { vy = fmaxf(vy, 0.f); vx *= bf; vz *= bf; }
            if (iy >= n_ - pad) { vy = fminf(vy, 0.f); vx *= bf; vz *= bf; }
            if (iz < pad) { vz = fmaxf(vz, 0.f); vx *= bf; vy *= bf; }
            if (iz >= n_ - pad) { vz = fminf(vz, 0.f); vx *= bf; vy *= bf; }
            out = make_float4(vx, vy, vz, m);
        }
    } else {
        if (m > 0.0f) {
            float vx = q.y / m, vy = q.z / m, vz = q.w / m;
            vy += sc.gravity_y * dt;
            if (below_floor(sc, iy)) {
                if (vy < 0.f) vy = 0.f;
                const float damp = 1.0f - sc.floor_friction;
                vx *= damp;
                vz *= damp;
            }
            out = make_float4(vx, vy, vz, m);
        }
    }
}

// RIGID: bodies are present (the imprint code stays out of the plain kernel).
// SPLIT = false: one wave per tile sums its candidate chunks' partials in
// ascending order (scenes with many more active tiles than blocks).
// SPLIT = true: the block's four waves sum four ranges of the candidate list
// and wave 0 adds the range sums in order, so the latency of the ordered
// partial loads is shared four ways (small scenes: a few hundred tiles, each
// overlapped by tens of chunks).  The host picks the shape from the particle
// count and the launch size only, so a run repeats bitwise; the two shapes
// agree to f32 rounding.
constexpr int GRID_RANGES = 4;
// One wave per tile runs at eight waves per SIMD with four partial loads in
// flight (51 VGPRs, the tile's record and scene in scalar registers): with
// ten thousand tiles or more, more tiles in flight beat deeper load batches.
// The split shape keeps eight loads in flight (dense small scenes have ~100
// overlapping chunks per tile).
template <bool RIGID, bool SPLIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RIGID ? 4 : (SPLIT ? 5 : 8)))) void k_grid_update(
    const SceneDev* __restrict__ scenes, GridDev g,
                                                     const uint32_t* __restrict__ dlist,
                                                     const uint32_t* __restrict__ dcount,
                                                     uint32_t* __restrict__ next_count, float dt,
                                                     int zero_after, int use_partials, float4* keep, RigidDev rg) {
    __shared__ float4 wsum[SPLIT ? GRID_RANGES - 1 : 1][64];
    const TileWalk walk(g, dcount);
    const int lane = threadIdx.x & 63, wave = wave_id();
    const uint32_t first = SPLIT ? blockIdx.x : blockIdx.x * 4 + uint32_t(wave);
    const uint32_t stride = SPLIT ? gridDim.x : gridDim.x * 4;
    // A static tile's record holds its id, scene and candidate ranges; the
    // next tile's record is requested before this tile's stores (which the
    // compiler cannot move it past), so a wave that takes several tiles
    // finds it loaded.
    TileRec cur, nxt;
    if (first < walk.n) load_tile_rec(g, walk.ns, first, cur);
    for (uint32_t i = first; i < walk.n; i += stride) {
        nxt.valid = false;
        if (i + stride < walk.n) load_tile_rec(g, walk.ns, i + stride, nxt);
        uint32_t t, st, scene;
        if (cur.valid) {
            t = cur.t;
            st = t;
            scene = cur.scene;
        } else {
            walk.at(g, dlist, i, t, st);
            scene = g.tile_scene[t];
        }
        const SceneDev sc = scenes[scene];
        int tx, ty, tz, ix, iy, iz;
        tile_cell(sc, t, lane, tx, ty, tz, ix, iy, iz);
        const uint32_t c = (st << 6) | uint32_t(lane);
        const uint32_t flag = g.tile_flag[t];
        // range 0 starts from the atomic channel
        float4 q = (flag & FLAG_GMP) && (!SPLIT || wave == 0) ? g.gmp[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (use_partials) {
            add_partials<(SPLIT || RIGID) ? 8 : 4>(g, sc, tx, ty, tz, ix, iy, iz, q, SPLIT ? wave : 0,
                                                   SPLIT ? GRID_RANGES : 1, cur.valid, cur.rng);
            if (SPLIT) {
                if (wave > 0) wsum[wave - 1][lane] = q;
                __syncthreads();
                if (wave == 0) {
#pragma unroll
                    for (int r = 0; r < GRID_RANGES - 1; ++r) {
                        const float4 v = wsum[r][lane];
                        q.x += v.x;
                        q.y += v.y;
                        q.z += v.z;
                        q.w += v.w;
                    }
                }
            }
        }
        if (!SPLIT || wave == 0) {
            float4 out;
            int owner = -1;
            if (RIGID && rg.count) {
                // rigid_imprint (mpm.wgsl:937-1006) between the P2G and the update:
                // a stamped cell keeps the body's velocity (grid_update :1034-1038)
                float3 force, torque;
                owner = in_grid(sc, ix, iy, iz) ? rigid_imprint(rg, sc, ix, iy, iz, dt, q, force, torque) : -1;
                if (rg.wrench) rigid_wrench_add(rg, owner, force, torque);
            }
            if (owner >= 0) {
                out = make_float4(q.y / q.x, q.z / q.x, q.w / q.x, q.x);
            } else {
                cell_update(sc, q, ix, iy, iz, dt, out);
            }
            g.gv[c] = out;
            if (keep) keep[c] = q;  // the P2G total, for the reverse pass
            if (zero_after) {
                if (flag & FLAG_GMP) g.gmp[c] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (lane == 0) g.tile_flag[t] = flag & FLAG_STATIC;
            } else {
                g.gmp[c] = q;  // the last P2G stays readable (read_grid_mass_momentum)
                if (lane == 0) g.tile_flag[t] = flag | FLAG_GMP;
            }
        }
        if (SPLIT && use_partials) __syncthreads();  // wsum is reused by the next tile
        cur = nxt;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && next_count) *next_count = 0u;
}

hipError_t launch_grid_update(const SceneDev* scenes, const GridDev& g, const uint32_t* dlist,
                              const uint32_t* dcount, uint32_t* next_count, float dt,
                              int zero_after, int use_partials, int blocks, hipStream_t s, float4* keep,
                              const RigidDev* rigid, int split) {
    const RigidDev rg = rigid ? *rigid : RigidDev{nullptr, nullptr, 0u, 0u};
#define MPM_GRID(R, S)                                                                                          \
    hipLaunchKernelGGL((k_grid_update<R, S>), dim3(blocks), dim3(256), 0, s, scenes, g, dlist, dcount, next_count, \
                       dt, zero_after, use_partials, keep, rg)
    if (rg.count) {
        if (split) MPM_GRID(true, true);
        else MPM_GRID(true, false);
    } else {
        if (split) MPM_GRID(false, true);
        else MPM_GRID(false, false);
    }
#undef MPM_GRID
    return hipGetLastError();
}

// Shell force of legacy rigid spheres (mpm_rigid_force_integrate,
// mpm.wgsl:1174-1226) after the grid update and the implicit sweeps: every
// unstamped cell with sand mass within one cell of a sphere's surface (full
// weight inside, tapering to zero one cell out) pushes on the sphere with
// -(m w / dt) (v_sphere - v_cell).  One wave per active tile.
__global__ __launch_bounds__(256) void k_rigid_shell(const SceneDev* __restrict__ scenes, GridDev g,
                                                     const uint32_t* __restrict__ dlist,
                                                     const uint32_t* __restrict__ dcount, float dt, RigidDev rg) {
    const TileWalk walk(g, dcount);
    const int lane = threadIdx.x & 63;
    for (uint32_t i = blockIdx.x * 4 + uint32_t(wave_id()); i < walk.n; i += gridDim.x * 4) {
        uint32_t t, st;
        walk.at(g, dlist, i, t, st);
        const SceneDev sc = scenes[g.tile_scene[t]];
        int tx, ty, tz, ix, iy, iz;
        tile_cell(sc, t, lane, tx, ty, tz, ix, iy, iz);
        const float4 cv = g.gv[(st << 6) | uint32_t(lane)];
        const bool live = in_grid(sc, ix, iy, iz) && cv.w > 1.0e-9f && cv.w == cv.w &&
                          !rigid_masked_legacy(rg, sc, ix, iy, iz);
        const float3 node = node_position(sc, ix, iy, iz);
        for (uint32_t r = 0; r < rg.count; ++r) {
            const float4* b = rg.body + RIGID_ROWS * r;
            const float3 arm = make_float3(node.x - b[0].x, node.y - b[0].y, node.z - b[0].z);
            const float d = sqrtf(arm.x * arm.x + arm.y * arm.y + arm.z * arm.z);
            const bool near = live && b[4].w == 0.0f && !(d > b[0].w + sc.dx);
            float3 force = make_float3(0.f, 0.f, 0.f), torque = force;
            if (near) {
                float w = 1.0f;
                if (d > b[0].w) w = fmaxf(0.0f, 1.0f - (d - b[0].w) / sc.dx);
                const float k = -(cv.w * w / fmaxf(dt, 1.0e-6f));
                force = make_float3(k * (b[1].x - cv.x), k * (b[1].y - cv.y), k * (b[1].z - cv.z));
                torque = cross3(arm, force);
            }
            if (__ballot(near)) rigid_wrench_add(rg, near ? int(r) : -1, force, torque);
        }
    }
}

hipError_t launch_rigid_shell(const SceneDev* scenes, const GridDev& g, const uint32_t* dlist,
                              const uint32_t* dcount, float dt, const RigidDev& rg, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_rigid_shell, dim3(blocks), dim3(256), 0, s, scenes, g, dlist, dcount, dt, rg);
    return hipGetLastError();
}

__global__ void k_zero_f64(double* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

hipError_t launch_zero_f64(double* p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Zero the accumulation grid of the static and dynamic lists; keep only the
// static bit.
// ---------------------------------------------------------------------------
// Semi-implicit Jacobi sweep (mpm_implicit_apply, mpm.wgsl:1060-1153) over the
// active tiles, one wave per tile, one lane per cell, reading `src` and
// writing `dst` (the host ping-pongs, mpm_implicit_swap):
//   v' = (m v + a v_avg) / (m + a),  a = dt^2 m (E / 1500) / dx^2,
// v_avg = mean velocity of the face neighbours with m > 1e-9 (v if none);
// empty cells are copied.  A neighbour in a tile that is not on this
// substep's lists (neither FLAG_STATIC nor FLAG_DYN) holds no P2G mass, as
// in the reference's sparse grid, whatever its stale contents.
__global__ __launch_bounds__(256) void k_implicit_sweep(const SceneDev* __restrict__ scenes, GridDev g,
                                                        const float4* __restrict__ src, float4* __restrict__ dst,
                                                        const uint32_t* __restrict__ dlist,
                                                        const uint32_t* __restrict__ dcount, float dt, RigidDev rg) {
    const TileWalk walk(g, dcount);
    const int lane = threadIdx.x & 63;
    for (uint32_t i = blockIdx.x * 4 + uint32_t(wave_id()); i < walk.n; i += gridDim.x * 4) {
        uint32_t t, st;
        walk.at(g, dlist, i, t, st);
        const SceneDev sc = scenes[g.tile_scene[t]];
        int tx, ty, tz, ix, iy, iz;
        tile_cell(sc, t, lane, tx, ty, tz, ix, iy, iz);
        const uint32_t c = (st << 6) | uint32_t(lane);
        const float4 me = src[c];
        const float m = me.w;
        // empty cells and cells stamped by a legacy sphere keep their velocity
        if (m <= 1.0e-9f || m != m || rigid_masked_legacy(rg, sc, ix, iy, iz)) {
            dst[c] = me;
            continue;
        }
        const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        float ax = 0.f, ay = 0.f, az = 0.f, cnt = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int jx = ix + off[k][0], jy = iy + off[k][1], jz = iz + off[k][2];
            if (!in_grid(sc, jx, jy, jz)) continue;
            uint32_t nc;
            if (g.tile_slot) {
                if (!cell_storage(sc, g.tile_slot, jx, jy, jz, nc)) continue;  // no slot: no P2G mass
            } else {
                nc = cell_index(sc, jx, jy, jz);
                if ((g.tile_flag[nc >> 6] & (FLAG_STATIC | FLAG_DYN)) == 0u) continue;
            }
            const float4 nb = src[nc];
            if (nb.w <= 1.0e-9f || nb.w != nb.w) continue;
            ax += nb.x;
            ay += nb.y;
            az += nb.z;
            cnt += 1.0f;
        }
        if (cnt > 0.f) {
            ax = ax / cnt;
            ay = ay / cnt;
            az = az / cnt;
        } else {
            ax = me.x;
            ay = me.y;
            az = me.z;
        }
        const float alpha = dt * dt * m * (sc.youngs / 1500.0f) * sc.inv_dx * sc.inv_dx;
        const float den = m + alpha;
        float vx = (m * me.x + alpha * ax) / den, vy = (m * me.y + alpha * ay) / den,
              vz = (m * me.z + alpha * az) / den;
        if (vx != vx || vy != vy || vz != vz) {
            vx = me.x;
            vy = me.y;
            vz = me.z;
        }
        dst[c] = make_float4(vx, vy, vz, m);
    }
}

__global__ __launch_bounds__(256) void k_copy_tiles(GridDev g, const float4* __restrict__ src, float4* __restrict__ dst,
                                                    const uint32_t* __restrict__ dlist,
                                                    const uint32_t* __restrict__ dcount) {
    const TileWalk walk(g, dcount);
    const int lane = threadIdx.x & 63;
    for (uint32_t i = blockIdx.x * 4 + uint32_t(wave_id()); i < walk.n; i += gridDim.x * 4) {
        uint32_t t, st;
        walk.at(g, dlist, i, t, st);
        dst[(st << 6) | uint32_t(lane)] = src[(st << 6) | uint32_t(lane)];
    }
}

hipError_t launch_implicit_sweeps(const SceneDev* scenes, const GridDev& g, float4* scratch, uint32_t iters,
                                  const uint32_t* dlist, const uint32_t* dcount, float dt, int blocks,
                                  hipStream_t s, const RigidDev* rigid) {
    const RigidDev rg = rigid ? *rigid : RigidDev{nullptr, nullptr, 0u, 0u};
    if (iters == 0) return hipSuccess;
    // odd sweep counts start from a copy so the result lands in g.gv
    float4* a = g.gv;
    float4* b = scratch;
    if (iters & 1u) {
        hipLaunchKernelGGL(k_copy_tiles, dim3(blocks), dim3(256), 0, s, g, g.gv, scratch, dlist, dcount);
        a = scratch;
        b = g.gv;
    }
    for (uint32_t it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_implicit_sweep, dim3(blocks), dim3(256), 0, s, scenes, g, a, b, dlist, dcount, dt, rg);
        float4* t = a;
        a = b;
        b = t;
    }
    return hipGetLastError();
}

// `also` (may be null) is a second per-cell buffer cleared on the same tiles.
__global__ __launch_bounds__(256) void k_clear_tiles(GridDev g, const uint32_t* __restrict__ dlist,
                                                     const uint32_t* __restrict__ dcount, float4* also) {
    const TileWalk walk(g, dcount);
    const int lane = threadIdx.x & 63;
    for (uint32_t i = blockIdx.x * 4 + uint32_t(wave_id()); i < walk.n; i += gridDim.x * 4) {
        uint32_t t, st;
        walk.at(g, dlist, i, t, st);
        g.gmp[(st << 6) | uint32_t(lane)] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (also) also[(st << 6) | uint32_t(lane)] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane == 0) g.tile_flag[t] &= FLAG_STATIC;
    }
}

// Resets a counter after every block of the previous launch has read it.
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

// Status bookkeeping of the differentiable path without host round trips:
// the escape and fallback counters into dst[0..2).
__global__ void k_snap_counters(const uint32_t* status, uint32_t* dst) {
    dst[0] = status[ST_ESCAPES];
    dst[1] = status[ST_FALLBACK];
}

hipError_t launch_snap_counters(const uint32_t* status, uint32_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_snap_counters, dim3(1), dim3(1), 0, s, status, dst);
    return hipGetLastError();
}

// Sum of the per-chunk (dJ/dmu, dJ/dlambda) pairs v[2i], v[2i+1], i < *count
// (clamped to cap), in a fixed order: thread t adds entries t, t+256, ...
// in sequence, then a fixed tree combines the 256 partials.  One block.
__global__ __launch_bounds__(256) void k_sum_pairs(const double* __restrict__ v, const uint32_t* count, uint32_t cap,
                                                   double* __restrict__ out) {
    __shared__ double sm[2][256];
    const uint32_t n = min(*count, cap);
    double a = 0.0, b = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        a += v[2 * (size_t)i];
        b += v[2 * (size_t)i + 1];
    }
    sm[0][threadIdx.x] = a;
    sm[1][threadIdx.x] = b;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + h];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0][0];
        out[1] = sm[1][0];
    }
}

hipError_t launch_sum_pairs(const double* v, const uint32_t* count, uint32_t cap, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, s, v, count, cap, out);
    return hipGetLastError();
}

hipError_t launch_clear_tiles(const GridDev& g, const uint32_t* dlist, uint32_t* dcount, int blocks,
                              hipStream_t s, float4* also) {
    hipLaunchKernelGGL(k_clear_tiles, dim3(blocks), dim3(256), 0, s, g, dlist, dcount, also);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, dcount, 0u);
    return hipGetLastError();
}

// ---- static active list of a sort ----
// tile_lb[k] = first sorted slot whose key is >= k, for k in [0, total].
__global__ void k_tile_lb(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ lb,
                          uint32_t total) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= total; k += gridDim.x * blockDim.x) {
        uint32_t first = 0, last = n;
        while (first < last) {
            const uint32_t mid = first + (last - first) / 2;
            if (keys[mid] < k) first = mid + 1;
            else last = mid;
        }
        lb[k] = first;
    }
}

// A tile is on the static list when any tile within one tile of it (each
// axis, same scene) holds a particle: P2G can reach it before the next sort.
// Occupancy of a 3-tile x-run is one difference of the lower-bound table.
__global__ void k_static_flags(const SceneDev* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                               const uint32_t* __restrict__ lb, uint32_t total, uint32_t* __restrict__ flag,
                               uint32_t* __restrict__ mark) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const SceneDev& sc = scenes[tile_scene[t]];
        const uint32_t tl = t - sc.tile_base;
        const int W = int(sc.W);
        const int tx = int(tl % sc.W), ty = int((tl / sc.W) % sc.W), tz = int(tl / (sc.W * sc.W));
        const int xa = max(tx - 1, 0), xb = min(tx + 1, W - 1);
        bool on = false;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy) {
                const int ny = ty + dy, nz = tz + dz;
                if (ny < 0 || nz < 0 || ny >= W || nz >= W) continue;
                const uint32_t row = sc.tile_base + uint32_t(W * (ny + W * nz));
                on = on || lb[row + xb + 1] > lb[row + xa];
            }
        mark[t] = on ? 1u : 0u;
        flag[t] = (flag[t] & ~FLAG_STATIC) | (on ? FLAG_STATIC : 0u);
    }
}

__global__ void k_static_emit(const uint32_t* __restrict__ mark, const uint32_t* __restrict__ pos, uint32_t total,
                              uint32_t* __restrict__ slist, uint32_t* __restrict__ scount) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        if (mark[t]) slist[pos[t]] = t;
        if (t + 1 == total) *scount = pos[t] + mark[t];
    }
}

hipError_t launch_static_rebuild(const SceneDev* scenes, const uint32_t* tile_scene,
                                 uint32_t* tile_flag, uint32_t* slist, uint32_t* scount,
                                 const uint32_t* sorted_keys, uint32_t n, uint32_t* tile_lb,
                                 uint32_t total_tiles, uint32_t* mark, uint32_t* pos, uint32_t* scan_scratch,
                                 hipStream_t s) {
    unsigned lbb = std::min<unsigned>((total_tiles + 256) / 256, 16384u);
    hipLaunchKernelGGL(k_tile_lb, dim3(lbb), dim3(256), 0, s, sorted_keys, n, tile_lb, total_tiles);
    const unsigned tb = std::max(1u, std::min<unsigned>((total_tiles + 255) / 256, 16384u));
    hipLaunchKernelGGL(k_static_flags, dim3(tb), dim3(256), 0, s, scenes, tile_scene, tile_lb, total_tiles, tile_flag,
                       mark);
    hipError_t e = exclusive_scan_u32(mark, pos, total_tiles, scan_scratch, nullptr, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_static_emit, dim3(tb), dim3(256), 0, s, mark, pos, total_tiles, slist, scount);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// One thread per (static tile, tile row): the row's candidate chunk range, as
// add_partials would find it from the tile bounds and the chunk of each slot.
__global__ void k_static_records(const SceneDev* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                                 const uint32_t* __restrict__ slist, const uint32_t* __restrict__ scount,
                                 uint32_t max_static, const uint32_t* __restrict__ lb,
                                 const uint32_t* __restrict__ chunk_of_slot, uint32_t part_capacity,
                                 int* __restrict__ srec) {
    const uint32_t ns = min(*scount, max_static);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < ns * 9u; j += gridDim.x * blockDim.x) {
        const uint32_t i = j / 9u, r = j % 9u;
        const uint32_t t = slist[i], scene = tile_scene[t];
        const SceneDev& sc = scenes[scene];
        const uint32_t tl = t - sc.tile_base;
        const int W = int(sc.W);
        const int tx = int(tl % sc.W), ty = int((tl / sc.W) % sc.W), tz = int(tl / (sc.W * sc.W));
        const int xa = max(tx - 1, 0), xb = min(tx + 1, W - 1);
        const int nz = tz + int(r) / 3 - 1, ny = ty + int(r) % 3 - 1;
        int c0 = 0, c1 = -1;
        if (nz >= 0 && nz < W && ny >= 0 && ny < W) {
            const uint32_t row = sc.tile_base + uint32_t(W * (ny + W * nz));
            const uint32_t s0 = lb[row + xa], s1 = lb[row + xb + 1];
            if (s0 < s1) {
                c0 = int(chunk_of_slot[s0]);
                c1 = min(int(chunk_of_slot[s1 - 1]), int(part_capacity) - 1);
            }
        }
        int* rec = srec + (size_t)i * SREC_WORDS;
        if (r == 0) {
            rec[0] = int(t);
            rec[1] = int(scene);
        }
        rec[2 + 2 * r] = c0;
        rec[3 + 2 * r] = c1;
    }
}

hipError_t launch_static_records(const SceneDev* scenes, const uint32_t* tile_scene, const uint32_t* slist,
                                 const uint32_t* scount, uint32_t max_static, const uint32_t* tile_lb,
                                 const uint32_t* chunk_of_slot, uint32_t part_capacity, int* srec, hipStream_t s) {
    if (max_static == 0) return hipSuccess;
    const unsigned blocks = std::max(1u, std::min<unsigned>((max_static * 9u + 255u) / 256u, 8192u));
    hipLaunchKernelGGL(k_static_records, dim3(blocks), dim3(256), 0, s, scenes, tile_scene, slist, scount,
                       max_static, tile_lb, chunk_of_slot, part_capacity, srec);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile keys (shaders/mpm_particle_order.wgsl:13-25): cell = floor((x - o) *
// inv_dx) - window, clamped to [0, G-1], /4, linearised x-fastest; the global
// key adds the scene's tile base.  Written at the particle's original id so
// that a stable sort over ids 0..n-1 reproduces the reference order, or (the
// layout sort) at the particle's slot, so ties keep the current slot order
// and the reorder gather reads nearly sequential memory.
__global__ __launch_bounds__(256) void k_tile_keys(StepArgs a, uint32_t* __restrict__ keys_out, int by_slot) {
    for (uint32_t ci = blockIdx.x; ci < a.n_chunks; ci += gridDim.x) {
        const Chunk ch = a.chunks[ci];
        if (threadIdx.x >= ch.count) continue;
        const SceneDev& sc = a.scenes[ch.scene];
        const uint32_t slot = ch.start + threadIdx.x;
        uint32_t t[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float c = floorf((a.ps[(size_t)(CX + k) * a.cap + slot] - sc.origin[k]) * sc.inv_dx);
            int ci2 = int(c) - sc.window[k];
            ci2 = min(max(ci2, 0), int(sc.G) - 1);
            t[k] = uint32_t(ci2) >> 2;
        }
        const uint32_t at = by_slot ? slot : __float_as_uint(a.ps[(size_t)CID * a.cap + slot]);
        keys_out[at] = sc.tile_base + t[0] + sc.W * (t[1] + sc.W * t[2]);
    }
}

hipError_t launch_tile_keys(const StepArgs& a, uint32_t* keys_out, int by_slot, hipStream_t s) {
    if (a.n_chunks == 0) return hipSuccess;
    unsigned blocks = a.n_chunks < 65535u ? a.n_chunks : 65535u;
    hipLaunchKernelGGL(k_tile_keys, dim3(blocks), dim3(256), 0, s, a, keys_out, by_slot);
    return hipGetLastError();
}

// src_slot[j] = slot_of_id[sorted_ids[j]]; then every plane is gathered with
// one (plane, slot) pair per thread so reads and writes stay coalesced.
__global__ void k_reorder_index(const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ slot_of_id,
                                uint32_t* __restrict__ src_slot, uint32_t n) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        src_slot[j] = slot_of_id[sorted_ids[j]];
}

hipError_t launch_reorder_index(const uint32_t* sorted_ids, const uint32_t* slot_of_id,
                                uint32_t* src_slot, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = std::min<unsigned>((n + 255) / 256, 8192u);
    hipLaunchKernelGGL(k_reorder_index, dim3(blocks), dim3(256), 0, s, sorted_ids, slot_of_id, src_slot, n);
    return hipGetLastError();
}

// One slot per thread through all planes: the index is read once and every
// thread keeps a plane group's loads in flight together.
__global__ __launch_bounds__(256) void k_reorder(const float* __restrict__ src, float* __restrict__ dst,
                                                 uint32_t cap, const uint32_t* __restrict__ src_slot,
                                                 uint32_t n) {
    constexpr int GROUP = 10;
    static_assert(NCOMP % GROUP == 0, "plane groups");
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const size_t from = src_slot[j];
        for (int c0 = 0; c0 < NCOMP; c0 += GROUP) {
            float v[GROUP];
#pragma unroll
            for (int c = 0; c < GROUP; ++c) v[c] = src[(size_t)(c0 + c) * cap + from];
#pragma unroll
            for (int c = 0; c < GROUP; ++c) dst[(size_t)(c0 + c) * cap + j] = v[c];
        }
    }
}

hipError_t launch_reorder(const float* src, float* dst, uint32_t cap, const uint32_t* src_slot,
                          uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = std::min<unsigned>((n + 255) / 256, 16384u);
    hipLaunchKernelGGL(k_reorder, dim3(blocks), dim3(256), 0, s, src, dst, cap, src_slot, n);
    return hipGetLastError();
}

// Spreads the particles of each chunk of the sorted layout over cells: the
// chunk's particles are ranked among those with the same stencil base cell
// (ties by slot) and laid out rank-major, base cells ascending within a rank.
// The sort alone leaves a chunk's particles in arbitrary order inside their
// tiles, so the 64 lanes of a wave fall on ~125 base cells at random and the
// scatter's LDS integer adds to one word from several lanes serialise; in
// rank-major order a wave's lanes mostly hold distinct, neighbouring bases.
// Only the order inside a chunk changes: the chunk list and its partial
// boxes stay valid, and the fixed-point scatter sums do not depend on order.
// src_in/src_out: previous slot of each sorted slot, before and after; the
// sorted tile keys are permuted with them (in place), so each slot keeps the
// key of its own particle for the scatter's escape test.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int mask) {
    const uint32_t lo = __shfl_xor(uint32_t(x), mask, 64);
    const uint32_t hi = __shfl_xor(uint32_t(x >> 32), mask, 64);
    return (uint64_t(hi) << 32) | lo;
}

// Ascending bitonic sort of one value per thread of an N-thread block:
// returns the value at sorted position t.  Partners closer than a wave are
// exchanged by shuffles; the steps across waves go through v (LDS).
// Every thread calls it; v may be read by other threads until the first
// barrier.
template <uint32_t N>
__device__ __forceinline__ uint64_t block_sort(uint64_t x, uint64_t* v, uint32_t t) {
    for (uint32_t k = 2; k <= N; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            uint64_t other;
            if (j >= 64u) {
                __syncthreads();
                v[t] = x;
                __syncthreads();
                other = v[t ^ j];
            } else {
                other = shfl_xor_u64(x, int(j));
            }
            // the lower index of a pair keeps the minimum in an ascending run
            const bool keep_min = (t < (t ^ j)) == ((t & k) == 0u);
            x = keep_min ? (x < other ? x : other) : (x < other ? other : x);
        }
    __syncthreads();  // the last cross-wave reads of v are done
    return x;
}

// One thread per particle of a chunk of up to N particles (N = 256 x the
// particles per thread of the pass): a thread of the pass takes positions t,
// t + 256, ..., so each of its passes over the block is rank-major too.
template <uint32_t N>
__global__ __launch_bounds__(N) void k_chunk_spread(const SceneDev* __restrict__ scenes,
                                                    const Chunk* __restrict__ chunks,
                                                    const uint32_t* __restrict__ n_chunks,
                                                    const float* __restrict__ ps, uint32_t cap,
                                                    const uint32_t* __restrict__ src_in,
                                                    uint32_t* __restrict__ src_out,
                                                    uint32_t* __restrict__ keys) {
    constexpr uint32_t TB = N == 256 ? 8 : 9;  // bits of a position
    constexpr uint64_t TMASK = (1ull << TB) - 1ull;
    __shared__ uint64_t order[N];
    __shared__ uint32_t rank[N];
    __shared__ uint32_t tiles[N];
    const uint32_t nch = *n_chunks;
    const uint32_t t = threadIdx.x;
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        const Chunk ch = chunks[ci];
        const uint32_t cnt = min(ch.count, N);
        const SceneDev& sc = scenes[ch.scene];
        uint32_t src = 0u, tile = 0u;
        uint64_t base = ~0ull;  // 30-bit base cell key; unused threads sort last
        if (t < cnt) {
            src = src_in[ch.start + t];
            tile = keys[ch.start + t];
            uint32_t k = 0u;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float xp = (ps[(size_t)(CX + d) * cap + src] - sc.origin[d]) * sc.inv_dx;
                const int b = int(floorf(xp - 0.5f)) - sc.window[d] + 1;
                k |= (uint32_t(b) & 1023u) << (10 * d);
            }
            base = uint64_t(k);
        }
        // 1. (base, slot) ascending: a particle's rank is its distance from
        //    the first entry of its base's run
        const uint64_t e = block_sort<N>(base == ~0ull ? ~0ull : (base << TB) | t, order, t);
        order[t] = e;
        __syncthreads();
        if (e != ~0ull) {
            const uint64_t first = e & ~TMASK;
            uint32_t lo = 0u, hi = t;  // lower bound of the run in order[0, t]
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (order[mid] < first) lo = mid + 1u;
                else hi = mid;
            }
            rank[uint32_t(e & TMASK)] = t - lo;
        }
        __syncthreads();
        // 2. (rank, base, slot) ascending: the new position of every slot
        const uint64_t f = block_sort<N>(
            base == ~0ull ? ~0ull : (uint64_t(rank[t]) << (30 + TB)) | (base << TB) | t, order, t);
        const uint32_t from = uint32_t(f & TMASK);
        // hand the slot's previous slot and tile key to its new position
        // (rank is free: every thread read its own entry before the sort's barriers)
        rank[t] = src;
        tiles[t] = tile;
        __syncthreads();
        if (t < cnt) {
            src_out[ch.start + t] = rank[from];
            keys[ch.start + t] = tiles[from];
        }
        __syncthreads();
    }
}

hipError_t launch_chunk_spread(const SceneDev* scenes, const Chunk* chunks, const uint32_t* n_chunks,
                               uint32_t max_chunks, const float* ps, uint32_t cap, const uint32_t* src_in,
                               uint32_t* src_out, uint32_t* keys, uint32_t chunk_cap, hipStream_t s) {
    if (max_chunks == 0) return hipSuccess;
    const unsigned blocks = std::min<unsigned>(max_chunks, 8192u);
    if (chunk_cap <= CHUNK_PARTICLES)
        hipLaunchKernelGGL(k_chunk_spread<CHUNK_PARTICLES>, dim3(blocks), dim3(CHUNK_PARTICLES), 0, s, scenes,
                           chunks, n_chunks, ps, cap, src_in, src_out, keys);
    else if (chunk_cap <= CHUNK_PARTICLES * MAX_PPT)
        hipLaunchKernelGGL(k_chunk_spread<CHUNK_PARTICLES * MAX_PPT>, dim3(blocks), dim3(CHUNK_PARTICLES * MAX_PPT),
                           0, s, scenes, chunks, n_chunks, ps, cap, src_in, src_out, keys);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

__global__ void k_index_slots(const float* __restrict__ ps, uint32_t cap, uint32_t* __restrict__ slot_of_id,
                              uint32_t n) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        slot_of_id[__float_as_uint(ps[(size_t)CID * cap + j])] = j;
}

hipError_t launch_index_slots(const float* ps, uint32_t cap, uint32_t* slot_of_id, uint32_t n,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_index_slots, dim3(blocks), dim3(256), 0, s, ps, cap, slot_of_id, n);
    return hipGetLastError();
}

__global__ void k_unpermute(const float* __restrict__ ps, uint32_t cap, int comp0, int ncomp,
                            const uint32_t* __restrict__ slot_of_id, float* __restrict__ out, uint32_t n) {
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
        const uint32_t s = slot_of_id[id];
        for (int c = 0; c < ncomp; ++c) out[(size_t)id * ncomp + c] = ps[(size_t)(comp0 + c) * cap + s];
    }
}

hipError_t launch_unpermute(const float* ps, uint32_t cap, int comp0, int ncomp,
                            const uint32_t* slot_of_id, float* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_unpermute, dim3(blocks), dim3(256), 0, s, ps, cap, comp0, ncomp, slot_of_id, out, n);
    return hipGetLastError();
}

__global__ void k_permute_in(float* __restrict__ ps, uint32_t cap, int comp0, int ncomp,
                             const uint32_t* __restrict__ slot_of_id, const float* __restrict__ in, uint32_t n) {
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
        const uint32_t s = slot_of_id[id];
        for (int c = 0; c < ncomp; ++c) ps[(size_t)(comp0 + c) * cap + s] = in[(size_t)id * ncomp + c];
    }
}

hipError_t launch_permute_in(float* ps, uint32_t cap, int comp0, int ncomp,
                             const uint32_t* slot_of_id, const float* in, uint32_t n,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_permute_in, dim3(blocks), dim3(256), 0, s, ps, cap, comp0, ncomp, slot_of_id, in, n);
    return hipGetLastError();
}

// ---- soil checkpoint streams (mpm_checkpoint.rs:78-83, :352-369) ----------
// The archive holds n particles as nine streams, stream-major, 40 words per
// particle: position+w (4), velocity+w (4), C (12), F (12), volume, mass,
// log-Jp, material id, tire compaction (4).  C and F are three 16-byte
// columns, word 4j+i = M_ij, word 4j+3 a pad.  The planes hold M row-major
// (plane 3i+j).  `aux` keeps, by particle id, the words the planes do not:
// position w [n], velocity w [n], C pads [3n], F pads [3n], tire [4n]; a null
// aux reads as zeros.  Every stream offset is a multiple of 4 words from a
// 256-byte-aligned base, so the 4-word groups move as 16-byte accesses.
__global__ void k_ckpt_pack(const float* __restrict__ ps, uint32_t cap, const uint32_t* __restrict__ slot_of_id,
                            const uint32_t* __restrict__ aux, int zero_w, uint32_t n, uint32_t* __restrict__ out) {
    const size_t N = n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const size_t s = slot_of_id[p];
        auto P = [&](int c) { return __float_as_uint(ps[(size_t)c * cap + s]); };
        auto A = [&](size_t i) { return aux ? aux[i] : 0u; };
        // a step stores w = 0 (mpm.wgsl:1349-1350)
        reinterpret_cast<uint4*>(out)[p] = make_uint4(P(CX), P(CX + 1), P(CX + 2), zero_w ? 0u : A(p));
        reinterpret_cast<uint4*>(out + 4 * N)[p] = make_uint4(P(CV), P(CV + 1), P(CV + 2), zero_w ? 0u : A(N + p));
        for (int j = 0; j < 3; ++j) {
            reinterpret_cast<uint4*>(out + 8 * N)[3 * p + j] =
                make_uint4(P(CC + j), P(CC + 3 + j), P(CC + 6 + j), A(2 * N + 3 * p + j));
            reinterpret_cast<uint4*>(out + 20 * N)[3 * p + j] =
                make_uint4(P(CF + j), P(CF + 3 + j), P(CF + 6 + j), A(5 * N + 3 * p + j));
        }
        out[32 * N + p] = P(CVOL);
        out[33 * N + p] = P(CMASS);
        out[34 * N + p] = P(CJP);
        out[35 * N + p] = P(CMAT);
        const size_t t = 8 * N + 4 * (size_t)p;
        reinterpret_cast<uint4*>(out + 36 * N)[p] = make_uint4(A(t), A(t + 1), A(t + 2), A(t + 3));
    }
}

// Inverse of k_ckpt_pack into slot = id order; E is the scene's, the id plane
// the identity.  `aux` (12n words) receives the words the planes do not hold.
__global__ void k_ckpt_unpack(const uint32_t* __restrict__ in, uint32_t n, float* __restrict__ ps, uint32_t cap,
                              float youngs, uint32_t* __restrict__ aux) {
    const size_t N = n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        auto put = [&](int c, uint32_t w) { ps[(size_t)c * cap + p] = __uint_as_float(w); };
        const uint4 x = reinterpret_cast<const uint4*>(in)[p];
        const uint4 v = reinterpret_cast<const uint4*>(in + 4 * N)[p];
        put(CX, x.x), put(CX + 1, x.y), put(CX + 2, x.z);
        put(CV, v.x), put(CV + 1, v.y), put(CV + 2, v.z);
        aux[p] = x.w;
        aux[N + p] = v.w;
        for (int j = 0; j < 3; ++j) {
            const uint4 c = reinterpret_cast<const uint4*>(in + 8 * N)[3 * p + j];
            const uint4 f = reinterpret_cast<const uint4*>(in + 20 * N)[3 * p + j];
            put(CC + j, c.x), put(CC + 3 + j, c.y), put(CC + 6 + j, c.z);
            put(CF + j, f.x), put(CF + 3 + j, f.y), put(CF + 6 + j, f.z);
            aux[2 * N + 3 * p + j] = c.w;
            aux[5 * N + 3 * p + j] = f.w;
        }
        put(CVOL, in[32 * N + p]);
        put(CMASS, in[33 * N + p]);
        put(CJP, in[34 * N + p]);
        put(CMAT, in[35 * N + p]);
        const uint4 t = reinterpret_cast<const uint4*>(in + 36 * N)[p];
        const size_t a = 8 * N + 4 * (size_t)p;
        aux[a] = t.x, aux[a + 1] = t.y, aux[a + 2] = t.z, aux[a + 3] = t.w;
        ps[(size_t)CE * cap + p] = youngs;
        put(CID, p);
    }
}

hipError_t launch_ckpt_pack(const float* ps, uint32_t cap, const uint32_t* slot_of_id, const uint32_t* aux,
                            int zero_w, uint32_t n, uint32_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = std::min<unsigned>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_ckpt_pack, dim3(blocks), dim3(256), 0, s, ps, cap, slot_of_id, aux, zero_w, n, out);
    return hipGetLastError();
}

hipError_t launch_ckpt_unpack(const uint32_t* in, uint32_t n, float* ps, uint32_t cap, float youngs,
                              uint32_t* aux, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = std::min<unsigned>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_ckpt_unpack, dim3(blocks), dim3(256), 0, s, in, n, ps, cap, youngs, aux);
    return hipGetLastError();
}

__global__ void k_grid_to_dense(const float4* __restrict__ gmp, const SceneDev* __restrict__ scenes,
                                uint32_t scene, uint32_t G, float* __restrict__ mass, float* __restrict__ mom) {
    const SceneDev sc = scenes[scene];
    const size_t total = (size_t)G * G * G;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < total; c += (size_t)gridDim.x * blockDim.x) {
        int ix = int(c % G), iy = int((c / G) % G), iz = int(c / ((size_t)G * G));
        float4 q = gmp[cell_index(sc, ix, iy, iz)];
        mass[c] = q.x;
        mom[3 * c] = q.y;
        mom[3 * c + 1] = q.z;
        mom[3 * c + 2] = q.w;
    }
}

hipError_t launch_grid_to_dense(const float4* gmp, const SceneDev* scenes, uint32_t scene,
                                uint32_t G, float* mass, float* mom, hipStream_t s) {
    size_t total = (size_t)G * G * G;
    unsigned blocks = (unsigned)((total + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_grid_to_dense, dim3(blocks), dim3(256), 0, s, gmp, scenes, scene, G, mass, mom);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Deterministic gather-order P2G (mpm_deterministic_p2g.wgsl:12-150).
// Per particle: affine matrix A (planes 0-8) and m v (planes 9-11).
template <int MODEL>
__global__ __launch_bounds__(256) void k_gather_prepare(StepArgs a, float* __restrict__ aff) {
    const uint32_t cap = a.cap;
    const float* ps = a.ps;
    for (uint32_t ci = blockIdx.x; ci < a.n_chunks; ci += gridDim.x) {
        const Chunk ch = a.chunks[ci];
        if (threadIdx.x >= ch.count) continue;
        const SceneDev sc = a.scenes[ch.scene];
        const uint32_t slot = ch.start + threadIdx.x;
        float F[9], C[9];
        for (int k = 0; k < 9; ++k) {
            F[k] = ps[(size_t)(CF + k) * cap + slot];
            C[k] = ps[(size_t)(CC + k) * cap + slot];
        }
        const float m = ps[(size_t)CMASS * cap + slot], vol = ps[(size_t)CVOL * cap + slot];
        const float jp = ps[(size_t)CJP * cap + slot];
        const float* mrow = mat_row(a, sc, __float_as_uint(ps[(size_t)CMAT * cap + slot]));
        float mu = sc.mu, lam = sc.lam;
        if (a.has_E) {
            float E = ps[(size_t)CE * cap + slot];
            mu = E / (2.0f * (1.0f + sc.poisson));
            lam = E * sc.poisson / ((1.0f + sc.poisson) * (1.0f - 2.0f * sc.poisson));
        }
        Svd32 d;
        svd3<MODEL>(F, d);
        float A[9];
        affine_from_svd<MODEL>(sc, mrow, d, mu, lam, jp, a.dt, vol, m, C, A);
        for (int k = 0; k < 9; ++k) aff[(size_t)k * cap + slot] = A[k];
        for (int k = 0; k < 3; ++k) aff[(size_t)(9 + k) * cap + slot] = m * ps[(size_t)(CV + k) * cap + slot];
    }
}

hipError_t launch_gather_prepare(const StepArgs& a, float* affine, int blocks, hipStream_t s) {
    if (a.n_chunks == 0) return hipSuccess;
    if (blocks <= 0 || uint32_t(blocks) > a.n_chunks) blocks = int(a.n_chunks);
    // the model is uniform per context; scenes[0] tells which
    SceneDev h;
    (void)h;
    hipLaunchKernelGGL((k_gather_prepare<MODEL_SHADER>), dim3(blocks), dim3(256), 0, s, a, affine);
    return hipGetLastError();
}

// One wave per active tile, one lane per cell.  For each cell, particles of
// the 27 neighbour tiles are summed in tile order dx + 3 dy + 9 dz, then in
// ascending sorted slot (mpm_deterministic_p2g.wgsl:100-144): a fixed order,
// so the result is bitwise repeatable.
__global__ __launch_bounds__(64) void k_gather_p2g(StepArgs a, const uint32_t* __restrict__ tile_scene,
                                                   const uint32_t* __restrict__ lb,
                                                   const float* __restrict__ aff,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count) {
    const uint32_t n = *count;
    const int lane = threadIdx.x;
    const uint32_t cap = a.cap;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t t = list[i];
        const SceneDev sc = a.scenes[tile_scene[t]];
        const uint32_t tl = t - sc.tile_base;
        const int W = int(sc.W);
        const int tx = int(tl % sc.W), ty = int((tl / sc.W) % sc.W), tz = int(tl / (sc.W * sc.W));
        const int cx = tx * 4 + (lane & 3), cy = ty * 4 + ((lane >> 2) & 3), cz = tz * 4 + (lane >> 4);
        float ms = 0.f, p0 = 0.f, p1 = 0.f, p2 = 0.f;
        for (int ti = 0; ti < 27; ++ti) {
            const int nx = tx + ti % 3 - 1, ny = ty + (ti / 3) % 3 - 1, nz = tz + ti / 9 - 1;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= W || ny >= W || nz >= W) continue;
            const uint32_t key = sc.tile_base + uint32_t(nx + W * (ny + W * nz));
            const uint32_t jb = lb[key], je = lb[key + 1];
            for (uint32_t j = jb; j < je; ++j) {
                float xp[3], fx[3];
                int off[3];
                bool in = true;
                for (int k = 0; k < 3; ++k) {
                    xp[k] = (a.ps[(size_t)(CX + k) * cap + j] - sc.origin[k]) * sc.inv_dx;
                    float bf = floorf(xp[k] - 0.5f);
                    fx[k] = xp[k] - bf;
                    const int c = k == 0 ? cx : (k == 1 ? cy : cz);
                    off[k] = c - (int(bf) - sc.window[k]);
                    in = in && off[k] >= 0 && off[k] <= 2;
                }
                if (!in) continue;
                float w0[3], w1[3], w2[3];
                quad_weights(fx[0], w0);
                quad_weights(fx[1], w1);
                quad_weights(fx[2], w2);
                const float wt = w0[off[0]] * w1[off[1]] * w2[off[2]];
                const float d0 = (float(off[0]) - fx[0]) * sc.dx, d1 = (float(off[1]) - fx[1]) * sc.dx,
                            d2 = (float(off[2]) - fx[2]) * sc.dx;
                float A[9];
                for (int k = 0; k < 9; ++k) A[k] = aff[(size_t)k * cap + j];
                const float m = a.ps[(size_t)CMASS * cap + j];
                ms += wt * m;
                p0 += wt * (aff[(size_t)9 * cap + j] + A[0] * d0 + A[1] * d1 + A[2] * d2);
                p1 += wt * (aff[(size_t)10 * cap + j] + A[3] * d0 + A[4] * d1 + A[5] * d2);
                p2 += wt * (aff[(size_t)11 * cap + j] + A[6] * d0 + A[7] * d1 + A[8] * d2);
            }
        }
        const bool valid = cx < int(sc.G) && cy < int(sc.G) && cz < int(sc.G);
        a.gmp[(t << 6) | uint32_t(lane)] = valid ? make_float4(ms, p0, p1, p2) : make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane == 0) a.tile_flag[t] |= FLAG_GMP;
    }
}

hipError_t launch_gather_p2g(const StepArgs& a, const uint32_t* tile_scene, const uint32_t* tile_lb,
                             const float* affine, const uint32_t* slist, const uint32_t* scount,
                             int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_p2g, dim3(blocks), dim3(64), 0, s, a, tile_scene, tile_lb, affine, slist, scount);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sparse tile pool (MpmSimulation::new_sparse, mpm.rs:848-853; tile marking
// mpm_active_tiles.wgsl:79-130).  Every substep the particles that scatter
// mark the tiles within one cell of their stencils (block_scatter, from its
// LDS box; deferred particles directly), legacy rigid spheres mark the tiles
// their shell can reach, and one pass over the tile table gives the marked
// tiles the pool's slots, in arrival order, and clears the marks.  A demand
// above the capacity latches FAULT_OVERFLOW: every later kernel of the
// substep sees the latch and leaves the particles as they were.
__global__ __launch_bounds__(256) void k_pool_compact(uint32_t* __restrict__ tile_mark, uint32_t total,
                                                      uint32_t* __restrict__ tile_slot,
                                                      uint32_t* __restrict__ slot_tile, uint32_t capacity,
                                                      uint32_t* __restrict__ status) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        uint32_t s1 = 0u;
        if (tile_mark[t]) {
            tile_mark[t] = 0u;
            const uint32_t slot = atomicAdd(&status[ST_POOL], 1u);
            if (slot < capacity) {
                s1 = slot + 1u;
                slot_tile[slot] = t;
            } else {
                latch(status, FAULT_OVERFLOW);
            }
        }
        tile_slot[t] = s1;
    }
}

hipError_t launch_pool_compact(uint32_t* tile_mark, uint32_t total_tiles, uint32_t* tile_slot, uint32_t* slot_tile,
                               uint32_t capacity, uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, status + ST_POOL, 0u);
    const unsigned blocks = std::max(1u, std::min<unsigned>((total_tiles + 255) / 256, 16384u));
    hipLaunchKernelGGL(k_pool_compact, dim3(blocks), dim3(256), 0, s, tile_mark, total_tiles, tile_slot, slot_tile,
                       capacity, status);
    return hipGetLastError();
}

// One block per legacy sphere: the tiles of its cells within radius + 2 cells
// (mark_rigids, mpm_active_tiles.wgsl:108-124).  A sphere whose tile range
// alone exceeds the pool latches the overflow without marking.
__global__ __launch_bounds__(256) void k_mark_spheres(const SceneDev* __restrict__ scenes, RigidDev rg,
                                                      uint32_t* __restrict__ tile_mark, uint32_t capacity,
                                                      uint32_t* __restrict__ status) {
    const float4* b = rg.body + RIGID_ROWS * blockIdx.x;
    if (b[4].w != 0.0f) return;  // prescribed solids mark nothing
    const SceneDev& sc = scenes[0];
    int lo[3], n[3];
    const float c[3] = {b[0].x, b[0].y, b[0].z};
    const float extent = fmaxf(b[0].w, 0.0f) * sc.inv_dx + 2.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float ck = (c[k] - sc.origin[k]) * sc.inv_dx;
        const int l = max(int(floorf(ck - extent)) - sc.window[k], 0);
        const int h = min(int(ceilf(ck + extent)) - sc.window[k], int(sc.G) - 1);
        if (l > h) return;
        lo[k] = l >> 2;
        n[k] = (h >> 2) - lo[k] + 1;
    }
    const uint64_t count = (uint64_t)n[0] * n[1] * n[2];
    if (count > capacity) {
        if (threadIdx.x == 0) latch(status, FAULT_OVERFLOW);
        return;
    }
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const uint32_t x = i % n[0], y = (i / n[0]) % n[1], z = i / (n[0] * n[1]);
        tile_mark[sc.tile_base + uint32_t(lo[0]) + x + sc.W * (uint32_t(lo[1]) + y + sc.W * (uint32_t(lo[2]) + z))] = 1u;
    }
}

hipError_t launch_mark_spheres(const SceneDev* scenes, const RigidDev& rg, uint32_t* tile_mark, uint32_t capacity,
                               uint32_t* status, hipStream_t s) {
    if (rg.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mark_spheres, dim3(rg.count), dim3(256), 0, s, scenes, rg, tile_mark, capacity, status);
    return hipGetLastError();
}

// The global-atomic P2G of the particles a sparse pool deferred (escaped
// stencils, chunks that could not reduce in LDS, the global-atomic mode):
// the P2G pass's particle math (stress from the FP64 SVD of F,
// mpm.wgsl:532-599) once the compaction has given their tiles slots.  Sparse
// pools hold one scene of the shader model.
__global__ __launch_bounds__(256) void k_deferred_p2g(StepArgs a) {
    if (a.status[ST_FAULT] != 0u) return;
    const uint32_t n = a.status[ST_DEFER], cap = a.cap;
    const float* ps = a.ps;
    const SceneDev sc = a.scenes[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t slot = a.defer_list[i], boff = 4u * slot;
        float x[3], v[3], C[9], F[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = plane_ld(ps, cap, CX + k, boff);
            v[k] = plane_ld(ps, cap, CV + k, boff);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            C[k] = plane_ld(ps, cap, CC + k, boff);
            F[k] = plane_ld(ps, cap, CF + k, boff);
        }
        const float jp = plane_ld(ps, cap, CJP, boff), m = plane_ld(ps, cap, CMASS, boff),
                    vol = plane_ld(ps, cap, CVOL, boff);
        const float* mrow = mat_row(a, sc, __float_as_uint(plane_ld(ps, cap, CMAT, boff)));
        float mu = sc.mu, lam = sc.lam;
        if (a.has_E) {
            const float E = plane_ld(ps, cap, CE, boff);
            mu = E / (2.0f * (1.0f + sc.poisson));
            lam = E * sc.poisson / ((1.0f + sc.poisson) * (1.0f - 2.0f * sc.poisson));
        }
        Scatter sp;
        Svd32 d;
        svd3<MODEL_SHADER>(F, d);
        affine_from_svd<MODEL_SHADER>(sc, mrow, d, mu, lam, jp, a.dt, vol, m, C, sp.A);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float xp = (x[k] - sc.origin[k]) * sc.inv_dx;
            const float bf = floorf(xp - 0.5f);
            sp.b[k] = int(bf) - sc.window[k];
            sp.fx[k] = xp - bf;
            sp.mv[k] = m * v[k];
        }
        sp.m = m;
        sp.on = true;
        sp.esc = true;
        sp.slot = slot;
        scatter_global(a, sc, sp);
    }
}

hipError_t launch_deferred_p2g(const StepArgs& a, int model, hipStream_t s) {
    if (model != MODEL_SHADER) return hipErrorInvalidValue;
    const unsigned blocks = std::max(1u, std::min<unsigned>((a.cap + 255) / 256, 4096u));
    hipLaunchKernelGGL(k_deferred_p2g, dim3(blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, a.status + ST_DEFER, 0u);
    return hipGetLastError();
}

__global__ void k_storage_layout(const float4* __restrict__ gmp, uint32_t cells, float* __restrict__ mass,
                                 float* __restrict__ mom) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += gridDim.x * blockDim.x) {
        const float4 q = gmp[i];
        mass[i] = q.x;
        mom[3 * (size_t)i] = q.y;
        mom[3 * (size_t)i + 1] = q.z;
        mom[3 * (size_t)i + 2] = q.w;
    }
}

hipError_t launch_storage_to_host_layout(const float4* gmp, uint32_t cells, float* mass, float* mom, hipStream_t s) {
    const unsigned blocks = std::max(1u, std::min<unsigned>((cells + 255) / 256, 8192u));
    hipLaunchKernelGGL(k_storage_layout, dim3(blocks), dim3(256), 0, s, gmp, cells, mass, mom);
    return hipGetLastError();
}

}  // namespace mpm
