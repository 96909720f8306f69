// mpm_diff.hip -- reverse pass of one CPUDIFF substep on gfx950.
//
// Follows mpm_adjoint.rs backward_step (:142-496) for the forward of
// mpm_cpu_diff.rs step_once_record (:421-588): corotated-SVK stress, floor
// plane, explicit APIC.  The forward substep is
//   P2G (mass, momentum + stress) -> grid (v = p/m, gravity, floor)
//   -> G2P (v, C) -> x += dt v, F = (I + dt C) F,
// and its reverse runs as three launches over the substep's pre-state, which
// the host recomputes from a checkpoint (k_adj_fused merges the reverse P2G
// of one substep with the reverse G2P of the previous one):
//   k_adj_g2p   reverse of G2P and of the x/F advance.  Gathers grid v like
//               G2P and scatters dJ/d(grid v) = sum w (gv' + 4/dx^2 gC' d)
//               with the forward's block scatter (the same form as P2G's
//               w (m v + A d), with m = 0), so it reuses the LDS fixed-point
//               partials and the sorted chunk list.
//   k_adj_grid  reverse of the grid update per active cell:
//               dJ/dp = M(dJ/dv) / m, dJ/dm = -p . M(dJ/dv) / m^2, where M is
//               the floor's mask replayed from the pre-update momentum.
//   k_adj_p2g   reverse of P2G, of the stress and of the SVD: gathers
//               (dJ/dm, dJ/dp), and runs the PK1 and SVD adjoints in FP64
//               (mpm_cpu_diff.rs:330-379, dense3.rs:321-372).
// Particle adjoints live in NADJ planes in the state's slot order; the
// per-particle dJ/dE and the per-chunk (dJ/dmu, dJ/dlambda) sums accumulate
// over substeps (StateAdjoint's g_mu/g_lambda, mpm_adjoint.rs:76-113).
#include "mpm_scatter.hpp"

namespace mpm {

namespace {

// 1/x in f64 from the f32 reciprocal and two Newton steps (x > 0, normal f32).
__device__ __forceinline__ double rcp_f64(double x) {
    double r = (double)__frcp_rn((float)x);
    r = r * (2.0 - x * r);
    r = r * (2.0 - x * r);
    return r;
}

// safe_inv_sq_diff (dense3.rs:496-499) with ADJ_SVD_EPS = 1e-12 (:295).
__device__ __forceinline__ double safe_inv_sq_diff(double a, double b) {
    const double d = b * b - a * a;
    return d * rcp_f64(d * d + 1.0e-24);
}

struct Dm3 {
    double a[9];
    __device__ double& operator()(int i, int j) { return a[3 * i + j]; }
    __device__ double operator()(int i, int j) const { return a[3 * i + j]; }
};

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// Reverse of G2P and of the x/F advance for the particle in `slot` of the
// pre-state a.ps with grid velocity a.gv (mpm_adjoint.rs:196-305).  On entry
// gx, gv, gC, gF are dJ/d(post-state); on return gx holds the G2P part of
// dJ/dx (plus the pass-through), gF = (I + dt C')^T dJ/dF', and sp the
// scatter of dJ/d(grid v).
__device__ __forceinline__ void adj_g2p_particle(const StepArgs& a, const SceneDev& sc, const Chunk& ch, uint32_t slot,
                                                 float* gx, const float* gv, const float* gC, float* gF, Scatter& sp) {
    const uint32_t cap = a.cap;
    const float* __restrict__ ps = a.ps;
    const float dt = a.dt;
    float x[3], F[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = ps[(size_t)(CX + k) * cap + slot];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = ps[(size_t)(CF + k) * cap + slot];
    // x' = x + dt v', F' = (I + dt C') F  =>  dJ/dv' += dt dJ/dx',
    // dJ/dC' += dt dJ/dF' F^T (mpm_adjoint.rs:196-247)
    float gvp[3], B[9];
    const float s4 = 4.0f * sc.inv_dx * sc.inv_dx;
#pragma unroll
    for (int k = 0; k < 3; ++k) gvp[k] = gx[k] * dt + gv[k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[3 * i + j] = s4 * (gC[3 * i + j] + dt * (gF[3 * i] * F[3 * j] + gF[3 * i + 1] * F[3 * j + 1] +
                                                       gF[3 * i + 2] * F[3 * j + 2]));
    float xp[3], fx[3], w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xp[k] = (x[k] - sc.origin[k]) * sc.inv_dx;
        const float bf = floorf(xp[k] - 0.5f);
        b[k] = int(bf);
        fx[k] = xp[k] - bf;
        quad_weights(fx[k], w[k]);
        dw[k][0] = -(1.5f - fx[k]);
        dw[k][1] = -2.0f * (fx[k] - 1.0f);
        dw[k][2] = fx[k] - 0.5f;
    }
    // G2P reverse (mpm_adjoint.rs:249-305): with B = 4/dx^2 dJ/dC',
    //   dJ/dx += sum grad(w) (v_i . (dJ/dv' + B d)) - B^T sum w v_i,
    // summed row by row (fixed j, k): with d = ((i, j, k) - fx) dx the node
    // term is u_jk + i b0 (b0 = dx B[:, 0]); with Sw = sum_i w0 h and
    // Sdw = sum_i dw0 h the gradient sum is
    //   idx (sum w1 w2 Sdw, sum dw1 w2 Sw, sum w1 dw2 Sw),
    // and C' = 4/dx^2 sum w v_i d^T = 4/dx (Mv - vn fx^T), Mv = sum w v_i (i, j, k)^T.
    const float dx = sc.dx, idx = sc.inv_dx;
    float b0[3], u0[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        b0[r] = dx * B[3 * r];
        u0[r] = gvp[r] - (b0[r] * fx[0] + dx * B[3 * r + 1] * fx[1] + dx * B[3 * r + 2] * fx[2]);
    }
    const float wi[3] = {0.f, w[0][1], 2.0f * w[0][2]};
    float vn[3] = {0.f, 0.f, 0.f}, Mv[9], acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 9; ++k) Mv[k] = 0.f;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            const float wjk = w[1][j] * w[2][k];
            float u[3], rs[3] = {0.f, 0.f, 0.f}, ri[3] = {0.f, 0.f, 0.f}, Sw = 0.f, Sdw = 0.f;
#pragma unroll
            for (int r = 0; r < 3; ++r) u[r] = u0[r] + float(j) * (dx * B[3 * r + 1]) + float(k) * (dx * B[3 * r + 2]);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int ix = b[0] + i, iy = b[1] + j, iz = b[2] + k;
                if (!in_grid(sc, ix, iy, iz)) continue;
                const float4 g = a.gv[cell_index(sc, ix, iy, iz)];
                const float h = g.x * (u[0] + float(i) * b0[0]) + g.y * (u[1] + float(i) * b0[1]) +
                                g.z * (u[2] + float(i) * b0[2]);
                Sw += w[0][i] * h;
                Sdw += dw[0][i] * h;
                rs[0] += w[0][i] * g.x;
                rs[1] += w[0][i] * g.y;
                rs[2] += w[0][i] * g.z;
                if (i > 0) {
                    ri[0] += wi[i] * g.x;
                    ri[1] += wi[i] * g.y;
                    ri[2] += wi[i] * g.z;
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                vn[r] += wjk * rs[r];
                Mv[3 * r] += wjk * ri[r];
                Mv[3 * r + 1] += float(j) * wjk * rs[r];
                Mv[3 * r + 2] += float(k) * wjk * rs[r];
            }
            acc[0] += wjk * Sdw;
            acc[1] += dw[1][j] * w[2][k] * Sw;
            acc[2] += w[1][j] * dw[2][k] * Sw;
        }
    float Cn[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Cn[3 * r + c] = s4 * dx * (Mv[3 * r + c] - vn[r] * fx[c]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        acc[k] = acc[k] * idx - (B[k] * vn[0] + B[3 + k] * vn[1] + B[6 + k] * vn[2]);
        gx[k] += acc[k];
    }
    // dJ/dF (pre) = (I + dt C')^T dJ/dF'
    float nF[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            nF[3 * i + j] = gF[3 * i + j] + dt * (Cn[i] * gF[j] + Cn[3 + i] * gF[3 + j] + Cn[6 + i] * gF[6 + j]);
#pragma unroll
    for (int k = 0; k < 9; ++k) gF[k] = nF[k];
    // dJ/d(grid v_i) += w (dJ/dv' + B d): the P2G form with m = 0
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sp.b[k] = b[k];
        sp.fx[k] = fx[k];
        sp.mv[k] = gvp[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) sp.A[k] = B[k];
    sp.m = 0.f;
    sp.on = true;
    sp.esc = stencil_escapes(sc, ch, a.sorted_keys[slot], sp.b);
}

// Reverse of P2G, of the stress and of the SVD for the particle in `slot`
// of the pre-state a.ps, with the P2G affine A the forward scattered (the
// planes `aff`), gathering (dJ/dm, dJ/dp) from gadj
// (mpm_adjoint.rs:389-487, mpm_cpu_diff.rs:330-379, dense3.rs:321-372).  On
// entry gx, gF hold the reverse G2P's output; on return gx, gv, gC, gF are
// dJ/d(pre-state), gE has dJ/dE_p added, and (amu, alam) are the particle's
// dJ/dmu, dJ/dlambda.
__device__ __forceinline__ void adj_p2g_particle(const StepArgs& a, const float4* __restrict__ gadj,
                                                 const float* __restrict__ aff, const SceneDev& sc, uint32_t slot,
                                                 float* gx, float* gF, float* gv, float* gC, float& gE, double& amu,
                                                 double& alam) {
    const uint32_t cap = a.cap;
    const float* __restrict__ ps = a.ps;
    const float dt = a.dt;
    float x[3], v[3], A[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = ps[(size_t)(CX + k) * cap + slot];
        v[k] = ps[(size_t)(CV + k) * cap + slot];
    }
    // A = m C + 4/dx^2 stress (stress = -dt vol P F^T) as the forward scattered
    // it: the tape holds it, so the FP64 SVD is not live across the gather
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = aff[(size_t)k * cap + slot];
    const float m = ps[(size_t)CMASS * cap + slot];
    float mv[3] = {m * v[0], m * v[1], m * v[2]};
    float xp[3], fx[3], w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xp[k] = (x[k] - sc.origin[k]) * sc.inv_dx;
        const float bf = floorf(xp[k] - 0.5f);
        b[k] = int(bf);
        fx[k] = xp[k] - bf;
        quad_weights(fx[k], w[k]);
        dw[k][0] = -(1.5f - fx[k]);
        dw[k][1] = -2.0f * (fx[k] - 1.0f);
        dw[k][2] = fx[k] - 0.5f;
    }
    // P2G reverse (mpm_adjoint.rs:389-487): with gw = sum w dJ/dp_i and
    // T = sum w dJ/dp_i d^T:  dJ/dv = m gw, dJ/dC = m T, dJ/dstress = 4/dx^2 T,
    // dJ/dx = sum grad(w) (dJ/dp_i . (m v + A d) + m dJ/dm_i) - A^T gw,
    // summed row by row as in the reverse G2P (A d = u_jk - m v + i a0,
    // a0 = dx A[:, 0]) with T = dx (Mt - gw fx^T), Mt = sum w dJ/dp_i (i, j, k)^T.
    const float dx = sc.dx, idx = sc.inv_dx;
    float a0[3], u0[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        a0[r] = dx * A[3 * r];
        u0[r] = mv[r] - (a0[r] * fx[0] + dx * A[3 * r + 1] * fx[1] + dx * A[3 * r + 2] * fx[2]);
    }
    const float wi[3] = {0.f, w[0][1], 2.0f * w[0][2]};
    float gw[3] = {0.f, 0.f, 0.f}, Mt[9], gpos[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 9; ++k) Mt[k] = 0.f;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            const float wjk = w[1][j] * w[2][k];
            float u[3], rs[3] = {0.f, 0.f, 0.f}, ri[3] = {0.f, 0.f, 0.f}, Sw = 0.f, Sdw = 0.f;
#pragma unroll
            for (int r = 0; r < 3; ++r) u[r] = u0[r] + float(j) * (dx * A[3 * r + 1]) + float(k) * (dx * A[3 * r + 2]);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int ix = b[0] + i, iy = b[1] + j, iz = b[2] + k;
                if (!in_grid(sc, ix, iy, iz)) continue;
                const float4 g = gadj[cell_index(sc, ix, iy, iz)];
                const float h = m * g.x + g.y * (u[0] + float(i) * a0[0]) + g.z * (u[1] + float(i) * a0[1]) +
                                g.w * (u[2] + float(i) * a0[2]);
                Sw += w[0][i] * h;
                Sdw += dw[0][i] * h;
                rs[0] += w[0][i] * g.y;
                rs[1] += w[0][i] * g.z;
                rs[2] += w[0][i] * g.w;
                if (i > 0) {
                    ri[0] += wi[i] * g.y;
                    ri[1] += wi[i] * g.z;
                    ri[2] += wi[i] * g.w;
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                gw[r] += wjk * rs[r];
                Mt[3 * r] += wjk * ri[r];
                Mt[3 * r + 1] += float(j) * wjk * rs[r];
                Mt[3 * r + 2] += float(k) * wjk * rs[r];
            }
            gpos[0] += wjk * Sdw;
            gpos[1] += dw[1][j] * w[2][k] * Sw;
            gpos[2] += w[1][j] * dw[2][k] * Sw;
        }
    float T[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) T[3 * r + c] = dx * (Mt[3 * r + c] - gw[r] * fx[c]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gpos[k] = gpos[k] * idx - (A[k] * gw[0] + A[3 + k] * gw[1] + A[6 + k] * gw[2]);
        gx[k] += gpos[k];
        gv[k] = m * gw[k];
    }
    // corotated-SVK PK1 from the FP64 SVD (mpm_cpu_diff.rs:279-287)
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = ps[(size_t)(CF + k) * cap + slot];
    const float vol = ps[(size_t)CVOL * cap + slot];
    float mu = sc.mu, lam = sc.lam;
    if (a.has_E) {
        const float E = ps[(size_t)CE * cap + slot];
        mu = E / (2.0f * (1.0f + sc.poisson));
        lam = E * sc.poisson / ((1.0f + sc.poisson) * (1.0f - 2.0f * sc.poisson));
    }
    double U[9], s[3], V[9];
    svd3_f64_raw(F, U, s, V);
    svd3_signs<MODEL_CPUDIFF>(U, s, V);
    const double dmu = mu, dlam = lam;
    const double trsi = s[0] + s[1] + s[2] - 3.0;
    Dm3 P;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double r = U[3 * i] * V[3 * j] + U[3 * i + 1] * V[3 * j + 1] + U[3 * i + 2] * V[3 * j + 2];
            const double fr = U[3 * i] * s[0] * V[3 * j] + U[3 * i + 1] * s[1] * V[3 * j + 1] +
                              U[3 * i + 2] * s[2] * V[3 * j + 2];
            P(i, j) = 2.0 * dmu * (fr - r) + dlam * trsi * r;
        }
    const float s4 = 4.0f * sc.inv_dx * sc.inv_dx;
    const double coef = -(double)dt * (double)vol;
#pragma unroll
    for (int k = 0; k < 9; ++k) gC[k] = m * T[k];
    // stress = coef P F^T with dJ/dstress = s4 T:
    //   dJ/dP = coef s4 T F,  dJ/dF += coef s4 T^T P
    Dm3 aP, gf;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            aP(i, j) = coef * (double)s4 *
                       ((double)T[3 * i] * F[j] + (double)T[3 * i + 1] * F[3 + j] + (double)T[3 * i + 2] * F[6 + j]);
            gf(i, j) = (double)gF[3 * i + j] +
                       coef * (double)s4 *
                           ((double)T[i] * P(0, j) + (double)T[3 + i] * P(1, j) + (double)T[6 + i] * P(2, j));
        }
    // corotated_pk1_adjoint (mpm_cpu_diff.rs:330-379), in the rotated
    // frame: with Ut aP V = Q, aU = aP V (2mu Sd + cr I) - ..., the SVD
    // adjoint only needs Ut aU, Vt aV, so form them directly.
    Dm3 Q;  // U^T aP V
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double pv = aP(r, 0) * V[j] + aP(r, 1) * V[3 + j] + aP(r, 2) * V[6 + j];
                acc += U[3 * r + i] * pv;
            }
            Q(i, j) = acc;
        }
    const double cr = -2.0 * dmu + dlam * trsi;
    const double r_dot = Q(0, 0) + Q(1, 1) + Q(2, 2);  // <R, aP> = tr(U^T aP V)
    // aU = aP V (2mu Sd) + cr aP V  ->  Ut aU = Q (2mu Sd + cr I)
    // aV = aP^T U (2mu Sd) + cr aP^T U  ->  Vt aV = Q^T (2mu Sd + cr I)
    Dm3 UaU, VaV;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double dj = 2.0 * dmu * s[j] + cr;
            UaU(i, j) = Q(i, j) * dj;
            VaV(i, j) = Q(j, i) * dj;
        }
    double aS[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) aS[i] = 2.0 * dmu * Q(i, i) + dlam * r_dot;
    // <Fr - R, aP> = sum_i (s_i - 1) Q_ii
    amu = 2.0 * ((s[0] - 1.0) * Q(0, 0) + (s[1] - 1.0) * Q(1, 1) + (s[2] - 1.0) * Q(2, 2));
    alam = trsi * r_dot;
    // adj_svd3 (dense3.rs:321-372): dJ/dF += U M V^T,
    //   M = diag(aS) + (f o om_u) Sd + Sd (f o om_v),
    //   om_u = Ut aU - aU^T U = UaU - UaU^T, om_v = VaV - VaV^T
    const double f01 = safe_inv_sq_diff(s[0], s[1]);
    const double f02 = safe_inv_sq_diff(s[0], s[2]);
    const double f12 = safe_inv_sq_diff(s[1], s[2]);
    Dm3 Mx;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double f = 0.0;
            if (i == 0 && j == 1) f = f01;
            if (i == 0 && j == 2) f = f02;
            if (i == 1 && j == 2) f = f12;
            if (i == 1 && j == 0) f = -f01;
            if (i == 2 && j == 0) f = -f02;
            if (i == 2 && j == 1) f = -f12;
            const double omu = UaU(i, j) - UaU(j, i);
            const double omv = VaV(i, j) - VaV(j, i);
            Mx(i, j) = (i == j ? aS[i] : 0.0) + f * omu * s[j] + s[i] * f * omv;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = gf(i, j);
#pragma unroll
            for (int r = 0; r < 3; ++r)
                acc += U[3 * i + r] * (Mx(r, 0) * V[3 * j] + Mx(r, 1) * V[3 * j + 1] + Mx(r, 2) * V[3 * j + 2]);
            gF[3 * i + j] = (float)acc;
        }
    // dJ/dE_p (per-particle E, chain rule through lame, mpm_cpu_diff.rs:217-222)
    const double nu = sc.poisson;
    const double ge = amu / (2.0 * (1.0 + nu)) + alam * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    gE += (float)ge;
}

// Block sum of the particles' (dJ/dmu, dJ/dlambda) added to chunk ci's
// accumulator: one block owns a chunk, so the order is fixed.  Every thread
// of the block calls it.
__device__ __forceinline__ void add_chunk_mu_lambda(double* gml, uint32_t ci, double amu, double alam,
                                                    double (*red)[WAVES]) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    amu = warp_sum(amu);
    alam = warp_sum(alam);
    if (lane == 0) {
        red[0][wave] = amu;
        red[1][wave] = alam;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0, sl = 0.0;
        for (int w = 0; w < WAVES; ++w) {
            sm += red[0][w];
            sl += red[1][w];
        }
        gml[2 * (size_t)ci] += sm;
        gml[2 * (size_t)ci + 1] += sl;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_adj_g2p(AdjArgs q) {
    __shared__ ScatterLdsT<true> L;
    const StepArgs& a = q.a;
    const uint32_t cap = a.cap;
    float* __restrict__ adj = q.adj;
    const uint32_t nch = *a.n_chunks_dev;
    for (ChunkWalk w = chunk_walk(nch); w.ci < w.end; w.ci += w.step) {
        const uint32_t ci = w.ci;
        const Chunk ch = a.chunks[ci];
        const SceneDev sc = a.scenes[ch.scene];
        const uint32_t slot = ch.start + threadIdx.x;
        Scatter sp;
        sp.on = false;
        sp.esc = false;
        if (threadIdx.x < ch.count) {
            float gx[3], gv[3], gC[9], gF[9];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                gx[k] = adj[(size_t)(AX + k) * cap + slot];
                gv[k] = adj[(size_t)(AV + k) * cap + slot];
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                gC[k] = adj[(size_t)(AC + k) * cap + slot];
                gF[k] = adj[(size_t)(AF + k) * cap + slot];
            }
            adj_g2p_particle(a, sc, ch, slot, gx, gv, gC, gF, sp);
#pragma unroll
            for (int k = 0; k < 3; ++k) adj[(size_t)(AX + k) * cap + slot] = gx[k];
#pragma unroll
            for (int k = 0; k < 9; ++k) adj[(size_t)(AF + k) * cap + slot] = gF[k];
        }
        block_scatter<false, true>(a, sc, sp, ci, L);
    }
}

// ---------------------------------------------------------------------------
// Grid reverse (mpm_adjoint.rs:324-363).  gmp holds the forward's P2G total
// (m, p) of every listed tile; gadj holds the atomic part of dJ/d(grid v)
// and receives (dJ/dm, dJ/dp).
// SPLIT: the block's four waves sum four ranges of each tile's candidate
// chunks and wave 0 adds them in order (small scenes; as k_grid_update).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_adj_grid(const SceneDev* __restrict__ scenes, GridDev g, float4* gadj,
                                                  const uint32_t* __restrict__ dlist,
                                                  const uint32_t* __restrict__ dcount, float dt, int use_esc) {
    constexpr int R = 4;
    __shared__ float4 wsum[SPLIT ? R - 1 : 1][64];
    const uint32_t ns = *g.scount, nd = *dcount;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const uint32_t first = SPLIT ? blockIdx.x : blockIdx.x * 4 + uint32_t(wave);
    const uint32_t stride = SPLIT ? gridDim.x : gridDim.x * 4;
    for (uint32_t i = first; i < ns + nd; i += stride) {
        TileRec r;
        load_tile_rec(g, ns, i, r);
        const uint32_t t = r.valid ? r.t : (i < ns ? g.slist[i] : dlist[i - ns]);
        const SceneDev sc = scenes[r.valid ? r.scene : g.tile_scene[t]];
        int tx, ty, tz, ix, iy, iz;
        tile_cell(sc, t, lane, tx, ty, tz, ix, iy, iz);
        const uint32_t c = (t << 6) | uint32_t(lane);
        float4 q = use_esc && (!SPLIT || wave == 0) ? gadj[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        add_partials(g, sc, tx, ty, tz, ix, iy, iz, q, SPLIT ? wave : 0, SPLIT ? R : 1, r.valid, r.rng);
        if (SPLIT) {
            if (wave > 0) wsum[wave - 1][lane] = q;
            __syncthreads();
            if (wave == 0) {
#pragma unroll
                for (int r = 0; r < R - 1; ++r) {
                    const float4 v = wsum[r][lane];
                    q.x += v.x;
                    q.y += v.y;
                    q.z += v.z;
                    q.w += v.w;
                }
            }
            __syncthreads();  // wsum is reused by the next tile
            if (wave != 0) continue;
        }
        const float4 mp = g.gmp[c];
        const float m = mp.x;
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        if (m > 0.0f) {
            float gx = q.y, gy = q.z, gz = q.w;
            if (below_floor(sc, iy)) {
                float vy = mp.z / m;  // same operations as the forward's floor test
                vy += sc.gravity_y * dt;
                if (vy < 0.f) gy = 0.f;
                const float damp = 1.0f - sc.floor_friction;
                gx *= damp;
                gz *= damp;
            }
            const float im = 1.0f / m;
            out.y = gx * im;
            out.z = gy * im;
            out.w = gz * im;
            out.x = -(mp.y * gx + mp.z * gy + mp.w * gz) * im * im;
        }
        gadj[c] = out;
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_adj_p2g(AdjArgs q) {
    __shared__ double red[2][WAVES];
    const StepArgs& a = q.a;
    const uint32_t cap = a.cap;
    float* __restrict__ adj = q.adj;
    const uint32_t nch = *a.n_chunks_dev;
    for (ChunkWalk w = chunk_walk(nch); w.ci < w.end; w.ci += w.step) {
        const uint32_t ci = w.ci;
        const Chunk ch = a.chunks[ci];
        const SceneDev sc = a.scenes[ch.scene];
        const uint32_t slot = ch.start + threadIdx.x;
        double amu = 0.0, alam = 0.0;
        if (threadIdx.x < ch.count) {
            float gx[3], gv[3], gC[9], gF[9];
#pragma unroll
            for (int k = 0; k < 3; ++k) gx[k] = adj[(size_t)(AX + k) * cap + slot];
#pragma unroll
            for (int k = 0; k < 9; ++k) gF[k] = adj[(size_t)(AF + k) * cap + slot];
            float gE = adj[(size_t)AE * cap + slot];
            adj_p2g_particle(a, q.gadj, q.aff, sc, slot, gx, gF, gv, gC, gE, amu, alam);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                adj[(size_t)(AX + k) * cap + slot] = gx[k];
                adj[(size_t)(AV + k) * cap + slot] = gv[k];
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                adj[(size_t)(AC + k) * cap + slot] = gC[k];
                adj[(size_t)(AF + k) * cap + slot] = gF[k];
            }
            adj[(size_t)AE * cap + slot] = gE;
        }
        add_chunk_mu_lambda(q.gml, ci, amu, alam, red);
    }
}

// Reverse P2G of substep k fused with the reverse G2P of substep k-1 for
// the same particle: dJ/d(state k) stays in registers.  Valid when no
// particle of either substep scatters with global atomics (the host checks
// the recompute's escape and fallback counters), so gadj is only read here.
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_adj_fused(AdjArgs q, const float* ps_prev, const float4* gv_prev) {
    __shared__ ScatterLdsT<true> L;
    __shared__ double red[2][WAVES];
    const StepArgs& a = q.a;
    StepArgs ap = q.a;
    ap.ps = ap.ps_out = const_cast<float*>(ps_prev);
    ap.gv = const_cast<float4*>(gv_prev);
    const uint32_t cap = a.cap;
    float* __restrict__ adj = q.adj;
    const uint32_t nch = *a.n_chunks_dev;
    for (ChunkWalk w = chunk_walk(nch); w.ci < w.end; w.ci += w.step) {
        const uint32_t ci = w.ci;
        const Chunk ch = a.chunks[ci];
        const SceneDev sc = a.scenes[ch.scene];
        const uint32_t slot = ch.start + threadIdx.x;
        const bool active = threadIdx.x < ch.count;
        double amu = 0.0, alam = 0.0;
        float gx[3], gv[3], gC[9], gF[9];
        if (active) {
#pragma unroll
            for (int k = 0; k < 3; ++k) gx[k] = adj[(size_t)(AX + k) * cap + slot];
#pragma unroll
            for (int k = 0; k < 9; ++k) gF[k] = adj[(size_t)(AF + k) * cap + slot];
            float gE = adj[(size_t)AE * cap + slot];
            adj_p2g_particle(a, q.gadj, q.aff, sc, slot, gx, gF, gv, gC, gE, amu, alam);
            adj[(size_t)AE * cap + slot] = gE;
        }
        add_chunk_mu_lambda(q.gml, ci, amu, alam, red);
        Scatter sp;
        sp.on = false;
        sp.esc = false;
        if (active) {
            adj_g2p_particle(ap, sc, ch, slot, gx, gv, gC, gF, sp);
#pragma unroll
            for (int k = 0; k < 3; ++k) adj[(size_t)(AX + k) * cap + slot] = gx[k];
#pragma unroll
            for (int k = 0; k < 9; ++k) adj[(size_t)(AF + k) * cap + slot] = gF[k];
        }
        block_scatter<false, true>(ap, sc, sp, ci, L);
    }
}

hipError_t launch_adj_g2p(const AdjArgs& q, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (blocks > int(NUM_XCD)) blocks &= ~int(NUM_XCD - 1);
    hipLaunchKernelGGL(k_adj_g2p, dim3(blocks), dim3(BLOCK), 0, s, q);
    return hipGetLastError();
}

hipError_t launch_adj_grid(const SceneDev* scenes, const GridDev& g, float4* gadj, const uint32_t* dlist,
                           const uint32_t* dcount, float dt, int use_esc, int blocks, hipStream_t s, int split) {
    if (split)
        hipLaunchKernelGGL(k_adj_grid<true>, dim3(blocks), dim3(256), 0, s, scenes, g, gadj, dlist, dcount, dt,
                           use_esc);
    else
        hipLaunchKernelGGL(k_adj_grid<false>, dim3(blocks), dim3(256), 0, s, scenes, g, gadj, dlist, dcount, dt,
                           use_esc);
    return hipGetLastError();
}

hipError_t launch_adj_fused(const AdjArgs& q, const float* ps_prev, const float4* gv_prev, int blocks,
                            hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (blocks > int(NUM_XCD)) blocks &= ~int(NUM_XCD - 1);
    hipLaunchKernelGGL(k_adj_fused, dim3(blocks), dim3(BLOCK), 0, s, q, ps_prev, gv_prev);
    return hipGetLastError();
}

hipError_t launch_adj_p2g(const AdjArgs& q, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (blocks > int(NUM_XCD)) blocks &= ~int(NUM_XCD - 1);
    hipLaunchKernelGGL(k_adj_p2g, dim3(blocks), dim3(BLOCK), 0, s, q);
    return hipGetLastError();
}

}  // namespace mpm
