// mpm_device.hpp -- per-particle device math for the MPM substep on gfx950.
//
// * quadratic B-spline weights (mpm.wgsl:287-294)
// * FP64 3x3 SVD by cyclic Jacobi on F^T F: the rotation angle is estimated
//   with the hardware f32 rcp/sqrt/rsq and the rotation is renormalised and
//   applied in f64, so the factorisation is accurate to f64 rounding while
//   avoiding f64 division and square root (each slow on CDNA).  sigma_i and
//   1/sigma_i come from an f32 rsqrt seed refined by two f64 Newton steps.
// * sign conventions of the two reference SVDs (shader: det U = det V = +1
//   with the smallest sigma negated for inverted F, mpm.wgsl:445-466; CPU:
//   sigma >= 0 with the smallest-sigma columns of U and V flipped together,
//   dense3.rs:126-140)
// * Kirchhoff-form stresses: for F = U diag(s) V^T,
//     fixed-corotated  (mpm.wgsl:764-826):  P F^T = U diag(tau) U^T,
//       tau_i = 2 mu h (s_i - 1) s_i + lam h (J - 1) J   (s clamped [0.1,10])
//     corotated-SVK    (mpm_cpu_diff.rs:279-287, :462-464):
//       tau_i = (2 mu (s_i - 1) + lam (sum s - 3)) s_i
//   so the stress needs only U (no V, no cofactor matrix).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

struct Svd32 {
    float U[9];
    float s[3];
    float V[9];
};

__device__ __forceinline__ void quad_weights(float fx, float* w) {
    float a = 1.5f - fx, b = fx - 1.0f, c = fx - 0.5f;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - b * b;
    w[2] = 0.5f * c * c;
}

__device__ __forceinline__ bool finitef(float v) {
    return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u;
}

// Accumulates x * 0, which is 0 for every finite x and NaN for an infinite
// or NaN one: one FMA per value, one test at the end (finite_acc_ok).
__device__ __forceinline__ float finite_acc(float acc, float x) { return __builtin_fmaf(x, 0.0f, acc); }
__device__ __forceinline__ bool finite_acc_ok(float acc) { return acc == acc; }

__device__ __forceinline__ double rsqrt_f64(double x) {
    double r = (double)__frsqrt_rn((float)x);
    r = r * (1.5 - 0.5 * x * r * r);
    r = r * (1.5 - 0.5 * x * r * r);
    return r;
}

// One Jacobi rotation on pair (p,q) with third index k of the symmetric
// matrix held in (a_pp, a_qq, a_pq, a_kp, a_kq); V columns rotated likewise.
struct JacobiPair {
    __device__ static __forceinline__ void rot(double& app, double& aqq, double& apq, double& akp,
                                               double& akq, double* vp, double* vq, double tol) {
        if (fabs(apq) <= tol) return;
        // Angle from the hardware rcp/sqrt/rsq estimates: any rotation near
        // the Jacobi angle converges, and (c, s) is renormalised in f64 below.
        float th = (float)(aqq - app) * __builtin_amdgcn_rcpf(2.0f * (float)apq);
        float t = __builtin_amdgcn_rcpf(fabsf(th) + __builtin_amdgcn_sqrtf(1.0f + th * th));
        t = th >= 0.0f ? t : -t;
        float cf = __builtin_amdgcn_rsqf(1.0f + t * t);
        double c = cf, s = (double)(t * cf);
        double r = c * c + s * s;
        double k = 1.5 - 0.5 * r;
        c *= k;
        s *= k;
        double cc = c * c, ss = s * s, cs = c * s;
        double napp = cc * app - 2.0 * cs * apq + ss * aqq;
        double naqq = ss * app + 2.0 * cs * apq + cc * aqq;
        double napq = cs * (app - aqq) + (cc - ss) * apq;
        double nakp = c * akp - s * akq;
        double nakq = s * akp + c * akq;
        app = napp;
        aqq = naqq;
        apq = napq;
        akp = nakp;
        akq = nakq;
#pragma unroll
        for (int r2 = 0; r2 < 3; ++r2) {
            double a0 = vp[r2], b0 = vq[r2];
            vp[r2] = c * a0 - s * b0;
            vq[r2] = s * a0 + c * b0;
        }
    }
};

// F (row-major f32) -> U, sigma (>= 0), V in f64 with V proper.  The sweeps
// stop when every off-diagonal entry A_ij of the rotated F^T F is below
// rel_tol * trace: 1e-15 keeps the factors at f64 accuracy (the adjoint uses
// them in f64).  For results rounded to f32 a looser bound is exact enough:
// V stays orthonormal to f64 rounding (it is a product of rotations), and
// U = F V diag(1/sigma) makes U diag(sigma) V^T = F to f64 rounding whatever
// the bound, while u_i . u_j = A_ij / (sigma_i sigma_j) bounds U's departure
// from orthonormality; with 1e-8 that is ~3e-8 for sigma near 1, below f32
// rounding (6e-8), growing as trace / (sigma_i sigma_j) for a strongly
// stretched or compressed F (2e-7 at sigma = (0.1, 1, 1)).  The reference
// shader's own f32 Jacobi stops at 1e-7 (mpm.wgsl:349-468).
__device__ __forceinline__ void svd3_f64_raw(const float* F, double* U, double* sg, double* V,
                                             double rel_tol = 1.0e-15) {
    // F stays f32 (exact in f64) and is widened where used, so no f64 copy
    // of it is live across the sweeps.
#define MPM_F(i) ((double)F[i])
    // A = F^T F; column j of F is (F[j], F[3+j], F[6+j]).
    double a00 = MPM_F(0) * MPM_F(0) + MPM_F(3) * MPM_F(3) + MPM_F(6) * MPM_F(6);
    double a11 = MPM_F(1) * MPM_F(1) + MPM_F(4) * MPM_F(4) + MPM_F(7) * MPM_F(7);
    double a22 = MPM_F(2) * MPM_F(2) + MPM_F(5) * MPM_F(5) + MPM_F(8) * MPM_F(8);
    double a01 = MPM_F(0) * MPM_F(1) + MPM_F(3) * MPM_F(4) + MPM_F(6) * MPM_F(7);
    double a02 = MPM_F(0) * MPM_F(2) + MPM_F(3) * MPM_F(5) + MPM_F(6) * MPM_F(8);
    double a12 = MPM_F(1) * MPM_F(2) + MPM_F(4) * MPM_F(5) + MPM_F(7) * MPM_F(8);
    // V stored by columns: v0, v1, v2.
    double v0[3] = {1.0, 0.0, 0.0}, v1[3] = {0.0, 1.0, 0.0}, v2[3] = {0.0, 0.0, 1.0};
    const double tol = rel_tol * (a00 + a11 + a22);
    for (int sweep = 0; sweep < 8; ++sweep) {
        double off = fmax(fabs(a01), fmax(fabs(a02), fabs(a12)));
        if (off <= tol) break;
        JacobiPair::rot(a00, a11, a01, a02, a12, v0, v1, tol);  // (0,1), k = 2
        JacobiPair::rot(a00, a22, a02, a01, a12, v0, v2, tol);  // (0,2), k = 1
        JacobiPair::rot(a11, a22, a12, a01, a02, v1, v2, tol);  // (1,2), k = 0
    }
    double* vc[3] = {v0, v1, v2};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double w0 = MPM_F(0) * vc[i][0] + MPM_F(1) * vc[i][1] + MPM_F(2) * vc[i][2];
        double w1 = MPM_F(3) * vc[i][0] + MPM_F(4) * vc[i][1] + MPM_F(5) * vc[i][2];
        double w2 = MPM_F(6) * vc[i][0] + MPM_F(7) * vc[i][1] + MPM_F(8) * vc[i][2];
        double n2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (n2 > 1.0e-24) {
            double r = rsqrt_f64(n2);
            sg[i] = n2 * r;
            U[0 * 3 + i] = w0 * r;
            U[1 * 3 + i] = w1 * r;
            U[2 * 3 + i] = w2 * r;
        } else {
            sg[i] = sqrt(n2);
            U[0 * 3 + i] = i == 0 ? 1.0 : 0.0;
            U[1 * 3 + i] = i == 1 ? 1.0 : 0.0;
            U[2 * 3 + i] = i == 2 ? 1.0 : 0.0;
        }
        V[0 * 3 + i] = vc[i][0];
        V[1 * 3 + i] = vc[i][1];
        V[2 * 3 + i] = vc[i][2];
    }
#undef MPM_F
}

__device__ __forceinline__ double det3d(const double* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

__device__ __forceinline__ int argmin3(double a, double b, double c) {
    int k = 0;
    double m = a;
    if (b < m) { m = b; k = 1; }
    if (c < m) { k = 2; }
    return k;
}

// Sign fix of svd3_f64_raw's output.  model 0: shader convention; model 1:
// CPU convention.
template <int MODEL>
__device__ __forceinline__ void svd3_signs(double* U, double* s, double* V) {
    if (det3d(U) < 0.0) {
        int k = argmin3(s[0], s[1], s[2]);
        U[k] = -U[k];
        U[3 + k] = -U[3 + k];
        U[6 + k] = -U[6 + k];
        if (MODEL == 0) {
            s[k] = -s[k];
        } else {
            V[k] = -V[k];
            V[3 + k] = -V[3 + k];
            V[6 + k] = -V[6 + k];
        }
    }
}

// FP64 SVD with the sign convention of MODEL, rounded to f32.  The shader
// model's factors are only used in f32 (bound 1e-8: a sweep fewer than 1e-10
// for strained elastic particles, 6.7 % off the fused pass of a 512-scene
// batch of elastic cubes); the differentiable model's forward must match its
// adjoint's f64 recompute, so it keeps the f64 tolerance.
template <int MODEL>
__device__ __forceinline__ void svd3(const float* F, Svd32& o) {
    double U[9], s[3], V[9];
    svd3_f64_raw(F, U, s, V, MODEL == 0 ? 1.0e-8 : 1.0e-15);
    svd3_signs<MODEL>(U, s, V);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        o.U[i] = (float)U[i];
        o.V[i] = (float)V[i];
    }
    o.s[0] = (float)s[0];
    o.s[1] = (float)s[1];
    o.s[2] = (float)s[2];
}

// out = U diag(d) V^T (row-major), f32.
__device__ __forceinline__ void udvt(const float* U, const float* d, const float* V, float* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            out[3 * i + j] = U[3 * i] * d[0] * V[3 * j] + U[3 * i + 1] * d[1] * V[3 * j + 1] +
                             U[3 * i + 2] * d[2] * V[3 * j + 2];
}

// out = U diag(d) U^T, symmetric (row-major, all 9 written).
__device__ __forceinline__ void udut(const float* U, const float* d, float* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            float v = U[3 * i] * d[0] * U[3 * j] + U[3 * i + 1] * d[1] * U[3 * j + 1] +
                      U[3 * i + 2] * d[2] * U[3 * j + 2];
            out[3 * i + j] = v;
            out[3 * j + i] = v;
        }
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    return fminf(fmaxf(v, lo), hi);
}

}  // namespace mpm
