// mpm_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of molla's MPM substep, written from the reference's
// source (bresilla/molla) so the HIP product path can be checked against
// it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load this library; the product (molla_amd/) never does.
//
// Two semantics are restated, selected by orc_scene::model:
//   * SHADER : crates/molla-solvers/shaders/mpm.wgsl (the GPU step:
//              fixed-corotated, sigma clamps, 6 padded walls, plasticity).
//   * CPUDIFF: crates/molla-autograd/src/mpm_cpu_diff.rs (the reference CPU
//              path: corotated-SVK, floor plane, no clamps) and its adjoint
//              crates/molla-autograd/src/mpm_adjoint.rs.
//
// Deliberate, documented deviations from a literal restatement:
//   * Per-particle mass/volume arrays and an optional per-particle Young's
//     modulus (the reference uses config-global values; uniform arrays
//     reproduce it exactly).
//   * Rheology 3 ("fluid") has no reference counterpart: fixed-corotated with
//     mu = 0, lambda = bulk and F reset to cbrt(det F) * I after each substep.
//   * The f64 instantiation of the shader SVD tightens the Jacobi tolerance
//     from the f32 shader's 1e-7 to 1e-14 so that it is a true FP64 reference.
//   * kdiag (implicit-Jacobi preconditioner, unused by explicit stepping) is
//     not accumulated.
//
// Parity status per function is listed in the oracle's test file header.
// Compiled with -ffp-contract=off so every operation rounds once.

#include "mpm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <omp.h>

// Planted mistakes (orc_set_mutation), for the tests that show a closed-form
// check fails when the rule it pins is perturbed.  0 = the faithful oracle.
enum Mutation : int {
    MUT_NONE = 0,
    MUT_PAD_ONE = 1,          // wall pad of 1 cell instead of 2
    MUT_FRICTION_NORMAL = 2,  // boundary friction also scales the normal component
    MUT_FRICTION_ONE_TAN = 3, // boundary friction on one tangential component only
    MUT_CLAMP_LO = 4,         // position clamp low bound at window + 1
    MUT_CLAMP_KEEP_V = 5,     // position clamp keeps the offending velocity
    MUT_STRESS_CLAMP = 6,     // stress sigma clamp at [0.05, 10]
    MUT_PLASTIC_CLAMP = 7,    // plasticity sigma clamp at [0.05, 10]
    MUT_FLUID_SQRT = 8,       // fluid reset to sqrt(J) I
    MUT_STRESS_COF = 9,       // volumetric stress term on R instead of the cofactor
};
static int g_mut = MUT_NONE;

namespace {

template <typename T>
struct M3 {
    T a[9];
    T& operator()(int i, int j) { return a[3 * i + j]; }
    T operator()(int i, int j) const { return a[3 * i + j]; }
};

template <typename T>
struct V3 {
    T x[3];
    T& operator[](int i) { return x[i]; }
    T operator[](int i) const { return x[i]; }
};

template <typename T>
M3<T> zero3() {
    M3<T> m;
    for (int i = 0; i < 9; ++i) m.a[i] = T(0);
    return m;
}
template <typename T>
M3<T> eye3() {
    M3<T> m = zero3<T>();
    m(0, 0) = m(1, 1) = m(2, 2) = T(1);
    return m;
}
template <typename T>
M3<T> mul(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T s = A(i, 0) * B(0, j);
            s = s + A(i, 1) * B(1, j);
            s = s + A(i, 2) * B(2, j);
            C(i, j) = s;
        }
    return C;
}
template <typename T>
M3<T> tr(const M3<T>& A) {
    M3<T> B;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B(i, j) = A(j, i);
    return B;
}
template <typename T>
M3<T> add(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int i = 0; i < 9; ++i) C.a[i] = A.a[i] + B.a[i];
    return C;
}
template <typename T>
M3<T> sub(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int i = 0; i < 9; ++i) C.a[i] = A.a[i] - B.a[i];
    return C;
}
template <typename T>
M3<T> scale(T s, const M3<T>& A) {
    M3<T> C;
    for (int i = 0; i < 9; ++i) C.a[i] = s * A.a[i];
    return C;
}
template <typename T>
V3<T> mv(const M3<T>& A, const V3<T>& v) {
    V3<T> r;
    for (int i = 0; i < 3; ++i) {
        T s = A(i, 0) * v[0];
        s = s + A(i, 1) * v[1];
        s = s + A(i, 2) * v[2];
        r[i] = s;
    }
    return r;
}
template <typename T>
T det3(const M3<T>& m) {
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
           m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
template <typename T>
M3<T> udv(const M3<T>& U, const V3<T>& d, const M3<T>& V) {
    // U * diag(d) * V^T
    M3<T> R;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T s = U(i, 0) * d[0] * V(j, 0);
            s = s + U(i, 1) * d[1] * V(j, 1);
            s = s + U(i, 2) * d[2] * V(j, 2);
            R(i, j) = s;
        }
    return R;
}
template <typename T>
T frob(const M3<T>& A, const M3<T>& B) {
    T s = T(0);
    for (int i = 0; i < 9; ++i) s = s + A.a[i] * B.a[i];
    return s;
}
template <typename T>
bool finite3(const T* p, int k) {
    for (int i = 0; i < k; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// ---------------------------------------------------------------------------
// svd3_jacobi: crates/molla-solvers/shaders/mpm.wgsl:349-468 (givens-free
// two-sided Jacobi on M^T M, sigma_i = |M v_i|, U_i = M v_i / sigma_i, then
// det(U) and det(V) sign fixes that negate the smallest sigma).
template <typename T>
void svd3_shader(const M3<T>& M, M3<T>& U, V3<T>& s, M3<T>& V) {
    // f32: the shader's absolute 1e-7 on |a_pq| (a ~ 1 near F = I).
    // f64: the same test relative to the diagonal scale, at 1e-15.
    const bool f32 = sizeof(T) == 4;
    V = eye3<T>();
    M3<T> a = mul(tr(M), M);
    const T tol = f32 ? T(1.0e-7)
                      : T(1.0e-15) * std::max(T(1.0e-30), std::fabs(a(0, 0)) + std::fabs(a(1, 1)) +
                                                              std::fabs(a(2, 2)));
    for (int sweep = 0; sweep < 16; ++sweep) {
        T max_off = T(0);
        for (int p = 0; p < 3; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                T apq = a(p, q);
                if (std::fabs(apq) > max_off) max_off = std::fabs(apq);
                if (std::fabs(apq) < tol) continue;
                T app = a(p, p);
                T aqq = a(q, q);
                T theta = (aqq - app) / (T(2) * apq);
                T t;
                if (theta >= T(0))
                    t = T(1) / (theta + std::sqrt(T(1) + theta * theta));
                else
                    t = T(1) / (theta - std::sqrt(T(1) + theta * theta));
                T c = T(1) / std::sqrt(T(1) + t * t);
                T sn = t * c;
                a(p, p) = app - t * apq;
                a(q, q) = aqq + t * apq;
                a(p, q) = T(0);
                a(q, p) = T(0);
                for (int k = 0; k < 3; ++k) {
                    if (k != p && k != q) {
                        T akp = a(k, p);
                        T akq = a(k, q);
                        a(k, p) = c * akp - sn * akq;
                        a(k, q) = sn * akp + c * akq;
                        a(p, k) = a(k, p);
                        a(q, k) = a(k, q);
                    }
                }
                for (int k = 0; k < 3; ++k) {
                    T vkp = V(k, p);
                    T vkq = V(k, q);
                    V(k, p) = c * vkp - sn * vkq;
                    V(k, q) = sn * vkp + c * vkq;
                }
            }
        }
        if (max_off < tol) break;
    }
    U = zero3<T>();
    for (int i = 0; i < 3; ++i) {
        V3<T> vi = {{V(0, i), V(1, i), V(2, i)}};
        V3<T> m = mv(M, vi);
        T len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        s[i] = len;
        if (len > T(1.0e-12)) {
            for (int k = 0; k < 3; ++k) U(k, i) = m[k] / len;
        } else {
            for (int k = 0; k < 3; ++k) U(k, i) = (k == i) ? T(1) : T(0);
        }
    }
    if (det3(U) < T(0)) {
        int mi = 0;
        if (s[1] < s[mi]) mi = 1;
        if (s[2] < s[mi]) mi = 2;
        for (int k = 0; k < 3; ++k) U(k, mi) = -U(k, mi);
        s[mi] = -s[mi];
    }
    if (det3(V) < T(0)) {
        int mi = 0;
        if (std::fabs(s[1]) < std::fabs(s[mi])) mi = 1;
        if (std::fabs(s[2]) < std::fabs(s[mi])) mi = 2;
        for (int k = 0; k < 3; ++k) V(k, mi) = -V(k, mi);
        s[mi] = -s[mi];
    }
}

// svd3: crates/molla-math/src/dense3.rs:54-142 (one-sided cyclic Jacobi,
// <= 32 sweeps, eps 1e-14; non-negative sigma; det(U) < 0 flips the
// smallest-sigma column of both U and V).
void svd3_cpu(const M3<double>& M, M3<double>& U, V3<double>& s, M3<double>& V) {
    const double eps = 1.0e-14;
    M3<double> b = M;
    V = eye3<double>();
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = 0.0;
        for (auto& pq : pairs) {
            int p = pq[0], q = pq[1];
            double app = 0, aqq = 0, apq = 0;
            for (int k = 0; k < 3; ++k) {
                app += b(k, p) * b(k, p);
                aqq += b(k, q) * b(k, q);
                apq += b(k, p) * b(k, q);
            }
            off += apq * apq;
            if (std::fabs(apq) < eps * (std::sqrt(app) * std::sqrt(aqq) + eps)) continue;
            double tau = (aqq - app) / (2.0 * apq);
            double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                  : 1.0 / (tau - std::sqrt(1.0 + tau * tau));
            double c = 1.0 / std::sqrt(1.0 + t * t);
            double sn = t * c;
            for (int k = 0; k < 3; ++k) {
                double bp = b(k, p), bq = b(k, q);
                b(k, p) = c * bp - sn * bq;
                b(k, q) = sn * bp + c * bq;
                double vp = V(k, p), vq = V(k, q);
                V(k, p) = c * vp - sn * vq;
                V(k, q) = sn * vp + c * vq;
            }
        }
        if (off < eps * eps) break;
    }
    U = eye3<double>();
    for (int i = 0; i < 3; ++i) {
        double n = std::sqrt(b(0, i) * b(0, i) + b(1, i) * b(1, i) + b(2, i) * b(2, i));
        s[i] = n;
        if (n > 0.0) {
            for (int k = 0; k < 3; ++k) U(k, i) = b(k, i) / n;
        }
    }
    if (det3(U) < 0.0) {
        int k = 0;
        for (int i = 1; i < 3; ++i)
            if (s[i] < s[k]) k = i;
        for (int r = 0; r < 3; ++r) {
            U(r, k) = -U(r, k);
            V(r, k) = -V(r, k);
        }
    }
}

// adj_svd3: crates/molla-math/src/dense3.rs:321-372 with safe_inv_sq_diff
// (:496-499) and ADJ_SVD_EPS = 1e-12 (:295).
double safe_inv_sq_diff(double a, double b, double eps) {
    double d = b * b - a * a;
    return d / (d * d + eps * eps);
}
M3<double> adj_svd3(const M3<double>& U, const V3<double>& s, const M3<double>& V,
                    const M3<double>& aU, const V3<double>& aS, const M3<double>& aV) {
    const double eps = 1.0e-12;
    double f01 = safe_inv_sq_diff(s[0], s[1], eps);
    double f02 = safe_inv_sq_diff(s[0], s[2], eps);
    double f12 = safe_inv_sq_diff(s[1], s[2], eps);
    M3<double> f = zero3<double>();
    f(0, 1) = f01;
    f(0, 2) = f02;
    f(1, 2) = f12;
    f(1, 0) = -f01;
    f(2, 0) = -f02;
    f(2, 1) = -f12;
    M3<double> Ut = tr(U), Vt = tr(V);
    M3<double> sigma_term = udv(U, aS, V);
    M3<double> om_u = sub(mul(Ut, aU), mul(tr(aU), U));
    M3<double> om_v = sub(mul(Vt, aV), mul(tr(aV), V));
    M3<double> sk_u, sk_v;
    for (int i = 0; i < 9; ++i) {
        sk_u.a[i] = f.a[i] * om_u.a[i];
        sk_v.a[i] = f.a[i] * om_v.a[i];
    }
    M3<double> Sd = zero3<double>();
    Sd(0, 0) = s[0];
    Sd(1, 1) = s[1];
    Sd(2, 2) = s[2];
    M3<double> u_term = mul(mul(U, mul(sk_u, Sd)), Vt);
    M3<double> v_term = mul(mul(U, mul(Sd, sk_v)), Vt);
    return add(add(sigma_term, u_term), v_term);
}

// quad_weights: mpm.wgsl:287-294; mpm_cpu_diff.rs:233-249.
template <typename T>
void quad_w(T fx, T* w) {
    w[0] = T(0.5) * (T(1.5) - fx) * (T(1.5) - fx);
    w[1] = T(0.75) - (fx - T(1)) * (fx - T(1));
    w[2] = T(0.5) * (fx - T(0.5)) * (fx - T(0.5));
}
// quad_weights_dfx: mpm_cpu_diff.rs:251-260.
template <typename T>
void quad_dw(T fx, T* dw) {
    dw[0] = -(T(1.5) - fx);
    dw[1] = T(-2) * (fx - T(1));
    dw[2] = fx - T(0.5);
}

// Lame constants: mpm_cpu_diff.rs:217-222 and mpm.wgsl:542-543.
template <typename T>
void lame(T E, T nu, T& mu, T& lam) {
    mu = E / (T(2) * (T(1) + nu));
    lam = E * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
}

// ---------------------------------------------------------------------------
// Fixed-corotated stress block of p2g / p2g_two_pass / gather_prepare
// (mpm.wgsl:532-599, :779-826; mpm_deterministic_p2g.wgsl:12-63).
// Returns stress = -dt * vol * P(F_clamped) * F_clamped^T.
template <typename T>
M3<T> shader_stress(const M3<T>& F, T mu_h, T lam_h, T dt, T vol) {
    M3<T> U, V;
    V3<T> s;
    svd3_shader(F, U, s, V);
    const T slo = g_mut == MUT_STRESS_CLAMP ? T(0.05) : T(0.1);
    for (int i = 0; i < 3; ++i) s[i] = std::min(std::max(s[i], slo), T(10));
    M3<T> R = mul(U, tr(V));
    T J = s[0] * s[1] * s[2];
    M3<T> Fc = udv(U, s, V);
    V3<T> cd = {{s[1] * s[2], s[0] * s[2], s[0] * s[1]}};
    M3<T> cof = g_mut == MUT_STRESS_COF ? R : udv(U, cd, V);
    M3<T> pk1 = add(scale(T(2) * mu_h, sub(Fc, R)), scale(lam_h * (J - T(1)), cof));
    return scale(-dt * vol, mul(pk1, tr(Fc)));
}

// corotated_pk1: mpm_cpu_diff.rs:279-287.
M3<double> pk1_svk(const M3<double>& U, const V3<double>& s, const M3<double>& V, double mu,
                   double lam) {
    M3<double> R = mul(U, tr(V));
    double trsi = s[0] + s[1] + s[2] - 3.0;
    M3<double> Fr = udv(U, s, V);
    return add(scale(2.0 * mu, sub(Fr, R)), scale(lam * trsi, R));
}

// corotated_pk1_adjoint: mpm_cpu_diff.rs:330-379.
void pk1_svk_adjoint(const M3<double>& U, const V3<double>& s, const M3<double>& V, double mu,
                     double lam, const M3<double>& aP, M3<double>& aU, V3<double>& aS,
                     M3<double>& aV, double& a_mu, double& a_lam) {
    M3<double> Vt = tr(V);
    M3<double> R = mul(U, Vt);
    M3<double> Sd = zero3<double>();
    Sd(0, 0) = s[0];
    Sd(1, 1) = s[1];
    Sd(2, 2) = s[2];
    M3<double> Fr = mul(mul(U, Sd), Vt);
    double trsi = s[0] + s[1] + s[2] - 3.0;
    M3<double> aF = scale(2.0 * mu, aP);
    M3<double> au_f = mul(mul(aF, V), Sd);
    M3<double> av_f = mul(mul(tr(aF), U), Sd);
    M3<double> ms = mul(mul(tr(U), aF), V);
    double coef_r = -2.0 * mu + lam * trsi;
    M3<double> aR = scale(coef_r, aP);
    M3<double> au_r = mul(aR, V);
    M3<double> av_r = mul(tr(aR), U);
    double r_dot = frob(R, aP);
    aU = add(au_f, au_r);
    aV = add(av_f, av_r);
    for (int i = 0; i < 3; ++i) aS[i] = ms(i, i) + lam * r_dot;
    a_mu = 2.0 * frob(sub(Fr, R), aP);
    a_lam = trsi * r_dot;
}

template <typename T>
struct Mat {
    T rheo, dp_alpha, theta_c, theta_s, hardening, vdamp, adamp, bulk;
};

template <typename T>
Mat<T> mat_row(const orc_scene* s, const uint32_t* mat, uint32_t p) {
    uint32_t id = mat ? mat[p] : 0u;
    if (id >= s->n_materials) id = 0;
    const orc_material& m = s->materials[id];
    return Mat<T>{T(m.rheology), T(m.dp_alpha), T(m.theta_c), T(m.theta_s),
                  T(m.hardening), T(m.velocity_damping), T(m.affine_damping), T(m.bulk)};
}
template <typename T>
int rheo_id(const Mat<T>& m) {
    return int(m.rheo + T(0.5));
}

// Plasticity return maps: mpm.wgsl:1355-1479 (sand Drucker-Prager,
// snow envelope, viscous identity) + build-defined fluid.  Returns false on
// a numerical fault.  `s` are the (possibly signed) shader singular values.
template <typename T>
bool plastic(const orc_scene* sc, const Mat<T>& m, T youngs, const V3<T>& s, T jp_in,
             V3<T>& s_out, T& jp_out) {
    int rheo = rheo_id(m);
    V3<T> ns = s;
    T dq = T(0);
    bool jp_replace = false;
    T jp_new = T(0);
    if (rheo == ORC_SNOW) {
        for (int i = 0; i < 3; ++i)
            ns[i] = std::min(std::max(s[i], T(1) - m.theta_c), T(1) + m.theta_s);
        T det_old = std::max(s[0] * s[1] * s[2], T(1.0e-6));
        T det_new = std::max(ns[0] * ns[1] * ns[2], T(1.0e-6));
        T dlog = std::log(det_old) - std::log(det_new);
        jp_replace = true;
        jp_new = jp_in + dlog;
        if (!std::isfinite(jp_new)) return false;
        jp_new = std::min(std::max(jp_new, T(-5)), T(5));
    } else if (rheo == ORC_VISCOUS || rheo == ORC_FLUID) {
        ns = s;
        jp_replace = true;
        jp_new = T(0);
    } else {
        T mu, lam;
        lame<T>(youngs, T(sc->poisson), mu, lam);
        const T d = T(3);
        V3<T> e;
        for (int i = 0; i < 3; ++i) e[i] = std::log(std::max(s[i], T(1.0e-6)));
        T tre = e[0] + e[1] + e[2];
        T mean = tre / d;
        V3<T> dev = {{e[0] - mean, e[1] - mean, e[2] - mean}};
        T dn = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
        T kappa = (d * lam + T(2) * mu) / (T(2) * mu);
        if (tre > T(0)) {
            ns = V3<T>{{T(1), T(1), T(1)}};
            dq = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        } else if (dn == T(0)) {
            ns = s;
        } else {
            T fy = dn + kappa * tre * m.dp_alpha;
            if (fy <= T(0)) {
                ns = s;
            } else {
                T rc = std::max(T(0), -m.dp_alpha * kappa * tre);
                for (int i = 0; i < 3; ++i) ns[i] = std::exp(dev[i] * (rc / dn) + mean);
                dq = fy;
            }
        }
    }
    if (!finite3(ns.x, 3) || !std::isfinite(dq) || !std::isfinite(jp_in + dq)) return false;
    const T shi = g_mut == MUT_PLASTIC_CLAMP ? T(10) : T(20);
    for (int i = 0; i < 3; ++i) s_out[i] = std::min(std::max(ns[i], T(0.05)), shi);
    jp_out = jp_replace ? jp_new : std::min(std::max(jp_in + dq, T(0)), T(5));
    return true;
}

template <typename T>
T hardening(const Mat<T>& m, T jp) {
    int r = rheo_id(m);
    if (r == ORC_SNOW) return std::exp(std::min(std::max(-m.hardening * jp, T(-3)), T(3)));
    if (r == ORC_SAND) return std::exp(std::min(std::max(jp, T(-3)), T(3)));
    return T(1);
}

template <typename T>
struct Grid {
    uint32_t G = 0;
    std::vector<T> m, p, v;  // m[G3], p[3 G3], v[3 G3]
    std::vector<uint8_t> mask;  // g_rigid_mask: cell stamped by a rigid body
    void init(uint32_t g) {
        G = g;
        size_t n = size_t(g) * g * g;
        m.assign(n, T(0));
        p.assign(3 * n, T(0));
        v.assign(3 * n, T(0));
        mask.assign(n, 0);
    }
    void clear() {
        std::fill(m.begin(), m.end(), T(0));
        std::fill(p.begin(), p.end(), T(0));
        std::fill(v.begin(), v.end(), T(0));
        std::fill(mask.begin(), mask.end(), uint8_t(0));
    }
    size_t idx(int ix, int iy, int iz) const {
        return size_t(ix) + size_t(G) * (size_t(iy) + size_t(G) * size_t(iz));
    }
    bool in(int ix, int iy, int iz) const {
        int n = int(G);
        return ix >= 0 && iy >= 0 && iz >= 0 && ix < n && iy < n && iz < n;
    }
};

template <typename T>
struct Particles {
    uint32_t n;
    T *x, *v, *C, *F, *jp;
    const T *mass, *vol, *E;
    const uint32_t* mat;
};

template <typename T>
M3<T> load9(const T* p) {
    M3<T> m;
    for (int i = 0; i < 9; ++i) m.a[i] = p[i];
    return m;
}
template <typename T>
void store9(T* p, const M3<T>& m) {
    for (int i = 0; i < 9; ++i) p[i] = m.a[i];
}

// Per-particle stress in the SHADER model (p2g: mpm.wgsl:764-826).
template <typename T>
M3<T> particle_stress_shader(const orc_scene* s, const Particles<T>& P, uint32_t p, T dt) {
    Mat<T> m = mat_row<T>(s, P.mat, p);
    T E = P.E ? P.E[p] : T(s->youngs);
    T mu, lam;
    lame<T>(E, T(s->poisson), mu, lam);
    T h = hardening(m, P.jp[p]);
    T mu_h = mu * h, lam_h = lam * h;
    if (rheo_id(m) == ORC_FLUID) {
        mu_h = T(0);
        lam_h = m.bulk;
    }
    return shader_stress(load9(P.F + 9 * size_t(p)), mu_h, lam_h, dt, P.vol[p]);
}

// P2G of particle p in the SHADER model: mpm.wgsl:494-628 (math shared
// with p2g_two_pass).
template <typename T>
void p2g_particle_shader(const orc_scene* s, const Particles<T>& P, uint32_t p, T dt, Grid<T>& g) {
    const T dx = T(s->cell_size);
    const T inv_dx = T(1) / dx;
    const T inv_dx2 = inv_dx * inv_dx;
    {
        V3<T> xp, fx;
        int base[3];
        for (int a = 0; a < 3; ++a) {
            xp[a] = (P.x[3 * p + a] - T(s->origin[a])) * inv_dx;
            T bf = std::floor(xp[a] - T(0.5));
            base[a] = int(bf) - s->window[a];
            fx[a] = xp[a] - bf;
        }
        T w[3][3];
        for (int a = 0; a < 3; ++a) quad_w(fx[a], w[a]);
        M3<T> stress = particle_stress_shader(s, P, p, dt);
        M3<T> C = load9(P.C + 9 * size_t(p));
        T mass = P.mass[p];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    T wt = w[0][i] * w[1][j] * w[2][k];
                    int ix = base[0] + i, iy = base[1] + j, iz = base[2] + k;
                    if (!g.in(ix, iy, iz)) continue;
                    V3<T> d = {{(T(i) - fx[0]) * dx, (T(j) - fx[1]) * dx, (T(k) - fx[2]) * dx}};
                    V3<T> aff = mv(C, d);
                    V3<T> sf = mv(stress, d);
                    size_t c = g.idx(ix, iy, iz);
                    g.m[c] += wt * mass;
                    for (int a = 0; a < 3; ++a) {
                        T mom = wt * mass * (P.v[3 * p + a] + aff[a]) + wt * (sf[a] * (T(4) * inv_dx2));
                        g.p[3 * c + a] += mom;
                    }
                }
    }
}

// P2G over all particles in index order.
template <typename T>
void p2g_shader(const orc_scene* s, const Particles<T>& P, T dt, Grid<T>& g) {
    for (uint32_t p = 0; p < P.n; ++p) p2g_particle_shader(s, P, p, dt, g);
}

// grid_update: mpm.wgsl:1010-1058 (no rigid imprint).
template <typename T>
void grid_update_shader(const orc_scene* s, T dt, Grid<T>& g, bool threads = false) {
    const int n = int(g.G);
    const T bf = T(s->boundary_friction);
#pragma omp parallel for schedule(static) if (threads)
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                size_t c = g.idx(ix, iy, iz);
                T m = g.m[c];
                if (m <= T(1.0e-9) || !(m == m)) {
                    g.v[3 * c] = g.v[3 * c + 1] = g.v[3 * c + 2] = T(0);
                    continue;
                }
                T v[3] = {g.p[3 * c] / m, g.p[3 * c + 1] / m, g.p[3 * c + 2] / m};
                if (g.mask[c]) {  // rigid-imprinted: no integration (mpm.wgsl:1034-1038)
                    for (int a = 0; a < 3; ++a) g.v[3 * c + a] = v[a];
                    continue;
                }
                v[1] = v[1] + dt * T(s->gravity_y);
                const int pad = g_mut == MUT_PAD_ONE ? 1 : 2;
                // a wall: the normal component may only point inwards, both
                // tangential components are scaled by the friction
                auto wall = [&](int a, bool low) {
                    v[a] = low ? std::max(v[a], T(0)) : std::min(v[a], T(0));
                    if (g_mut == MUT_FRICTION_NORMAL) v[a] *= bf;
                    v[(a + 1) % 3] *= bf;
                    if (g_mut != MUT_FRICTION_ONE_TAN) v[(a + 2) % 3] *= bf;
                };
                if (ix < pad) wall(0, true);
                if (ix >= n - pad) wall(0, false);
                if (iy < pad) wall(1, true);
                if (iy >= n - pad) wall(1, false);
                if (iz < pad) wall(2, true);
                if (iz >= n - pad) wall(2, false);
                for (int a = 0; a < 3; ++a) g.v[3 * c + a] = v[a];
            }
}

// mpm_implicit_apply + mpm_implicit_swap: mpm.wgsl:1060-1153, host loop
// mpm.rs:2515-2530 (no rigid mask).  Each Jacobi sweep blends every cell with
// m > 1e-9 towards the mean velocity of its face neighbours with m > 1e-9:
// v' = (m v + a v_avg) / (m + a), a = dt^2 m (E / 1500) / dx^2; cells without
// such neighbours keep v; empty cells are copied.  Sweeps read the previous
// sweep's velocities (ping-pong).
template <typename T>
void implicit_sweeps(const orc_scene* s, T dt, Grid<T>& g) {
    const int n = int(g.G);
    const T inv_dx = T(1) / T(s->cell_size);
    const T c2 = T(s->youngs) / T(1500);
    std::vector<T> nv(g.v.size());
    for (uint32_t it = 0; it < s->jacobi_iters; ++it) {
        for (int iz = 0; iz < n; ++iz)
            for (int iy = 0; iy < n; ++iy)
                for (int ix = 0; ix < n; ++ix) {
                    const size_t c = g.idx(ix, iy, iz);
                    const T m = g.m[c];
                    if (m <= T(1.0e-9) || !(m == m) || g.mask[c]) {
                        for (int a = 0; a < 3; ++a) nv[3 * c + a] = g.v[3 * c + a];
                        continue;
                    }
                    const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                    T avg[3] = {T(0), T(0), T(0)};
                    T cnt = T(0);
                    for (int k = 0; k < 6; ++k) {
                        const int jx = ix + off[k][0], jy = iy + off[k][1], jz = iz + off[k][2];
                        if (!g.in(jx, jy, jz)) continue;
                        const size_t d = g.idx(jx, jy, jz);
                        const T nm = g.m[d];
                        if (nm <= T(1.0e-9) || !(nm == nm)) continue;
                        for (int a = 0; a < 3; ++a) avg[a] += g.v[3 * d + a];
                        cnt += T(1);
                    }
                    for (int a = 0; a < 3; ++a) avg[a] = cnt > T(0) ? avg[a] / cnt : g.v[3 * c + a];
                    const T alpha = dt * dt * m * c2 * inv_dx * inv_dx;
                    const T den = m + alpha;
                    T out[3];
                    bool fin = true;
                    for (int a = 0; a < 3; ++a) {
                        out[a] = (m * g.v[3 * c + a] + alpha * avg[a]) / den;
                        fin = fin && out[a] == out[a];
                    }
                    for (int a = 0; a < 3; ++a) nv[3 * c + a] = fin ? out[a] : g.v[3 * c + a];
                }
        g.v.swap(nv);
    }
}

// g2p: mpm.wgsl:1241-1353.  Returns false (and leaves the particle
// untouched) on a numerical fault, mirroring the shader's early return.
template <typename T>
bool g2p_shader(const orc_scene* s, Particles<T>& P, uint32_t p, T dt, const Grid<T>& g) {
    T* x = P.x + 3 * size_t(p);
    T* vel = P.v + 3 * size_t(p);
    T* Cp = P.C + 9 * size_t(p);
    T* Fp = P.F + 9 * size_t(p);
    if (!finite3(x, 3) || !finite3(vel, 3) || !finite3(Cp, 9) || !finite3(Fp, 9) ||
        !std::isfinite(P.jp[p]))
        return false;
    const T dx = T(s->cell_size);
    const T inv_dx = T(1) / dx;
    V3<T> xp, fx;
    int base[3];
    for (int a = 0; a < 3; ++a) xp[a] = (x[a] - T(s->origin[a])) * inv_dx;
    if (!finite3(xp.x, 3)) return false;
    for (int a = 0; a < 3; ++a) {
        T bf = std::floor(xp[a] - T(0.5));
        base[a] = int(bf) - s->window[a];
        fx[a] = xp[a] - bf;
    }
    T w[3][3];
    for (int a = 0; a < 3; ++a) quad_w(fx[a], w[a]);
    V3<T> vn = {{T(0), T(0), T(0)}};
    M3<T> Cn = zero3<T>();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                T wt = w[0][i] * w[1][j] * w[2][k];
                int ix = base[0] + i, iy = base[1] + j, iz = base[2] + k;
                if (!g.in(ix, iy, iz)) continue;
                size_t c = g.idx(ix, iy, iz);
                const T* gv = &g.v[3 * c];
                T d[3] = {(T(i) - fx[0]) * dx, (T(j) - fx[1]) * dx, (T(k) - fx[2]) * dx};
                T sc = T(4) * wt * inv_dx * inv_dx;
                for (int a = 0; a < 3; ++a) {
                    vn[a] = vn[a] + wt * gv[a];
                    for (int b = 0; b < 3; ++b) Cn(a, b) = Cn(a, b) + sc * gv[a] * d[b];
                }
            }
    Mat<T> m = mat_row<T>(s, P.mat, p);
    if (rheo_id(m) == ORC_VISCOUS) {
        T vd = std::exp(-std::max(m.vdamp, T(0)) * dt);
        T ad = std::exp(-std::max(m.adamp, T(0)) * dt);
        for (int a = 0; a < 3; ++a) vn[a] = vn[a] * vd;
        Cn = scale(ad, Cn);
    }
    T np[3];
    for (int a = 0; a < 3; ++a) np[a] = x[a] + dt * vn[a];
    if (!finite3(np, 3) || !finite3(vn.x, 3) || !finite3(Cn.a, 9)) return false;
    for (int a = 0; a < 3; ++a) {
        T lo = T(s->origin[a]) + T(s->window[a] + (g_mut == MUT_CLAMP_LO ? 1 : 2)) * dx;
        T hi = T(s->origin[a]) + T(s->window[a] + int(s->grid_size) - 2 - 1) * dx;
        if ((np[a] < lo || np[a] > hi) && g_mut != MUT_CLAMP_KEEP_V) vn[a] = T(0);
        np[a] = std::min(std::max(np[a], lo), hi);
    }
    M3<T> A = add(eye3<T>(), scale(dt, Cn));
    M3<T> Fn = mul(A, load9(Fp));
    if (!finite3(Fn.a, 9)) return false;
    for (int a = 0; a < 3; ++a) {
        x[a] = np[a];
        vel[a] = vn[a];
    }
    store9(Cp, Cn);
    store9(Fp, Fn);
    return true;
}

// plasticity: mpm.wgsl:1355-1479 (+ build-defined fluid reset).
template <typename T>
bool plasticity_shader(const orc_scene* s, Particles<T>& P, uint32_t p) {
    Mat<T> m = mat_row<T>(s, P.mat, p);
    T* Fp = P.F + 9 * size_t(p);
    if (!finite3(Fp, 9) || !std::isfinite(P.jp[p])) return false;
    M3<T> F = load9(Fp);
    if (rheo_id(m) == ORC_FLUID) {
        double J = det3(M3<double>{{double(F.a[0]), double(F.a[1]), double(F.a[2]), double(F.a[3]),
                                    double(F.a[4]), double(F.a[5]), double(F.a[6]), double(F.a[7]),
                                    double(F.a[8])}});
        double sc = g_mut == MUT_FLUID_SQRT ? std::sqrt(J) : std::cbrt(J);
        sc = std::min(std::max(sc, 0.05), 20.0);
        if (!std::isfinite(sc)) return false;
        M3<T> Fn = scale(T(sc), eye3<T>());
        store9(Fp, Fn);
        P.jp[p] = T(0);
        return true;
    }
    M3<T> U, V;
    V3<T> sv;
    svd3_shader(F, U, sv, V);
    if (!finite3(sv.x, 3) || !finite3(U.a, 9) || !finite3(V.a, 9)) return false;
    T E = P.E ? P.E[p] : T(s->youngs);
    V3<T> so;
    T jp_out;
    if (!plastic(s, m, E, sv, P.jp[p], so, jp_out)) return false;
    M3<T> Fn = udv(U, so, V);
    if (!finite3(Fn.a, 9)) return false;
    store9(Fp, Fn);
    P.jp[p] = jp_out;
    return true;
}

// ---------------------------------------------------------------------------
// Rigid bodies on the grid.  A body is five rows of four floats: center and
// bounding radius, linear velocity, angular velocity, unit quaternion (xyzw),
// shape dims and kind (0 legacy sphere, 1 sphere, 2 box, 3 cylinder along
// local y).
struct Rigids {
    const float* rows = nullptr;
    uint32_t count = 0;
    double* wrench = nullptr;  // 6 per body: force, torque (last substep)
    const float* b(uint32_t r) const { return rows + 20 * size_t(r); }
};

// prescribed_solid_contains (mpm.wgsl:927-935) and the legacy distance test
// (:956-960): the arm is rotated into the body frame by the conjugate of q,
// local = arm + w t + (-q) x t with t = 2 (-q) x arm.
template <typename T>
bool rigid_contains(const float* b, const T arm[3]) {
    const T kind = T(b[19]);
    auto len3 = [](T x, T y, T z) { return std::sqrt(x * x + y * y + z * z); };
    if (kind == T(0)) return len3(arm[0], arm[1], arm[2]) < T(b[3]);
    if (kind == T(1)) return len3(arm[0], arm[1], arm[2]) < T(b[16]);
    const T iq[3] = {-T(b[12]), -T(b[13]), -T(b[14])}, w = T(b[15]);
    T t[3] = {T(2) * (iq[1] * arm[2] - iq[2] * arm[1]), T(2) * (iq[2] * arm[0] - iq[0] * arm[2]),
              T(2) * (iq[0] * arm[1] - iq[1] * arm[0])};
    const T u[3] = {iq[1] * t[2] - iq[2] * t[1], iq[2] * t[0] - iq[0] * t[2], iq[0] * t[1] - iq[1] * t[0]};
    T local[3];
    for (int a = 0; a < 3; ++a) local[a] = arm[a] + w * t[a] + u[a];
    if (kind == T(2))
        return std::fabs(local[0]) < T(b[16]) && std::fabs(local[1]) < T(b[17]) && std::fabs(local[2]) < T(b[18]);
    return std::sqrt(local[0] * local[0] + local[2] * local[2]) < T(b[16]) && std::fabs(local[1]) < T(b[17]);
}

template <typename T>
void node_pos(const orc_scene* s, int ix, int iy, int iz, T out[3]) {
    const int i[3] = {ix, iy, iz};
    for (int a = 0; a < 3; ++a) out[a] = T(s->origin[a]) + T(i[a] + s->window[a]) * T(s->cell_size);
}

// rigid_imprint (mpm.wgsl:937-1006) on every cell of the dense grid: the
// first body containing the node stamps it (legacy spheres with mass 1e6,
// solids keep the sand mass and skip empty cells); the sand's impulse
// m_old v - p_old (solids: minus m_old g dt in y) reacts on the body as
// -impulse / dt (rigid_cell_imprint_wrench, mpm_oracle.rs:54-80).
template <typename T>
void rigid_imprint_grid(const orc_scene* s, const Rigids& R, T dt, Grid<T>& g) {
    const int n = int(g.G);
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                const size_t c = g.idx(ix, iy, iz);
                T node[3];
                node_pos(s, ix, iy, iz, node);
                for (uint32_t r = 0; r < R.count; ++r) {
                    const float* b = R.b(r);
                    const T arm[3] = {node[0] - T(b[0]), node[1] - T(b[1]), node[2] - T(b[2])};
                    if (!rigid_contains(b, arm)) continue;
                    const bool solid = b[19] != 0.0f;
                    T v[3] = {T(b[4]), T(b[5]), T(b[6])};
                    if (solid) {  // rigid_body_velocity_at_cell, mpm_oracle.rs:82-90
                        const T om[3] = {T(b[8]), T(b[9]), T(b[10])};
                        v[0] += om[1] * arm[2] - om[2] * arm[1];
                        v[1] += om[2] * arm[0] - om[0] * arm[2];
                        v[2] += om[0] * arm[1] - om[1] * arm[0];
                    }
                    const T m_old = g.m[c];
                    if (solid && m_old <= T(1.0e-9)) break;
                    const T stamp = solid ? m_old : T(1.0e6);
                    const T p_old[3] = {g.p[3 * c], g.p[3 * c + 1], g.p[3 * c + 2]};
                    g.m[c] = stamp;
                    for (int a = 0; a < 3; ++a) g.p[3 * c + a] = stamp * v[a];
                    g.mask[c] = 1;
                    if (m_old > T(1.0e-9) && R.wrench) {
                        T imp[3];
                        for (int a = 0; a < 3; ++a) imp[a] = m_old * v[a] - p_old[a];
                        if (solid) imp[1] -= m_old * T(s->gravity_y) * dt;
                        const T inv_dt = T(1) / (solid ? dt : std::max(dt, T(1.0e-6)));
                        const T f[3] = {-imp[0] * inv_dt, -imp[1] * inv_dt, -imp[2] * inv_dt};
                        double* w = R.wrench + 6 * size_t(r);
                        for (int a = 0; a < 3; ++a) w[a] += double(f[a]);
                        w[3] += double(arm[1] * f[2] - arm[2] * f[1]);
                        w[4] += double(arm[2] * f[0] - arm[0] * f[2]);
                        w[5] += double(arm[0] * f[1] - arm[1] * f[0]);
                    }
                    break;
                }
            }
}

// mpm_rigid_force_integrate (mpm.wgsl:1174-1226), legacy spheres only: every
// unstamped cell with sand mass within radius + dx of a sphere's center
// pushes on it with -(m w / dt)(v_sphere - v_cell), w = 1 inside, tapering
// linearly to 0 one cell outside.
template <typename T>
void rigid_shell_grid(const orc_scene* s, const Rigids& R, T dt, const Grid<T>& g) {
    const int n = int(g.G);
    const T dx = T(s->cell_size);
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                const size_t c = g.idx(ix, iy, iz);
                const T m = g.m[c];
                if (m <= T(1.0e-9) || !(m == m) || g.mask[c]) continue;
                T node[3];
                node_pos(s, ix, iy, iz, node);
                for (uint32_t r = 0; r < R.count; ++r) {
                    const float* b = R.b(r);
                    if (b[19] != 0.0f) continue;
                    const T arm[3] = {node[0] - T(b[0]), node[1] - T(b[1]), node[2] - T(b[2])};
                    const T d = std::sqrt(arm[0] * arm[0] + arm[1] * arm[1] + arm[2] * arm[2]);
                    const T radius = T(b[3]);
                    if (d > radius + dx) continue;
                    T w = T(1);
                    if (d > radius) w = std::max(T(0), T(1) - (d - radius) / dx);
                    const T k = -(m * w / std::max(dt, T(1.0e-6)));
                    const T f[3] = {k * (T(b[4]) - g.v[3 * c]), k * (T(b[5]) - g.v[3 * c + 1]),
                                    k * (T(b[6]) - g.v[3 * c + 2])};
                    double* wr = R.wrench + 6 * size_t(r);
                    for (int a = 0; a < 3; ++a) wr[a] += double(f[a]);
                    wr[3] += double(arm[1] * f[2] - arm[2] * f[1]);
                    wr[4] += double(arm[2] * f[0] - arm[0] * f[2]);
                    wr[5] += double(arm[0] * f[1] - arm[1] * f[0]);
                }
            }
}

template <typename T>
int step_shader(const orc_scene* s, Particles<T>& P, T dt, uint32_t substeps, uint32_t* status,
                T* out_m, T* out_p, const Rigids* R = nullptr) {
    Grid<T> g;
    g.init(s->grid_size);
    uint32_t fault = status ? status[0] : 0u;
    uint32_t gathers = status ? status[1] : 0u;
    const bool rigid = R && R->count > 0;
    for (uint32_t st = 0; st < substeps; ++st) {
        g.clear();
        if (fault == 0) p2g_shader(s, P, dt, g);
        // the wrench buffer is reset every substep: the last one remains
        if (rigid && R->wrench) std::fill(R->wrench, R->wrench + 6 * size_t(R->count), 0.0);
        if (rigid) rigid_imprint_grid(s, *R, dt, g);
        if (out_m && st + 1 == substeps) std::copy(g.m.begin(), g.m.end(), out_m);
        if (out_p && st + 1 == substeps) std::copy(g.p.begin(), g.p.end(), out_p);
        grid_update_shader(s, dt, g);
        implicit_sweeps(s, dt, g);
        if (rigid && R->wrench) rigid_shell_grid(s, *R, dt, g);
        if (fault == 0) {
            if (P.n > 0) gathers++;
            uint32_t f = 0;
            for (uint32_t p = 0; p < P.n; ++p)
                if (!g2p_shader(s, P, p, dt, g)) f |= ORC_FAULT_NUMERICAL;
            fault |= f;
        }
        if (fault == 0) {
            uint32_t f = 0;
            for (uint32_t p = 0; p < P.n; ++p)
                if (!plasticity_shader(s, P, p)) f |= ORC_FAULT_NUMERICAL;
            fault |= f;
        }
    }
    if (status) {
        status[0] = fault;
        status[1] = gathers;
    }
    return 0;
}

// The SHADER step on all host threads: a timing port for the CPU baseline,
// not an oracle (its sums run in another order than step_shader's).  The P2G
// scatter runs in two phases over z-slabs of SLAB cells: a particle's stencil
// spans its base cell and the next two, so with SLAB >= 2 two slabs of the
// same parity never touch the same node.  The grid update and the particle
// passes split across threads as they are.  No rigid bodies.
template <typename T>
int step_shader_threads(const orc_scene* s, Particles<T>& P, T dt, uint32_t substeps, uint32_t* status) {
    constexpr int SLAB = 4;
    Grid<T> g;
    g.init(s->grid_size);
    const int G = int(s->grid_size);
    const int nslab = (G + SLAB - 1) / SLAB;
    const T inv_dx = T(1) / T(s->cell_size);
    std::vector<int> slab(P.n);
    std::vector<uint32_t> order(P.n), start(size_t(nslab) + 1);
    uint32_t fault = status ? status[0] : 0u;
    uint32_t gathers = status ? status[1] : 0u;
    const long long ncell = (long long)g.m.size();
    for (uint32_t st = 0; st < substeps; ++st) {
#pragma omp parallel for schedule(static)
        for (long long c = 0; c < ncell; ++c) {
            g.m[c] = T(0);
            for (int a = 0; a < 3; ++a) g.p[3 * c + a] = g.v[3 * c + a] = T(0);
        }
        if (fault == 0) {
#pragma omp parallel for schedule(static)
            for (long long p = 0; p < (long long)P.n; ++p) {
                const T xp = (P.x[3 * p + 2] - T(s->origin[2])) * inv_dx;
                const T bf = std::floor(xp - T(0.5));
                const int bz = std::isfinite(bf) ? int(std::min(std::max(bf, T(-8)), T(1 << 30))) - s->window[2]
                                                 : 0;
                slab[p] = std::min(std::max(bz, 0), G - 1) / SLAB;
            }
            std::fill(start.begin(), start.end(), 0u);
            for (uint32_t p = 0; p < P.n; ++p) start[slab[p] + 1]++;
            for (int k = 0; k < nslab; ++k) start[k + 1] += start[k];
            std::vector<uint32_t> fill(start.begin(), start.end() - 1);
            for (uint32_t p = 0; p < P.n; ++p) order[fill[slab[p]]++] = p;
            for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
                for (int k = phase; k < nslab; k += 2)
                    for (uint32_t q = start[k]; q < start[k + 1]; ++q) p2g_particle_shader(s, P, order[q], dt, g);
            }
        }
        grid_update_shader(s, dt, g, true);
        implicit_sweeps(s, dt, g);
        if (fault == 0) {
            if (P.n > 0) gathers++;
            uint32_t f = 0;
#pragma omp parallel for schedule(static) reduction(| : f)
            for (long long p = 0; p < (long long)P.n; ++p)
                if (!g2p_shader(s, P, uint32_t(p), dt, g)) f |= ORC_FAULT_NUMERICAL;
            fault |= f;
        }
        if (fault == 0) {
            uint32_t f = 0;
#pragma omp parallel for schedule(static) reduction(| : f)
            for (long long p = 0; p < (long long)P.n; ++p)
                if (!plasticity_shader(s, P, uint32_t(p))) f |= ORC_FAULT_NUMERICAL;
            fault |= f;
        }
    }
    if (status) {
        status[0] = fault;
        status[1] = gathers;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// CPUDIFF model: mpm_cpu_diff.rs step_once_record (:421-588) without tape.
template <typename T>
struct DiffScratch {
    Grid<double> g;
    std::vector<double> mom_pre;
};

struct DiffCtx {
    const orc_scene* s;
    uint32_t n;
    const double *mass, *vol, *E;
};

void diff_pk1_stress(const DiffCtx& c, uint32_t p, const M3<double>& F, double dt, M3<double>& U,
                     V3<double>& sig, M3<double>& V, M3<double>& pk1, M3<double>& stress,
                     double& mu, double& lam) {
    double Ep = c.E ? c.E[p] : double(c.s->youngs);
    lame<double>(Ep, double(c.s->poisson), mu, lam);
    svd3_cpu(F, U, sig, V);
    pk1 = pk1_svk(U, sig, V, mu, lam);
    stress = scale(-dt * c.vol[p], mul(pk1, tr(F)));
}

struct Stencil {
    int base[3];
    double fx[3];
    double w[3][3];
    double dw[3][3];
};

Stencil stencil_diff(const orc_scene* s, const double* x) {
    Stencil st;
    double inv_dx = 1.0 / double(s->cell_size);
    for (int a = 0; a < 3; ++a) {
        double xp = (x[a] - double(s->origin[a])) * inv_dx;
        double bf = std::floor(xp - 0.5);
        st.base[a] = int(bf);
        st.fx[a] = xp - bf;
        quad_w(st.fx[a], st.w[a]);
        // adjoint recomputes fx as xp - base (mpm_adjoint.rs:257-262); same value.
        quad_dw(st.fx[a], st.dw[a]);
    }
    return st;
}

// P2G + grid update of the CPUDIFF model (mpm_cpu_diff.rs:440-542).
void diff_p2g_grid(const DiffCtx& c, const double* x, const double* v, const double* C,
                   const double* F, double dt, Grid<double>& g, std::vector<double>& mom_pre) {
    const orc_scene* s = c.s;
    const double dx = s->cell_size, inv_dx = 1.0 / dx;
    g.clear();
    for (uint32_t p = 0; p < c.n; ++p) {
        Stencil st = stencil_diff(s, x + 3 * size_t(p));
        M3<double> U, V, pk1, stress;
        V3<double> sig;
        double mu, lam;
        M3<double> Fp = load9(F + 9 * size_t(p));
        diff_pk1_stress(c, p, Fp, dt, U, sig, V, pk1, stress, mu, lam);
        M3<double> Cp = load9(C + 9 * size_t(p));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    double w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                    int ix = st.base[0] + i, iy = st.base[1] + j, iz = st.base[2] + k;
                    if (!g.in(ix, iy, iz)) continue;
                    size_t ci = g.idx(ix, iy, iz);
                    V3<double> d = {{double(s->origin[0]) + ix * dx - x[3 * p],
                                     double(s->origin[1]) + iy * dx - x[3 * p + 1],
                                     double(s->origin[2]) + iz * dx - x[3 * p + 2]}};
                    V3<double> aff = mv(Cp, d);
                    V3<double> sf = mv(stress, d);
                    for (int a = 0; a < 3; ++a) {
                        double mom = w * c.mass[p] * (v[3 * p + a] + aff[a]) +
                                     w * (sf[a] * (4.0 * inv_dx * inv_dx));
                        g.p[3 * ci + a] += mom;
                    }
                    g.m[ci] += w * c.mass[p];
                }
    }
    mom_pre = g.p;
    const int n = int(g.G);
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                size_t ci = g.idx(ix, iy, iz);
                double m = g.m[ci];
                if (m <= 0.0) continue;
                double vv[3] = {g.p[3 * ci] / m, g.p[3 * ci + 1] / m, g.p[3 * ci + 2] / m};
                vv[1] += double(s->gravity_y) * dt;
                double node_y = double(s->origin[1]) + double(iy) * dx;
                if (node_y < double(s->floor_y)) {
                    if (vv[1] < 0.0) vv[1] = 0.0;
                    double damp = 1.0 - double(s->floor_friction);
                    vv[0] *= damp;
                    vv[2] *= damp;
                }
                for (int a = 0; a < 3; ++a) g.v[3 * ci + a] = vv[a];
            }
}

// G2P gather of the CPUDIFF model (mpm_cpu_diff.rs:544-580).
void diff_gather(const orc_scene* s, const Grid<double>& g, const double* x, const Stencil& st,
                 V3<double>& vn, M3<double>& Cn) {
    const double dx = s->cell_size, inv_dx = 1.0 / dx;
    vn = V3<double>{{0, 0, 0}};
    Cn = zero3<double>();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                double w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                int ix = st.base[0] + i, iy = st.base[1] + j, iz = st.base[2] + k;
                if (!g.in(ix, iy, iz)) continue;
                size_t ci = g.idx(ix, iy, iz);
                const double* vnode = &g.v[3 * ci];
                double d[3] = {double(s->origin[0]) + ix * dx - x[0],
                               double(s->origin[1]) + iy * dx - x[1],
                               double(s->origin[2]) + iz * dx - x[2]};
                for (int a = 0; a < 3; ++a) {
                    vn[a] += w * vnode[a];
                    for (int b = 0; b < 3; ++b) Cn(a, b) += vnode[a] * d[b] * (4.0 * w * inv_dx * inv_dx);
                }
            }
}

int diff_forward(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                 const double* mass, const double* vol, const double* E, double dt,
                 uint32_t steps, double* tape) {
    DiffCtx c{s, n, mass, vol, E};
    Grid<double> g;
    g.init(s->grid_size);
    std::vector<double> mom_pre;
    const size_t per = size_t(n) * 24;
    for (uint32_t st = 0; st < steps; ++st) {
        if (tape) {
            double* t = tape + per * st;
            std::copy(x, x + 3 * size_t(n), t);
            std::copy(v, v + 3 * size_t(n), t + 3 * size_t(n));
            std::copy(C, C + 9 * size_t(n), t + 6 * size_t(n));
            std::copy(F, F + 9 * size_t(n), t + 15 * size_t(n));
        }
        diff_p2g_grid(c, x, v, C, F, dt, g, mom_pre);
        for (uint32_t p = 0; p < n; ++p) {
            Stencil stn = stencil_diff(s, x + 3 * size_t(p));
            V3<double> vn;
            M3<double> Cn;
            diff_gather(s, g, x + 3 * size_t(p), stn, vn, Cn);
            for (int a = 0; a < 3; ++a) {
                v[3 * p + a] = vn[a];
                x[3 * p + a] = x[3 * p + a] + vn[a] * dt;
            }
            store9(C + 9 * size_t(p), Cn);
            M3<double> Fn = mul(add(eye3<double>(), scale(dt, Cn)), load9(F + 9 * size_t(p)));
            store9(F + 9 * size_t(p), Fn);
        }
    }
    return 0;
}

// backward_step: mpm_adjoint.rs:142-496, with the grid tape recomputed from
// the pre-step particle state (the reference stores it; values are equal).
void diff_backward_step(const DiffCtx& c, double dt, const double* x, const double* v,
                        const double* C, const double* F, double* gx, double* gv, double* gC,
                        double* gF, double* gE, double* g_mu_lam) {
    const orc_scene* s = c.s;
    const uint32_t n = c.n;
    const double dx = s->cell_size, inv_dx = 1.0 / dx, i4 = 4.0 * inv_dx * inv_dx;
    Grid<double> g;
    g.init(s->grid_size);
    std::vector<double> mom_pre;
    diff_p2g_grid(c, x, v, C, F, dt, g, mom_pre);
    const size_t gt = g.m.size();
    std::vector<double> grad_gvel(3 * gt, 0.0), grad_gmom(3 * gt, 0.0), grad_gmass(gt, 0.0);
    std::vector<V3<double>> gvp(n), gpos_g2p(n);
    std::vector<M3<double>> gcp(n), gf_pre(n);
    std::vector<Stencil> sts(n);
    for (uint32_t p = 0; p < n; ++p) {
        sts[p] = stencil_diff(s, x + 3 * size_t(p));
        for (int a = 0; a < 3; ++a) gvp[p][a] = gx[3 * p + a] * dt + gv[3 * p + a];
        M3<double> Fp = load9(F + 9 * size_t(p));
        gcp[p] = add(load9(gC + 9 * size_t(p)), scale(dt, mul(load9(gF + 9 * size_t(p)), tr(Fp))));
        V3<double> vn;
        M3<double> Cn;
        diff_gather(s, g, x + 3 * size_t(p), sts[p], vn, Cn);
        gf_pre[p] = mul(tr(add(eye3<double>(), scale(dt, Cn))), load9(gF + 9 * size_t(p)));
    }
    // G2P reverse (mpm_adjoint.rs:249-305).
    for (uint32_t p = 0; p < n; ++p) {
        const Stencil& st = sts[p];
        V3<double> acc = {{0, 0, 0}};
        const double* xp = x + 3 * size_t(p);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    double w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                    int ix = st.base[0] + i, iy = st.base[1] + j, iz = st.base[2] + k;
                    if (!g.in(ix, iy, iz)) continue;
                    size_t ci = g.idx(ix, iy, iz);
                    V3<double> d = {{double(s->origin[0]) + ix * dx - xp[0],
                                     double(s->origin[1]) + iy * dx - xp[1],
                                     double(s->origin[2]) + iz * dx - xp[2]}};
                    V3<double> vnode = {{g.v[3 * ci], g.v[3 * ci + 1], g.v[3 * ci + 2]}};
                    V3<double> gcd = mv(gcp[p], d);
                    for (int a = 0; a < 3; ++a) grad_gvel[3 * ci + a] += w * gvp[p][a];
                    for (int a = 0; a < 3; ++a) grad_gvel[3 * ci + a] += (4.0 * w * inv_dx * inv_dx) * gcd[a];
                    V3<double> gctv = mv(tr(gcp[p]), vnode);
                    for (int a = 0; a < 3; ++a) acc[a] -= (4.0 * w * inv_dx * inv_dx) * gctv[a];
                    double dwv[3] = {st.dw[0][i] * st.w[1][j] * st.w[2][k] * inv_dx,
                                     st.w[0][i] * st.dw[1][j] * st.w[2][k] * inv_dx,
                                     st.w[0][i] * st.w[1][j] * st.dw[2][k] * inv_dx};
                    double gvdot = vnode[0] * gvp[p][0] + vnode[1] * gvp[p][1] + vnode[2] * gvp[p][2];
                    for (int a = 0; a < 3; ++a) acc[a] += dwv[a] * gvdot;
                    double od = vnode[0] * gcd[0] + vnode[1] * gcd[1] + vnode[2] * gcd[2];
                    for (int a = 0; a < 3; ++a) acc[a] += dwv[a] * (4.0 * inv_dx * inv_dx * od);
                }
        gpos_g2p[p] = acc;
    }
    // Grid reverse (mpm_adjoint.rs:324-363).
    const int G = int(g.G);
    const double damp = 1.0 - double(s->floor_friction);
    for (int iz = 0; iz < G; ++iz)
        for (int iy = 0; iy < G; ++iy)
            for (int ix = 0; ix < G; ++ix) {
                size_t ci = g.idx(ix, iy, iz);
                double m = g.m[ci];
                if (m <= 0.0) continue;
                double node_y = double(s->origin[1]) + double(iy) * dx;
                double gvv[3] = {grad_gvel[3 * ci], grad_gvel[3 * ci + 1], grad_gvel[3 * ci + 2]};
                if (node_y < double(s->floor_y)) {
                    double vpre_y = mom_pre[3 * ci + 1] / m;
                    vpre_y += double(s->gravity_y) * dt;
                    if (vpre_y < 0.0) gvv[1] = 0.0;
                    gvv[0] *= damp;
                    gvv[2] *= damp;
                }
                for (int a = 0; a < 3; ++a) grad_gmom[3 * ci + a] = gvv[a] / m;
                double dot = mom_pre[3 * ci] * gvv[0] + mom_pre[3 * ci + 1] * gvv[1] +
                             mom_pre[3 * ci + 2] * gvv[2];
                grad_gmass[ci] = -dot / (m * m);
            }
    (void)i4;
    // P2G reverse (mpm_adjoint.rs:389-487).
    double gmu_acc = 0.0, glam_acc = 0.0;
    for (uint32_t p = 0; p < n; ++p) {
        const Stencil& st = sts[p];
        const double* xp = x + 3 * size_t(p);
        const double* vp = v + 3 * size_t(p);
        M3<double> Cp = load9(C + 9 * size_t(p));
        M3<double> Fp = load9(F + 9 * size_t(p));
        M3<double> U, V, pk1, stress;
        V3<double> sig;
        double mu, lam;
        diff_pk1_stress(c, p, Fp, dt, U, sig, V, pk1, stress, mu, lam);
        double m = c.mass[p];
        V3<double> gpos, gvel = {{0, 0, 0}};
        for (int a = 0; a < 3; ++a) gpos[a] = gx[3 * p + a] + gpos_g2p[p][a];
        M3<double> gc = zero3<double>(), gstress = zero3<double>();
        M3<double> gf = gf_pre[p];
        M3<double> Ct = tr(Cp), St = tr(stress);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    double w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                    int ix = st.base[0] + i, iy = st.base[1] + j, iz = st.base[2] + k;
                    if (!g.in(ix, iy, iz)) continue;
                    size_t ci = g.idx(ix, iy, iz);
                    V3<double> d = {{double(s->origin[0]) + ix * dx - xp[0],
                                     double(s->origin[1]) + iy * dx - xp[1],
                                     double(s->origin[2]) + iz * dx - xp[2]}};
                    V3<double> gm = {{grad_gmom[3 * ci], grad_gmom[3 * ci + 1], grad_gmom[3 * ci + 2]}};
                    for (int a = 0; a < 3; ++a) gvel[a] += w * m * gm[a];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            gc(a, b) += w * m * gm[a] * d[b];
                            gstress(a, b) += (4.0 * w * inv_dx * inv_dx) * gm[a] * d[b];
                        }
                    V3<double> ctg = mv(Ct, gm), stg = mv(St, gm);
                    for (int a = 0; a < 3; ++a) gpos[a] -= w * m * ctg[a];
                    for (int a = 0; a < 3; ++a) gpos[a] -= (4.0 * w * inv_dx * inv_dx) * stg[a];
                    double dwv[3] = {st.dw[0][i] * st.w[1][j] * st.w[2][k] * inv_dx,
                                     st.w[0][i] * st.dw[1][j] * st.w[2][k] * inv_dx,
                                     st.w[0][i] * st.w[1][j] * st.dw[2][k] * inv_dx};
                    V3<double> cd = mv(Cp, d), sd = mv(stress, d);
                    double ddot = 0.0;
                    for (int a = 0; a < 3; ++a)
                        ddot += (m * (vp[a] + cd[a]) + (4.0 * inv_dx * inv_dx) * sd[a]) * gm[a];
                    for (int a = 0; a < 3; ++a) gpos[a] += dwv[a] * ddot;
                    for (int a = 0; a < 3; ++a) gpos[a] += dwv[a] * (m * grad_gmass[ci]);
                }
        double coef = -dt * c.vol[p];
        M3<double> gpk1 = scale(coef, mul(gstress, Fp));
        gf = add(gf, scale(coef, mul(tr(gstress), pk1)));
        M3<double> aU, aV;
        V3<double> aS;
        double amu, alam;
        pk1_svk_adjoint(U, sig, V, mu, lam, gpk1, aU, aS, aV, amu, alam);
        gf = add(gf, adj_svd3(U, sig, V, aU, aS, aV));
        gmu_acc += amu;
        glam_acc += alam;
        if (gE) {
            double nu = double(s->poisson);
            gE[p] += amu * (1.0 / (2.0 * (1.0 + nu))) + alam * (nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
        }
        for (int a = 0; a < 3; ++a) {
            gx[3 * p + a] = gpos[a];
            gv[3 * p + a] = gvel[a];
        }
        store9(gC + 9 * size_t(p), gc);
        store9(gF + 9 * size_t(p), gf);
    }
    if (g_mu_lam) {
        g_mu_lam[0] += gmu_acc;
        g_mu_lam[1] += glam_acc;
    }
}

template <typename T>
int step_dispatch(const orc_scene* s, uint32_t n, T* x, T* v, T* C, T* F, T* jp, const T* mass,
                  const T* vol, const uint32_t* mat, const T* E, T dt, uint32_t substeps,
                  uint32_t* status, T* gm, T* gp, const Rigids* R = nullptr) {
    if (!s || s->grid_size == 0) return -1;
    Particles<T> P{n, x, v, C, F, jp, mass, vol, E, mat};
    return step_shader(s, P, dt, substeps, status, gm, gp, R);
}

}  // namespace

extern "C" {
// implicit_sweeps on a caller-supplied dense grid (m[G^3], v[3 G^3], x
// fastest), for unit tests of the sweep itself.
void orc_implicit_sweeps(const orc_scene* s, double dt, const double* m, double* v) {
    Grid<double> g;
    g.init(s->grid_size);
    std::copy(m, m + g.m.size(), g.m.begin());
    std::copy(v, v + g.v.size(), g.v.begin());
    implicit_sweeps(s, dt, g);
    std::copy(g.v.begin(), g.v.end(), v);
}


// tile_keys: shaders/mpm_particle_order.wgsl:13-25 and the closed-form key()
// of src/mpm_particle_order_tests.rs:34-43: cell = floor((x - o) * inv_dx)
// - window (f32, multiply by the f32 reciprocal like the shader), clamped to
// [0, G-1], /4, linearised x-fastest with width ceil(G/4); padding keys
// 0xffffffff for ids in [n, capacity); ids[i] = i.
void orc_tile_keys(const orc_scene* s, uint32_t n, uint32_t capacity, const float* x,
                   uint32_t* keys, uint32_t* ids) {
    const float inv = 1.0f / s->cell_size;
    const uint32_t width = (s->grid_size + 3u) / 4u;
    for (uint32_t i = 0; i < capacity; ++i) {
        ids[i] = i;
        if (i >= n) {
            keys[i] = 0xffffffffu;
            continue;
        }
        uint32_t t[3];
        for (int a = 0; a < 3; ++a) {
            float c = std::floor((x[3 * size_t(i) + a] - s->origin[a]) * inv);
            int ci = int(c) - s->window[a];
            ci = std::min(std::max(ci, 0), int(s->grid_size) - 1);
            t[a] = uint32_t(ci) / 4u;
        }
        keys[i] = t[0] + width * (t[1] + width * t[2]);
    }
}

// Stable (key, id) sort: the order RadixSortKernel produces
// (molla-compute-primitives/src/radix_sort.rs:1-30; stability asserted by
// src/mpm_particle_order_tests.rs:128-150).
void orc_sort_pairs(uint32_t n, uint32_t* keys, uint32_t* ids) {
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::vector<uint32_t> k2(n), i2(n);
    for (uint32_t j = 0; j < n; ++j) {
        k2[j] = keys[idx[j]];
        i2[j] = ids[idx[j]];
    }
    std::copy(k2.begin(), k2.end(), keys);
    std::copy(i2.begin(), i2.end(), ids);
}

// Stencil base cell: base = floor((x - o) * inv_dx - 0.5) - window
// (mpm.wgsl:755-757), f32.
void orc_base_cells(const orc_scene* s, uint32_t n, const float* x, int32_t* base) {
    const float inv = 1.0f / s->cell_size;
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float xp = (x[3 * size_t(i) + a] - s->origin[a]) * inv;
            base[3 * size_t(i) + a] = int(std::floor(xp - 0.5f)) - s->window[a];
        }
}

void orc_svd3_shader_f64(const double* F, double* U, double* sigma, double* V) {
    M3<double> u, v;
    V3<double> s;
    svd3_shader(load9(F), u, s, v);
    store9(U, u);
    store9(V, v);
    for (int i = 0; i < 3; ++i) sigma[i] = s[i];
}
void orc_svd3_shader_f32(const float* F, float* U, float* sigma, float* V) {
    M3<float> u, v;
    V3<float> s;
    svd3_shader(load9(F), u, s, v);
    store9(U, u);
    store9(V, v);
    for (int i = 0; i < 3; ++i) sigma[i] = s[i];
}
void orc_svd3_cpu(const double* F, double* U, double* sigma, double* V) {
    M3<double> u, v;
    V3<double> s;
    svd3_cpu(load9(F), u, s, v);
    store9(U, u);
    store9(V, v);
    for (int i = 0; i < 3; ++i) sigma[i] = s[i];
}
void orc_adj_svd3(const double* U, const double* sigma, const double* V, const double* adjU,
                  const double* adjS, const double* adjV, double* adjF) {
    V3<double> s = {{sigma[0], sigma[1], sigma[2]}}, as = {{adjS[0], adjS[1], adjS[2]}};
    store9(adjF, adj_svd3(load9(U), s, load9(V), load9(adjU), as, load9(adjV)));
}
void orc_pk1_fixed_corotated(const double* F, double mu, double lam, double* P) {
    // stress block with dt = -1, vol = 1 gives P * Fc^T; recover P directly.
    M3<double> U, V;
    V3<double> s;
    svd3_shader(load9(F), U, s, V);
    for (int i = 0; i < 3; ++i) s[i] = std::min(std::max(s[i], 0.1), 10.0);
    M3<double> R = mul(U, tr(V));
    double J = s[0] * s[1] * s[2];
    M3<double> Fc = udv(U, s, V);
    V3<double> cd = {{s[1] * s[2], s[0] * s[2], s[0] * s[1]}};
    M3<double> cof = udv(U, cd, V);
    store9(P, add(scale(2.0 * mu, sub(Fc, R)), scale(lam * (J - 1.0), cof)));
}
void orc_pk1_corotated_svk(const double* F, double mu, double lam, double* P) {
    M3<double> U, V;
    V3<double> s;
    svd3_cpu(load9(F), U, s, V);
    store9(P, pk1_svk(U, s, V, mu, lam));
}
void orc_pk1_corotated_svk_adjoint(const double* F, double mu, double lam, const double* adjP,
                                   double* adjF, double* adj_mu, double* adj_lam) {
    M3<double> U, V, aU, aV;
    V3<double> s, aS;
    svd3_cpu(load9(F), U, s, V);
    pk1_svk_adjoint(U, s, V, mu, lam, load9(adjP), aU, aS, aV, *adj_mu, *adj_lam);
    store9(adjF, adj_svd3(U, s, V, aU, aS, aV));
}
void orc_quad_weights(double fx, double* w) { quad_w(fx, w); }
void orc_quad_weights_dfx(double fx, double* dw) { quad_dw(fx, dw); }

void orc_plastic_project(const orc_scene* s, const orc_material* m, const double* sigma,
                         double jp_in, double* sigma_out, double* jp_out) {
    Mat<double> mm{m->rheology, m->dp_alpha, m->theta_c, m->theta_s,
                   m->hardening, m->velocity_damping, m->affine_damping, m->bulk};
    V3<double> sv = {{sigma[0], sigma[1], sigma[2]}}, so;
    double jo = jp_in;
    if (!plastic<double>(s, mm, double(s->youngs), sv, jp_in, so, jo)) {
        for (int i = 0; i < 3; ++i) sigma_out[i] = NAN;
        *jp_out = NAN;
        return;
    }
    for (int i = 0; i < 3; ++i) sigma_out[i] = so[i];
    *jp_out = jo;
}

int orc_step_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                 double* jp, const double* mass, const double* vol, const uint32_t* mat,
                 const double* E, double dt, uint32_t substeps, uint32_t* status, double* grid_m,
                 double* grid_p) {
    if (s && s->model == ORC_MODEL_CPUDIFF)
        return diff_forward(s, n, x, v, C, F, mass, vol, E, dt, substeps, nullptr);
    return step_dispatch<double>(s, n, x, v, C, F, jp, mass, vol, mat, E, dt, substeps, status,
                                 grid_m, grid_p);
}
int orc_step_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F,
                 float* jp, const float* mass, const float* vol, const uint32_t* mat,
                 const float* E, float dt, uint32_t substeps, uint32_t* status, float* grid_m,
                 float* grid_p) {
    if (s && s->model == ORC_MODEL_CPUDIFF) return -2;
    return step_dispatch<float>(s, n, x, v, C, F, jp, mass, vol, mat, E, dt, substeps, status,
                                grid_m, grid_p);
}

// The threaded f32 timing port of the SHADER step (step_shader_threads);
// *threads receives the thread count used.
int orc_step_threads_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F, float* jp,
                         const float* mass, const float* vol, const uint32_t* mat, const float* E, float dt,
                         uint32_t substeps, uint32_t* status, uint32_t* threads) {
    if (!s || s->grid_size == 0 || s->model == ORC_MODEL_CPUDIFF) return -2;
    Particles<float> P{n, x, v, C, F, jp, mass, vol, E, mat};
    if (threads) *threads = uint32_t(omp_get_max_threads());
    return step_shader_threads<float>(s, P, dt, substeps, status);
}
// The same threaded step in f64: a fast FP64 reference for long runs of
// large scenes (its sums differ from orc_step_f64's only in order).
void orc_set_mutation(int m) { g_mut = m; }

int orc_step_threads_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F, double* jp,
                         const double* mass, const double* vol, const uint32_t* mat, const double* E, double dt,
                         uint32_t substeps, uint32_t* status, uint32_t* threads) {
    if (!s || s->grid_size == 0 || s->model == ORC_MODEL_CPUDIFF) return -2;
    Particles<double> P{n, x, v, C, F, jp, mass, vol, E, mat};
    if (threads) *threads = uint32_t(omp_get_max_threads());
    return step_shader_threads<double>(s, P, dt, substeps, status);
}

// orc_step_* with rigid bodies (rows: 20 floats per body, see Rigids).
int orc_step_rigid_f64(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                       double* jp, const double* mass, const double* vol, const uint32_t* mat,
                       const double* E, double dt, uint32_t substeps, uint32_t* status, double* grid_m,
                       double* grid_p, const float* rows, uint32_t n_bodies, double* wrench) {
    if (!s || s->model != ORC_MODEL_SHADER) return -2;
    Rigids R{rows, n_bodies, wrench};
    return step_dispatch<double>(s, n, x, v, C, F, jp, mass, vol, mat, E, dt, substeps, status, grid_m, grid_p,
                                 &R);
}
int orc_step_rigid_f32(const orc_scene* s, uint32_t n, float* x, float* v, float* C, float* F,
                       float* jp, const float* mass, const float* vol, const uint32_t* mat,
                       const float* E, float dt, uint32_t substeps, uint32_t* status, float* grid_m,
                       float* grid_p, const float* rows, uint32_t n_bodies, double* wrench) {
    if (!s || s->model != ORC_MODEL_SHADER) return -2;
    Rigids R{rows, n_bodies, wrench};
    return step_dispatch<float>(s, n, x, v, C, F, jp, mass, vol, mat, E, dt, substeps, status, grid_m, grid_p,
                                &R);
}

// The grid half of one substep with rigid bodies on a caller-supplied dense
// P2G grid (m[G^3], p[3 G^3], in/out): imprint, grid update, implicit sweeps,
// shell force.  v: 3 G^3 out; mask: G^3 out; wrench: 6 per body out.
void orc_rigid_grid_f64(const orc_scene* s, double dt, const float* rows, uint32_t n_bodies, double* m,
                        double* p, double* v, uint8_t* mask, double* wrench) {
    Grid<double> g;
    g.init(s->grid_size);
    std::copy(m, m + g.m.size(), g.m.begin());
    std::copy(p, p + g.p.size(), g.p.begin());
    Rigids R{rows, n_bodies, wrench};
    std::fill(wrench, wrench + 6 * size_t(n_bodies), 0.0);
    rigid_imprint_grid(s, R, dt, g);
    std::copy(g.m.begin(), g.m.end(), m);
    std::copy(g.p.begin(), g.p.end(), p);
    grid_update_shader(s, dt, g);
    implicit_sweeps(s, dt, g);
    rigid_shell_grid(s, R, dt, g);
    std::copy(g.v.begin(), g.v.end(), v);
    std::copy(g.mask.begin(), g.mask.end(), mask);
}

int orc_p2g_f64(const orc_scene* s, uint32_t n, const double* x, const double* v, const double* C,
                const double* F, const double* jp, const double* mass, const double* vol,
                const uint32_t* mat, const double* E, double dt, double* grid_m, double* grid_p) {
    if (!s || s->grid_size == 0) return -1;
    if (s->model == ORC_MODEL_CPUDIFF) {
        DiffCtx c{s, n, mass, vol, E};
        Grid<double> g;
        g.init(s->grid_size);
        std::vector<double> mom_pre;
        diff_p2g_grid(c, x, v, C, F, dt, g, mom_pre);
        std::copy(g.m.begin(), g.m.end(), grid_m);
        std::copy(mom_pre.begin(), mom_pre.end(), grid_p);
        return 0;
    }
    Grid<double> g;
    g.init(s->grid_size);
    Particles<double> P{n, const_cast<double*>(x), const_cast<double*>(v), const_cast<double*>(C),
                        const_cast<double*>(F), const_cast<double*>(jp), mass, vol, E, mat};
    p2g_shader(s, P, dt, g);
    std::copy(g.m.begin(), g.m.end(), grid_m);
    std::copy(g.p.begin(), g.p.end(), grid_p);
    return 0;
}

// Deterministic gather-order P2G, f32: mpm_deterministic_p2g.wgsl:65-150
// (gather_grid) / mpm_gather_masks.wgsl:1-105.  For each cell the particles of
// the 27 neighbour tiles are visited in tile order dx + 3 dy + 9 dz
// (d in {-1,0,1}) and, within a tile, in ascending sorted index; the stress
// is the gather_prepare block (:12-63).
int orc_p2g_gather_f32(const orc_scene* s, uint32_t n, const float* x, const float* v,
                       const float* C, const float* F, const float* jp, const float* mass,
                       const float* vol, const uint32_t* mat, const float* E, float dt,
                       float* grid_m, float* grid_p) {
    if (!s || s->grid_size == 0 || s->model != ORC_MODEL_SHADER) return -1;
    const uint32_t G = s->grid_size, W = (G + 3u) / 4u;
    std::vector<uint32_t> keys(n), ids(n);
    orc_tile_keys(s, n, n, x, keys.data(), ids.data());
    orc_sort_pairs(n, keys.data(), ids.data());
    std::vector<uint32_t> lo(size_t(W) * W * W, 0), hi(size_t(W) * W * W, 0);
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t k = keys[j];
        if (j == 0 || keys[j - 1] != k) lo[k] = j;
        hi[k] = j + 1;
    }
    Particles<float> P{n, const_cast<float*>(x), const_cast<float*>(v), const_cast<float*>(C),
                       const_cast<float*>(F), const_cast<float*>(jp), mass, vol, E, mat};
    std::vector<M3<float>> stress(n);
    for (uint32_t p = 0; p < n; ++p) stress[p] = particle_stress_shader<float>(s, P, p, dt);
    const float dx = s->cell_size, inv = 1.0f / s->cell_size, inv2 = inv * inv;
    for (uint32_t cz = 0; cz < G; ++cz)
        for (uint32_t cy = 0; cy < G; ++cy)
            for (uint32_t cx = 0; cx < G; ++cx) {
                int cell[3] = {int(cx), int(cy), int(cz)};
                int tile[3] = {int(cx / 4), int(cy / 4), int(cz / 4)};
                float ms = 0.0f, mom[3] = {0.0f, 0.0f, 0.0f};
                for (int ti = 0; ti < 27; ++ti) {
                    int nb[3] = {tile[0] + ti % 3 - 1, tile[1] + (ti / 3) % 3 - 1, tile[2] + ti / 9 - 1};
                    if (nb[0] < 0 || nb[1] < 0 || nb[2] < 0 || nb[0] >= int(W) || nb[1] >= int(W) ||
                        nb[2] >= int(W))
                        continue;
                    uint32_t key = uint32_t(nb[0]) + W * (uint32_t(nb[1]) + W * uint32_t(nb[2]));
                    for (uint32_t sj = lo[key]; sj < hi[key]; ++sj) {
                        uint32_t pid = ids[sj];
                        float xp[3], fx[3];
                        int off[3];
                        bool ok = true;
                        for (int a = 0; a < 3; ++a) {
                            xp[a] = (x[3 * size_t(pid) + a] - s->origin[a]) * inv;
                            float bf = std::floor(xp[a] - 0.5f);
                            off[a] = cell[a] - (int(bf) - s->window[a]);
                            if (off[a] < 0 || off[a] > 2) ok = false;
                            fx[a] = xp[a] - bf;
                        }
                        if (!ok) continue;
                        float wx[3], wy[3], wz[3];
                        quad_w(fx[0], wx);
                        quad_w(fx[1], wy);
                        quad_w(fx[2], wz);
                        float w = wx[off[0]] * wy[off[1]] * wz[off[2]];
                        float mass_p = mass[pid];
                        V3<float> d = {{(float(off[0]) - fx[0]) * dx, (float(off[1]) - fx[1]) * dx,
                                        (float(off[2]) - fx[2]) * dx}};
                        V3<float> aff = mv(load9(C + 9 * size_t(pid)), d);
                        V3<float> sf = mv(stress[pid], d);
                        for (int a = 0; a < 3; ++a)
                            mom[a] += w * mass_p * (v[3 * size_t(pid) + a] + aff[a]) + w * (sf[a] * (4.0f * inv2));
                        ms += w * mass_p;
                    }
                }
                size_t c = size_t(cx) + size_t(G) * (size_t(cy) + size_t(G) * size_t(cz));
                grid_m[c] = ms;
                for (int a = 0; a < 3; ++a) grid_p[3 * c + a] = mom[a];
            }
    return 0;
}

int orc_diff_forward(const orc_scene* s, uint32_t n, double* x, double* v, double* C, double* F,
                     const double* mass, const double* vol, const double* E, double dt,
                     uint32_t steps, double* tape) {
    if (!s || s->grid_size == 0) return -1;
    return diff_forward(s, n, x, v, C, F, mass, vol, E, dt, steps, tape);
}

int orc_diff_backward(const orc_scene* s, uint32_t n, const double* mass, const double* vol,
                      const double* E, double dt, uint32_t steps, const double* tape, double* gx,
                      double* gv, double* gC, double* gF, double* gE, double* g_mu_lam) {
    if (!s || s->grid_size == 0 || !tape) return -1;
    DiffCtx c{s, n, mass, vol, E};
    const size_t per = size_t(n) * 24;
    for (uint32_t st = steps; st-- > 0;) {
        const double* t = tape + per * st;
        diff_backward_step(c, dt, t, t + 3 * size_t(n), t + 6 * size_t(n), t + 15 * size_t(n), gx,
                           gv, gC, gF, gE, g_mu_lam);
    }
    return 0;
}

}  // extern "C"
