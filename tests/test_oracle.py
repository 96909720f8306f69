"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

The reference commits no golden vectors, so each oracle function is pinned
by something other than its own formula:

==========================  ==================================================
oracle function             pinned by
==========================  ==================================================
orc_tile_keys / sort        closed-form key() + Python sorted() on the
                            reference's points() fixture (mpm_particle_order
                            _tests.rs:34-43, :66-150)
quad weights (+ dfx)        partition of unity (mpm_cpu_diff.rs:640-649),
                            first/second moments (sum w(i-fx) = 0,
                            sum w(i-fx)^2 = 1/4), finite differences
svd3 (shader, cpu)          reconstruction, orthogonality, det signs,
                            numpy.linalg.svd singular values, identity /
                            zero / diagonal (dense3.rs:516-543)
adj_svd3                    closed form d(sum s^2)/dF = 2F; FD of the
                            polar factor <W, U V^T>
fixed-corotated PK1         P(I)=0; P = dPsi/dF by FD of the energy
                            mu sum(s-1)^2 + lam/2 (J-1)^2 (numpy SVD)
corotated-SVK PK1 (+adj)    P(I)=0 (mpm_cpu_diff.rs:651-666); FD of
                            mu sum(s-1)^2 + lam/2 (sum s - 3)^2; the
                            reference's adjoint FD tests (:668-773)
plasticity return maps      the reference's material-oracle assertions
                            (tests/mpm_material_oracles.rs:19-320)
shader step (P2G/grid/G2P)  mass = sum m_p and zero momentum at rest
                            (mpm_gpu.rs:840-885); exact affine-field
                            reproduction (APIC: v = a + B x, C = B);
                            free fall v = g t (mpm_freefall_regression_gpu
                            .rs:76-111); closed-system momentum (mpm_gpu.rs
                            :781-838); block stays in column (:113-128)
deterministic gather P2G    equals the scatter P2G (f64) to f32 rounding
cpu-diff forward/adjoint    the reference's adjoint tests (mpm_adjoint.rs
                            :612-889) and full central FD in f64 on x0, v0
                            and per-particle E
==========================  ==================================================
"""
import math

import numpy as np
import pytest

import oracle as orc
from molla_amd.config import (MODEL_CPUDIFF, Fluid, Implicit, MpmConfig, Sand, Snow, Viscous,
                              drucker_prager_alpha, material_row)
from tests.fixtures import fill_box, order_config, points, reference_key

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- ordering
@pytest.mark.parametrize("size", [35, 1536])
def test_tile_keys_match_reference_closed_form_and_stable_sort(size):
    cfg = order_config(size)
    for count in [513, 257, 1, 0, 63]:
        pos = points(count)
        if count > 1:
            pos[0] = [-1.0625] * 3
            pos[1] = [size * cfg.cell_size] * 3
        keys, ids = orc.sorted_tile_keys(cfg, pos, capacity=513)
        expected = sorted(
            [(reference_key(cfg, pos[i]) if i < count else 0xFFFFFFFF, i) for i in range(513)],
            key=lambda e: e[0])
        assert keys.tolist() == [e[0] for e in expected]
        assert ids.tolist() == [e[1] for e in expected]


def test_sort_pairs_is_stable_lexicographic():
    keys = RNG.integers(0, 7, size=1000).astype(np.uint32)
    ids = np.arange(1000, dtype=np.uint32)
    k, i = orc.sort_pairs(keys, ids)
    exp = sorted(zip(keys.tolist(), ids.tolist()))
    assert list(zip(k.tolist(), i.tolist())) == exp


def test_base_cells_closed_form():
    cfg = MpmConfig.sand()
    x = fill_box([-0.3, 0.2, -0.3], [0.3, 0.8, 0.3], 9, 3)
    b = orc.base_cells(cfg, x)
    inv = np.float32(1.0) / np.float32(cfg.cell_size)
    xp = (x - np.asarray(cfg.grid_origin, np.float32)) * inv
    assert np.array_equal(b, np.floor(xp - np.float32(0.5)).astype(np.int32))


# ---------------------------------------------------------------- weights
def test_quad_weights_partition_moments_and_derivative():
    for k in range(11):
        fx = 0.5 + 0.1 * k
        if fx >= 1.5:
            fx = 1.5 - 1e-12
        w = orc.quad_weights(fx)
        off = np.array([0.0, 1.0, 2.0]) - fx
        assert abs(w.sum() - 1.0) < 1e-12
        assert abs((w * off).sum()) < 1e-12
        assert abs((w * off * off).sum() - 0.25) < 1e-12
        h = 1e-6
        fd = (orc.quad_weights(fx + h) - orc.quad_weights(fx - h)) / (2 * h)
        assert np.allclose(orc.quad_weights_dfx(fx), fd, atol=1e-8)


# ---------------------------------------------------------------- SVD
def _rand_F(n, spread=0.3, inverted=False):
    out = []
    for _ in range(n):
        A = np.eye(3) + spread * RNG.standard_normal((3, 3))
        if inverted and np.linalg.det(A) > 0:
            A[:, 0] *= -1
        if not inverted and np.linalg.det(A) < 0:
            A[:, 0] *= -1
        out.append(A)
    return out


@pytest.mark.parametrize("which", ["shader", "cpu"])
def test_svd_reconstruction_orthogonality_and_library_sigma(which):
    fn = orc.svd_shader if which == "shader" else orc.svd_cpu
    mats = _rand_F(200) + _rand_F(50, inverted=True) + [np.eye(3), np.zeros((3, 3)),
                                                        np.diag([2.0, 3.0, 5.0])]
    for F in mats:
        U, s, V = fn(F)
        assert np.abs(U @ np.diag(s) @ V.T - F).max() < 1e-12 * max(1.0, np.abs(F).max())
        if np.abs(F).max() == 0:
            assert np.abs(s).max() < 1e-12
            continue
        assert np.abs(U.T @ U - np.eye(3)).max() < 1e-12
        assert np.abs(V.T @ V - np.eye(3)).max() < 1e-12
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.sort(np.abs(s)), np.sort(ref), atol=1e-12)
        if which == "shader":
            assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
            assert (np.sum(s < 0) == 1) == (np.linalg.det(F) < 0)
        else:
            assert np.all(s >= 0) and np.linalg.det(U) > 0


def test_svd_shader_f32_reconstruction():
    for F in _rand_F(100):
        U, s, V = orc.svd_shader(F.astype(np.float32), dtype=np.float32)
        assert np.abs(U.astype(np.float64) @ np.diag(s) @ V.T - F.astype(np.float32)).max() < 2e-5


def test_adj_svd_sigma_closed_form_and_polar_fd():
    for F in _rand_F(20):
        U, s, V = orc.svd_cpu(F)
        g = orc.adj_svd(U, s, V, np.zeros((3, 3)), 2 * s, np.zeros((3, 3)))
        assert np.allclose(g, 2 * F, atol=1e-10)
        W = RNG.standard_normal((3, 3))
        U2 = W @ V  # adj of <W, U V^T> wrt U is W V, wrt V is W^T U
        V2 = W.T @ U
        g = orc.adj_svd(U, s, V, U2, np.zeros(3), V2)
        fd = np.zeros((3, 3))
        h = 1e-6
        for i in range(3):
            for j in range(3):
                Fp = F.copy(); Fp[i, j] += h
                Fm = F.copy(); Fm[i, j] -= h
                Up, _, Vp = np.linalg.svd(Fp)
                Um, _, Vm = np.linalg.svd(Fm)
                fd[i, j] = (np.sum(W * (Up @ Vp)) - np.sum(W * (Um @ Vm))) / (2 * h)
        assert np.allclose(g, fd, atol=1e-6)


# ---------------------------------------------------------------- PK1
def _lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def _fd_energy(psi, F, h=1e-6):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp = F.copy(); Fp[i, j] += h
            Fm = F.copy(); Fm[i, j] -= h
            P[i, j] = (psi(Fp) - psi(Fm)) / (2 * h)
    return P


def test_fixed_corotated_pk1_is_energy_gradient():
    mu, lam = _lame(1.0e4, 0.2)
    assert np.abs(orc.pk1_fixed_corotated(np.eye(3), mu, lam)).max() < 1e-9

    def psi(F):
        s = np.linalg.svd(F, compute_uv=False)
        J = np.linalg.det(F)
        return mu * np.sum((s - 1) ** 2) + 0.5 * lam * (J - 1) ** 2

    for F in _rand_F(20, spread=0.15):
        P = orc.pk1_fixed_corotated(F, mu, lam)
        fd = _fd_energy(psi, F)
        assert np.abs(P - fd).max() < 1e-4 * max(1.0, np.abs(fd).max())


def test_corotated_svk_pk1_is_energy_gradient():
    mu, lam = _lame(1.0e4, 0.3)
    assert np.abs(orc.pk1_corotated_svk(np.eye(3), mu, lam)).max() < 1e-10

    def psi(F):
        s = np.linalg.svd(F, compute_uv=False)
        return mu * np.sum((s - 1) ** 2) + 0.5 * lam * (np.sum(s) - 3) ** 2

    for F in _rand_F(20, spread=0.15):
        P = orc.pk1_corotated_svk(F, mu, lam)
        fd = _fd_energy(psi, F)
        assert np.abs(P - fd).max() < 1e-4 * max(1.0, np.abs(fd).max())


def test_corotated_adjoint_fd_oracle():
    # mpm_cpu_diff.rs:668-733
    F = np.array([[1.2, 0.0, 0.03], [0.1, 1.1, -0.02], [-0.05, 0.07, 0.95]])
    mu, lam = _lame(2.0e4, 0.3)
    ptgt = np.array([[0.5, 0.0, 0.0], [0.0, -0.4, 0.0], [0.0, 0.1, 0.7]])
    adjP = 2 * (orc.pk1_corotated_svk(F, mu, lam) - ptgt)
    adjF, _, _ = orc.pk1_corotated_svk_adjoint(F, mu, lam, adjP)

    def loss(Fx):
        d = orc.pk1_corotated_svk(Fx, mu, lam) - ptgt
        return np.sum(d * d)

    h = 1e-7
    for i in range(3):
        for j in range(3):
            Fp = F.copy(); Fp[i, j] += h
            Fm = F.copy(); Fm[i, j] -= h
            fd = (loss(Fp) - loss(Fm)) / (2 * h)
            rel = abs(adjF[i, j] - fd) / max(abs(adjF[i, j]), abs(fd), 1.0)
            assert rel < 1e-3


def test_corotated_adjoint_mu_lambda_fd_oracle():
    # mpm_cpu_diff.rs:735-772
    F = np.array([[1.05, -0.01, 0.0], [0.02, 0.98, 0.04], [0.0, 0.03, 1.02]])
    mu, lam = _lame(1.5e4, 0.3)
    adjP = 2 * orc.pk1_corotated_svk(F, mu, lam)
    _, amu, alam = orc.pk1_corotated_svk_adjoint(F, mu, lam, adjP)

    def loss(m, l):
        P = orc.pk1_corotated_svk(F, m, l)
        return np.sum(P * P)

    fd_mu = (loss(mu + 1e-3, lam) - loss(mu - 1e-3, lam)) / 2e-3
    fd_l = (loss(mu, lam + 1e-3) - loss(mu, lam - 1e-3)) / 2e-3
    assert abs(amu - fd_mu) / max(abs(fd_mu), abs(amu), 1.0) < 1e-4
    assert abs(alam - fd_l) / max(abs(fd_l), abs(alam), 1.0) < 1e-4


# ---------------------------------------------------------------- plasticity
def _sand_row(phi):
    return (0.0, drucker_prager_alpha(phi), 0, 0, 0, 0, 0, 0)


def test_sand_return_map_reference_assertions():
    cfg = MpmConfig.sand().copy(youngs=1.0e5, poisson=0.3)
    s, jp = orc.plastic_project(cfg, _sand_row(0.5), [1.2, 1.1, 1.0], 0.0)
    assert np.allclose(s, 1.0) and jp > 0.0
    s, jp = orc.plastic_project(cfg, _sand_row(math.radians(30)), [0.95] * 3, 0.2)
    assert np.allclose(s, 0.95, atol=1e-15) and abs(jp - 0.2) < 1e-7
    sv = np.array([0.50, 0.90, 1.10])
    s, jp = orc.plastic_project(cfg, _sand_row(math.radians(5)), sv, 0.0)
    assert np.linalg.norm(s - sv) > 1e-3 and jp > 0.0 and s[0] > sv[0]
    sv = np.array([0.42, 0.82, 1.05])
    lo, jlo = orc.plastic_project(cfg, _sand_row(math.radians(5)), sv, 0.0)
    hi, jhi = orc.plastic_project(cfg, _sand_row(math.radians(35)), sv, 0.0)
    dev = lambda t: np.linalg.norm(t - t.mean())
    assert dev(hi) > dev(lo) and dev(hi) <= dev(sv) + 1e-12
    assert np.linalg.norm(hi - sv) < np.linalg.norm(lo - sv) and jhi < jlo
    _, jp = orc.plastic_project(cfg, _sand_row(0.5), [20.0] * 3, 4.9)
    assert jp == 5.0


def test_snow_return_map_reference_assertions():
    cfg = MpmConfig.snow()
    row = material_row(Snow(0.025, 0.0075, 10.0))
    s, jp = orc.plastic_project(cfg, row, [0.80, 1.20, 1.0], 0.0)
    assert abs(s[0] - (1 - 0.025)) < 1e-7 and abs(s[1] - (1 + 0.0075)) < 1e-7 and s[2] == 1.0
    sv = np.array([0.5, 1.5, 1.1])
    s, jp = orc.plastic_project(cfg, row, sv, 0.25)
    cl = np.clip(sv, 1 - 0.025, 1 + 0.0075)
    exp = 0.25 + math.log(max(np.prod(sv), 1e-6)) - math.log(max(np.prod(cl), 1e-6))
    assert abs(jp - min(max(exp, -5), 5)) < 1e-6
    jp = 0.0
    prev = jp
    for _ in range(12):
        s, jp = orc.plastic_project(cfg, row, [0.82, 0.94, 1.06], jp)
        assert np.all((s >= 1 - 0.025 - 1e-7) & (s <= 1 + 0.0075 + 1e-7))
        assert jp <= prev and -5 <= jp <= 5
        prev = jp
    assert jp < -0.5
    _, jhi = orc.plastic_project(cfg, row, [20.0] * 3, 4.9)
    _, jlo = orc.plastic_project(cfg, row, [1e-3] * 3, -4.9)
    assert jhi == 5.0 and jlo == -5.0


# ---------------------------------------------------------------- shader step
def _state(cfg, x, v=None, dtype=np.float64, **kw):
    n = len(x)
    return orc.State(x, v, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume),
                     dtype=dtype, **kw)


def test_grid_mass_matches_particle_mass_and_rest_momentum_zero():
    # tests/mpm_gpu.rs:840-885
    cfg = MpmConfig.sand().copy(gravity_y=0.0, grid_origin=(-1.6, -1.6, -1.6), grid_size=32,
                                cell_size=0.1, particle_mass=2.5e-3, boundary_friction=1.0)
    x = fill_box([-0.35] * 3, [0.35] * 3, 7, 19)
    st = _state(cfg, x)
    gm, gp = orc.step(cfg, st, 1 / 240, 1, want_grid=True)
    exp = cfg.particle_mass * len(x)
    assert abs(gm.sum() - exp) / exp < 1e-12
    assert np.abs(gp).max() < 1e-12


def test_affine_velocity_field_is_reproduced_exactly():
    # APIC with quadratic B-splines reproduces affine fields: after one
    # stress-free step v_p = a + B x_p (old x) and C_p = B.
    cfg = MpmConfig.viscous().copy(gravity_y=0.0, rheology=Viscous(0.0, 0.0), grid_size=32,
                                   cell_size=0.05, grid_origin=(-0.8, -0.8, -0.8))
    x = fill_box([-0.2] * 3, [0.2] * 3, 6, 5).astype(np.float64)
    a = np.array([0.3, -0.2, 0.1])
    B = np.array([[0.1, -0.4, 0.2], [0.4, 0.05, -0.3], [-0.2, 0.3, -0.15]])
    v = a + x @ B.T
    st = _state(cfg, x, v, C_=np.broadcast_to(B, (len(x), 3, 3)))
    st.vol[:] = 0.0  # stress-free
    x0 = st.x.copy()
    orc.step(cfg, st, 1e-4, 1)
    assert np.abs(st.v - (a + x0 @ B.T)).max() < 1e-12
    assert np.abs(st.C - B).max() < 1e-11
    assert np.abs(st.x - (x0 + 1e-4 * st.v)).max() < 1e-15


def test_single_particle_free_fall_is_exact():
    # tests/mpm_freefall_regression_gpu.rs:76-111
    cfg = MpmConfig.snow()
    for dtype in (np.float64, np.float32):
        st = _state(cfg, np.array([[0.013, 0.2, 0.017]]), dtype=dtype)
        dt = 1.0 / 480.0
        for step in range(20):
            orc.step(cfg, st, dt, 1)
            vy = -9.81 * dt * (step + 1)
            assert abs(st.v[0, 1] - vy) < 1e-4
            assert abs(st.v[0, 0]) < 1e-5 and abs(st.v[0, 2]) < 1e-5
            assert abs(st.x[0, 0] - 0.013) < 1e-5 and abs(st.x[0, 2] - 0.017) < 1e-5


def test_closed_system_momentum_drift():
    # tests/mpm_gpu.rs:781-838
    cfg = MpmConfig.sand().copy(gravity_y=0.0, grid_origin=(-1.6, -1.6, -1.6), grid_size=32,
                                cell_size=0.1, boundary_friction=1.0)
    x = fill_box([-0.4] * 3, [0.4] * 3, 8, 7)
    st = _state(cfg, x, dtype=np.float32)
    orc.step(cfg, st, 1 / 240, 60)
    p = np.linalg.norm(st.v.astype(np.float64).sum(0)) * cfg.particle_mass
    assert p < 0.01 * cfg.particle_mass * len(x)


def test_dropped_block_settles_in_its_column():
    # tests/mpm_freefall_regression_gpu.rs:113-128
    cfg = MpmConfig.snow()
    x = fill_box([-0.075, 0.12, -0.075], [0.075, 0.27, 0.075], 6, 22)
    st = _state(cfg, x, dtype=np.float32)
    orc.step(cfg, st, 1 / 480, 300)
    assert st.x[:, 1].max() < 0.35 and np.abs(st.x[:, 0]).max() < 0.25
    assert st.status[0] == 0


def test_f32_restatement_tracks_f64():
    cfg = MpmConfig.sand().copy(youngs=1.0e4, poisson=0.2, rheology=Viscous(0, 0),
                                boundary_friction=0.0)
    x = fill_box([-0.1, 0.1, -0.1], [0.1, 0.3, 0.1], 8, 11)
    s64 = _state(cfg, x)
    s32 = _state(cfg, x, dtype=np.float32)
    orc.step(cfg, s64, 1e-3, 50)
    orc.step(cfg, s32, 1e-3, 50)
    assert np.abs(s32.x - s64.x).max() < 1e-5
    assert np.abs(s32.F - s64.F).max() < 1e-3


def test_threaded_port_tracks_the_serial_step():
    # The CPU baseline's threaded f32 port (two-phase z-slab scatter) against
    # the serial f32 and f64 steps: a mixed elastic/fluid/sand block on the
    # floor, started partly outside the grid's low z face so that clamped
    # base cells land in the first slab.
    cfg = MpmConfig.sand().copy(youngs=1.0e4, poisson=0.2, boundary_friction=0.0,
                                materials=[Viscous(0, 0), Fluid(1.0e4), Sand()])
    x = fill_box([-0.3, 0.02, -0.83], [0.3, 0.3, -0.3], 10, 12)
    assert (x[:, 2] < cfg.grid_origin[2]).any()
    mat = (np.arange(len(x)) % 3).astype(np.uint32)
    runs = {}
    for name, dtype in (("f64", np.float64), ("f32", np.float32), ("threads", np.float32)):
        st = _state(cfg, x, dtype=dtype)
        st.mat = mat.copy()
        if name == "threads":
            assert orc.step_threads(cfg, st, 1e-3, 40) >= 1
        else:
            orc.step(cfg, st, 1e-3, 40)
        assert st.status[0] == 0 and st.status[1] == 40
        runs[name] = st
    for f in ("x", "v", "F"):
        ref = getattr(runs["f64"], f)
        e32 = np.abs(getattr(runs["f32"], f) - ref).max()
        eth = np.abs(getattr(runs["threads"], f) - ref).max()
        assert eth <= 8 * e32 + 1e-6 * np.abs(ref).max(), (f, eth, e32)
    # the f64 threaded step is the serial f64 step up to summation order
    st = _state(cfg, x)
    st.mat = mat.copy()
    orc.step_threads(cfg, st, 1e-3, 40)
    for f in ("x", "v", "F"):
        ref = getattr(runs["f64"], f)
        assert np.abs(getattr(st, f) - ref).max() <= 1e-11 * np.abs(ref).max(), f


def test_fault_latch_freezes_state():
    cfg = MpmConfig.sand()
    x = fill_box([-0.1, 0.3, -0.1], [0.1, 0.5, 0.1], 4, 2)
    st = _state(cfg, x, dtype=np.float32)
    st.v[3, 0] = np.nan
    x0 = st.x.copy()
    orc.step(cfg, st, 1 / 240, 3)
    assert st.status[0] & 2
    assert np.isnan(st.v[3, 0])
    # substep 1 faulted inside G2P: the faulting particle is untouched and
    # later substeps do nothing (gather counter frozen at 1).
    assert np.array_equal(st.x[3], x0[3]) and st.status[1] == 1


def test_fluid_rheology_keeps_isotropic_F():
    cfg = MpmConfig.sand().copy(rheology=Fluid(1.0e4), youngs=1.0e4)
    x = fill_box([-0.1, 0.1, -0.1], [0.1, 0.3, 0.1], 6, 4)
    st = _state(cfg, x)
    orc.step(cfg, st, 2e-4, 20)
    off = st.F - np.einsum("p,ij->pij", st.F[:, 0, 0], np.eye(3))
    assert np.abs(off).max() < 1e-15 and np.all(np.isfinite(st.x))


def test_deterministic_gather_equals_scatter():
    cfg = MpmConfig.sand().copy(youngs=1.0e4, rheology=Viscous(0, 0))
    x = fill_box([-0.2, 0.2, -0.2], [0.2, 0.6, 0.2], 10, 9)
    n = len(x)
    Fs = np.stack(_rand_F(n, spread=0.05))
    v = RNG.standard_normal((n, 3)) * 0.1
    Cm = RNG.standard_normal((n, 3, 3)) * 0.1
    st = _state(cfg, x.astype(np.float32), v, C_=Cm, F=Fs, dtype=np.float32)
    gm32, gp32 = orc.p2g_gather_f32(cfg, st, 1e-3)
    st64 = _state(cfg, st.x.astype(np.float64), st.v, C_=st.C, F=st.F)
    gm, gp = orc.p2g(cfg, st64, 1e-3)
    assert np.abs(gm32 - gm).max() < 1e-6 * gm.max()
    assert np.abs(gp32 - gp).max() < 1e-5 * np.abs(gp).max()


# ---------------------------------------------------------------- cpu-diff + adjoint
def _diff_cfg(**kw):
    base = dict(grid_size=12, cell_size=0.1, grid_origin=(-0.6, -0.05, -0.6), gravity_y=-9.81,
                youngs=1.0e4, poisson=0.3, particle_volume=1.0e-3, particle_mass=1.0e-2,
                floor_y=0.0, floor_friction=0.4, model=MODEL_CPUDIFF)
    base.update(kw)
    return MpmConfig(**base)


def _cube27():
    # mpm_adjoint.rs:582-610 make_cube_particles
    pos, fs = [], []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                pos.append([(i - 1) * 0.1, 0.20 + j * 0.1, (k - 1) * 0.1])
                fs.append(np.diag([1 + 0.10 * (i - 1), 1 + 0.05 * (j - 1), 1 + 0.08 * (k - 1)]))
    return np.array(pos), np.array(fs)


def _diff_state(cfg, x, v=None, F=None, E=None):
    n = len(x)
    return orc.State(x, v, F=F, mass=np.full(n, cfg.particle_mass),
                     vol=np.full(n, cfg.particle_volume), E=E)


def _centroid_loss(x, target):
    d = x.mean(0) - target
    return float(d @ d), np.broadcast_to(2 * d / len(x), x.shape).copy()


def test_cpu_diff_forward_falls_under_gravity():
    # mpm_cpu_diff.rs:612-637
    cfg = MpmConfig.cpu_diff_default().copy(particle_volume=1.0e-5)
    st = _diff_state(cfg, np.array([[0.0, 0.5, 0.0]]))
    orc.diff_forward(cfg, st, 1 / 240, 30, record=False)
    assert np.isfinite(st.x[0, 1]) and st.x[0, 1] < 0.5


def test_backward_e_zero_when_f_is_identity():
    cfg = _diff_cfg(grid_size=8, grid_origin=(-0.4, -0.05, -0.4), gravity_y=0.0, floor_y=-10.0,
                    floor_friction=0.0)
    st = _diff_state(cfg, np.array([[0.05, 0.05, 0.05]]), np.array([[0.1, 0.0, 0.0]]))
    st0 = st.copy()
    tape = orc.diff_forward(cfg, st, 0.01, 1)
    _, g = _centroid_loss(st.x, np.zeros(3))
    out = orc.diff_backward(cfg, st0, 0.01, 1, tape, g)
    assert abs(out["youngs"]) < 1e-10


@pytest.mark.parametrize("case", ["two_step", "multi_particle"])
def test_backward_e_displacement_fd(case):
    # mpm_adjoint.rs:637-800
    if case == "two_step":
        cfg = _diff_cfg(grid_size=10, grid_origin=(-0.5, -0.5, -0.5), gravity_y=0.0,
                        floor_y=-10.0, floor_friction=0.0)
        dt, fs_scale = 0.005, (0.15, 0.10, None)
    else:
        cfg = _diff_cfg(grid_size=12, grid_origin=(-0.6, -0.6, -0.6), gravity_y=0.0,
                        floor_y=-10.0, floor_friction=0.0)
        dt, fs_scale = 0.001, (0.2, 0.1, 0.05)
    pos, vel, fs = [], [], []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                pos.append([-0.05 + 0.1 * i, -0.05 + 0.1 * j, -0.05 + 0.1 * k])
                vel.append([0.2 * (i - 0.5), -0.05, 0.0] if case == "two_step" else [0, 0, 0])
                sz = 1 + 0.05 * k if fs_scale[2] is None else 1 + fs_scale[2] * (k - 0.5)
                fs.append(np.diag([1 + fs_scale[0] * (i - 0.5), 1 + fs_scale[1] * (j - 0.5), sz]))
    st = _diff_state(cfg, np.array(pos), np.array(vel), np.array(fs))
    start = st.x.copy()
    st0 = st.copy()
    tape = orc.diff_forward(cfg, st, dt, 3)
    g = 2 * (st.x - start)
    ge = orc.diff_backward(cfg, st0, dt, 3, tape, g)["youngs"]

    def loss(E):
        s = st0.copy()
        orc.diff_forward(cfg.copy(youngs=E), s, dt, 3, record=False)
        return float(np.sum((s.x - start) ** 2))

    h = cfg.youngs * 1e-3
    fd = (loss(cfg.youngs + h) - loss(cfg.youngs - h)) / (2 * h)
    assert abs(ge) > 1e-16
    assert abs(ge - fd) / max(abs(ge), abs(fd)) < 5e-2


def test_backward_e_matches_fd_oracle():
    # mpm_adjoint.rs:852-889
    cfg = _diff_cfg()
    pos, fs = _cube27()
    st = _diff_state(cfg, pos, F=fs)
    st0 = st.copy()
    dt, steps, target = 1 / 240, 5, np.array([0.05, 0.05, 0.0])
    tape = orc.diff_forward(cfg, st, dt, steps)
    _, g = _centroid_loss(st.x, target)
    ge = orc.diff_backward(cfg, st0, dt, steps, tape, g)["youngs"]

    def loss(E):
        s = st0.copy()
        orc.diff_forward(cfg.copy(youngs=E), s, dt, steps, record=False)
        return _centroid_loss(s.x, target)[0]

    h = cfg.youngs * 1e-4
    fd = (loss(cfg.youngs + h) - loss(cfg.youngs - h)) / (2 * h)
    assert abs(ge - fd) / max(abs(ge), abs(fd), 1e-30) < 1e-2


def test_adjoint_full_central_fd_x0_v0_and_per_particle_E():
    cfg = _diff_cfg()
    pos, fs = _cube27()
    n = len(pos)
    E = 1.0e4 * np.exp(0.3 * RNG.standard_normal(n))
    v0 = 0.2 * RNG.standard_normal((n, 3))
    st = _diff_state(cfg, pos, v0, fs, E=E)
    st0 = st.copy()
    # dt = 1e-3 keeps the 27-particle cube in the stable regime; at the
    # reference demo's dt = 1/240 it blows up by step 6 and FD is meaningless.
    dt, steps, target = 1e-3, 12, np.array([0.05, 0.05, 0.0])
    tape = orc.diff_forward(cfg, st, dt, steps)
    w = RNG.standard_normal((n, 3))

    def loss_of(s):
        return _centroid_loss(s.x, target)[0] + float(np.sum(w * s.x))

    _, g = _centroid_loss(st.x, target)
    out = orc.diff_backward(cfg, st0, dt, steps, tape, g + w)

    def run(mod):
        s = st0.copy()
        mod(s)
        orc.diff_forward(cfg, s, dt, steps, record=False)
        return loss_of(s)

    checks = []
    for p, a in [(0, 0), (5, 1), (13, 2), (26, 1)]:
        h = 1e-7
        def mx(s, sgn): s.x[p, a] += sgn * h
        def mv(s, sgn): s.v[p, a] += sgn * h
        fdx = (run(lambda s: mx(s, 1)) - run(lambda s: mx(s, -1))) / (2 * h)
        fdv = (run(lambda s: mv(s, 1)) - run(lambda s: mv(s, -1))) / (2 * h)
        checks += [(out["x"][p, a], fdx), (out["v"][p, a], fdv)]
    for p in (1, 8, 20):
        h = E[p] * 1e-6
        def me(s, sgn): s.E[p] += sgn * h
        fde = (run(lambda s: me(s, 1)) - run(lambda s: me(s, -1))) / (2 * h)
        checks.append((out["E"][p], fde))
    for an, fd in checks:
        assert abs(an - fd) <= 1e-5 * max(abs(an), abs(fd)) + 1e-10, (an, fd)


def test_implicit_sweep_closed_forms():
    # mpm_implicit_apply / mpm_implicit_swap (mpm.wgsl:1060-1153):
    # v' = (m v + a v_avg) / (m + a), a = dt^2 m (E / 1500) / dx^2, v_avg over
    # face neighbours with m > 1e-9; sweeps ping-pong.
    G, dt, E, dx = 8, 1e-2, 2.0e4, 0.1
    cfg1 = MpmConfig.sand().copy(grid_size=G, cell_size=dx, youngs=E, time_step=Implicit(1))
    cfg2 = cfg1.copy(time_step=Implicit(2))

    def idx(x, y, z):
        return x + G * (y + G * z)

    m = np.zeros(G ** 3)
    v = np.zeros((G ** 3, 3))
    iso, a_, b_, tiny = idx(1, 1, 1), idx(4, 4, 4), idx(5, 4, 4), idx(4, 5, 4)
    m[iso], v[iso] = 0.5, (1.0, 2.0, 3.0)
    m[a_], v[a_] = 0.2, (1.0, 0.0, 0.0)
    m[b_], v[b_] = 0.8, (-1.0, 0.5, 0.0)
    m[tiny], v[tiny] = 1e-10, (100.0, 100.0, 100.0)  # below the mass threshold: ignored, copied

    dx32 = float(np.float32(dx))  # the scene stores cell_size as f32

    def alpha(mm):
        return dt * dt * mm * (E / 1500.0) / (dx32 * dx32)

    def one(va, vb):
        return ((m[a_] * va + alpha(m[a_]) * vb) / (m[a_] + alpha(m[a_])),
                (m[b_] * vb + alpha(m[b_]) * va) / (m[b_] + alpha(m[b_])))

    out1 = orc.implicit_sweeps(cfg1, dt, m, v)
    ea, eb = one(v[a_], v[b_])
    np.testing.assert_allclose(out1[a_], ea, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(out1[b_], eb, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(out1[iso], v[iso], rtol=1e-14)  # v_avg = v: unchanged up to rounding
    np.testing.assert_array_equal(out1[tiny], v[tiny])
    out2 = orc.implicit_sweeps(cfg2, dt, m, v)
    ea2, eb2 = one(ea, eb)
    np.testing.assert_allclose(out2[a_], ea2, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(out2[b_], eb2, rtol=1e-13, atol=1e-15)
    # momentum of the pair moves towards the common velocity, never past it
    assert np.all(np.abs(out1[a_] - out1[b_]) <= np.abs(v[a_] - v[b_]) + 1e-15)


def test_implicit_step_keeps_uniform_translation():
    # A uniform grid velocity is a fixed point of the sweep (v_avg = v), so the
    # semi-implicit step moves a freely translating block like the explicit one.
    cfg = MpmConfig.sand().copy(gravity_y=0.0, youngs=2.0e4, boundary_friction=1.0,
                                rheology=Viscous(0, 0))
    x = fill_box([-0.15, 0.3, -0.15], [0.15, 0.6, 0.15], 8, 3)
    v = np.tile(np.array([0.2, -0.1, 0.05]), (len(x), 1))
    se = orc.State(x, v, dtype=np.float64)
    si = orc.State(x, v, dtype=np.float64)
    orc.step(cfg, se, 1e-3, 20)
    orc.step(cfg.copy(time_step=Implicit(3)), si, 1e-3, 20)
    assert np.abs(se.x - si.x).max() < 1e-12
    assert np.abs(se.v - si.v).max() < 1e-10
