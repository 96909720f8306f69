"""Closed-form pins of the oracle's boundary and clamp rules.

The shader model's wall rule, position clamp, the two singular-value clamps
and the build-defined fluid reset are each checked against what they must
give on inputs where the answer is known in closed form:

=============================  ================================================
rule (reference)               pinned by
=============================  ================================================
6-wall, pad-2 grid rule with   every cell of a 16^3 grid with random (m, p):
boundary_friction              a cell in the walls S gets, per component a,
(mpm.wgsl:1043-1055)           max/min(v_a, 0) * bf^(|S|-1) for a in S and
                               v_a * bf^|S| otherwise (bf >= 0), gravity
                               added first; cells 2..G-3 untouched; and a
                               particle whose grid nodes all lie in one pad
                               (or straddle pad rows 1 and 2) for each face
                               and bf in {0, 0.3, 1}
position clamp                 x ends at origin + (window + 2) dx or
(mpm.wgsl:1321-1334)           origin + (window + G - 3) dx exactly, the
                               offending velocity component is 0, the other
                               two are the free particle's
stress sigma clamp [0.1, 10]   P2G of one particle: F with sigma 0.05 (12)
(mpm.wgsl:527-538)             scatters what sigma 0.1 (10) scatters; for a
                               diagonal F the Kirchhoff stress is
                               tau_k = 2 mu (s_k - 1) s_k + lam (J - 1) J of
                               the clamped s (fixed-corotated energy)
plasticity clamp [0.05, 20]    one stress-free step of F = R diag(0.01, 1, 30)
(mpm.wgsl:1459-1468)           Q^T ends at R diag(0.05, 1, 20) Q^T
fluid reset (build-defined)    F -> cbrt(det F) I (det F preserved, clamped
                               to [0.05, 20]); pressure tau = bulk (J - 1) J
=============================  ================================================

Each check is also run with a mistake planted in the oracle (a 1-cell pad,
friction on the normal or on one tangential only, an off-by-one clamp, a
clamp that keeps the velocity, a wrong sigma clamp, a sqrt(J) fluid reset,
the volumetric term on R instead of the cofactor) and must fail there.
"""
import contextlib

import numpy as np
import pytest

import oracle as orc
from molla_amd.config import Fluid, MpmConfig, Viscous

G = 16
DX = 1.0 / 16.0


@contextlib.contextmanager
def planted(m):
    orc.set_mutation(m)
    try:
        yield
    finally:
        orc.set_mutation(0)


def _cfg(**kw):
    base = dict(grid_size=G, cell_size=DX, grid_origin=(0.0, 0.0, 0.0), gravity_y=0.0, youngs=1.0e4,
                poisson=0.2, boundary_friction=0.0, particle_mass=1e-3, particle_volume=1e-3,
                max_particles=1, rheology=Viscous(0.0, 0.0))
    base.update(kw)
    return MpmConfig(**base)


def _one(cfg, x, v=(0.0, 0.0, 0.0), F=None, C_=None, vol=0.0, mass=1e-3):
    F = np.eye(3) if F is None else np.asarray(F, np.float64)
    return orc.State(np.array([x], np.float64), np.array([v], np.float64),
                     C_=None if C_ is None else np.array([C_]), F=np.array([F]),
                     mass=np.array([mass]), vol=np.array([vol]))


def _rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    return q if np.linalg.det(q) > 0 else -q


# ---------------------------------------------------------------- walls
def _walls_of(ix, iy, iz):
    """(axis, low side) of every wall pad the cell lies in."""
    out = []
    for a, c in enumerate((ix, iy, iz)):
        if c < 2:
            out.append((a, True))
        elif c >= G - 2:
            out.append((a, False))
    return out


def check_wall_cells():
    rng = np.random.default_rng(11)
    dt = 1e-3
    for bf in (0.0, 0.3, 1.0):
        for g in (0.0, -9.8):
            cfg = _cfg(boundary_friction=bf, gravity_y=g)
            m = np.full(G ** 3, 2.0)
            v = rng.uniform(-1.0, 1.0, (G ** 3, 3))
            # a few cells just above the floor whose v_y + dt g changes sign
            v[::97, 1] = 0.004
            vo = orc.grid_update(cfg, dt, m, m[:, None] * v)
            for iz in range(G):
                for iy in range(G):
                    for ix in range(G):
                        c = ix + G * (iy + G * iz)
                        vin = v[c].copy()
                        vin[1] += dt * float(np.float32(g))  # the scene holds gravity in f32
                        walls = _walls_of(ix, iy, iz)
                        exp = np.empty(3)
                        axes = {a: low for a, low in walls}
                        f = float(np.float32(bf))  # the scene holds the friction in f32
                        for a in range(3):
                            if a in axes:
                                n = max(vin[a], 0.0) if axes[a] else min(vin[a], 0.0)
                                exp[a] = n * f ** (len(walls) - 1)
                            else:
                                exp[a] = vin[a] * f ** len(walls)
                        assert np.allclose(vo[c], exp, rtol=1e-13, atol=1e-15), (bf, g, ix, iy, iz, vo[c], exp)
    # an empty cell (m <= 1e-9) reads zero velocity
    cfg = _cfg(boundary_friction=1.0, gravity_y=-9.8)
    m = np.full(G ** 3, 1e-9)
    vo = orc.grid_update(cfg, dt, m, np.ones((G ** 3, 3)))
    assert not vo.any()


def check_wall_particles():
    """One particle whose in-grid nodes all lie in one face's pad, and one
    whose nodes straddle pad row 1 and interior row 2, per face."""
    dt = 1e-4
    w_lo = orc.quad_weights(1.3)  # fx = 1.3: the base node -1 lies outside
    w_hi = orc.quad_weights(0.7)  # fx = 0.7: the last node G lies outside
    for bf in (0.0, 0.3, 1.0):
        cfg = _cfg(boundary_friction=bf)
        bf = float(np.float32(bf))  # the scene holds the friction in f32
        for a in range(3):
            for low in (True, False):
                # all nodes in the pad: cell coordinate 0.3 (nodes -1, 0, 1)
                # or its mirror G - 1.3 (nodes G-2, G-1, G); normal velocity
                # inwards 0.4 m/s, tangentials (0.7, -0.5) in cyclic order
                x = np.full(3, 8.3 * DX)
                x[a] = (0.3 if low else G - 1.3) * DX
                v = np.zeros(3)
                v[a] = 0.4 if low else -0.4
                v[(a + 1) % 3], v[(a + 2) % 3] = 0.7, -0.5
                st = _one(cfg, x, v)
                orc.step(cfg, st, dt, 1)
                # every in-grid node carries the wall's velocity; the node
                # outside the grid drops its weight (fx = 1.3, or 0.7 mirrored)
                keep = 1.0 - (w_lo[0] if low else w_hi[2])
                exp = np.empty(3)
                exp[a] = 0.0  # the particle lies beyond the clamp: v_a zeroed
                exp[(a + 1) % 3] = keep * bf * 0.7
                exp[(a + 2) % 3] = keep * bf * -0.5
                assert np.allclose(st.v[0], exp, rtol=1e-12, atol=1e-15), (bf, a, low, st.v[0], exp)
                lo, hi = 2 * DX, (G - 3) * DX
                assert st.x[0, a] == (lo if low else hi)
                # nodes on pad row 1 and interior rows 2, 3 (cell coordinate
                # 2.2, fx = 1.2; mirror G - 3.2: rows G-4, G-3 and pad row
                # G-2), moving away from the wall: the tangentials are scaled
                # by bf on the pad row only
                x = np.full(3, 8.3 * DX)
                x[a] = (2.2 if low else G - 3.2) * DX
                v = np.zeros(3)
                v[a] = 0.3 if low else -0.3
                v[(a + 1) % 3], v[(a + 2) % 3] = 0.7, -0.5
                st = _one(cfg, x, v)
                orc.step(cfg, st, dt, 1)
                w_pad = orc.quad_weights(1.2)[0] if low else orc.quad_weights(0.8)[2]
                f = w_pad * bf + (1.0 - w_pad)
                exp = np.empty(3)
                exp[a] = v[a]
                exp[(a + 1) % 3] = f * 0.7
                exp[(a + 2) % 3] = f * -0.5
                assert np.allclose(st.v[0], exp, rtol=1e-12, atol=1e-15), (bf, a, low, st.v[0], exp)


def test_wall_rule_on_every_cell():
    check_wall_cells()


def test_wall_rule_through_a_particle():
    check_wall_particles()


# ---------------------------------------------------------------- position clamp
def check_position_clamp():
    dt = 1e-3
    for window in ((0, 0, 0), (3, -5, 7)):
        cfg = _cfg(boundary_friction=1.0, grid_window_cells=window)
        for a in range(3):
            for low in (True, False):
                # 0.3 cells inside the clamp bound, moving 0.6 cells per step
                # towards the wall; friction 1 leaves the tangentials alone
                cell = np.array([8.25, 8.25, 8.25]) + np.asarray(window, np.float64)
                bound = window[a] + (2 if low else G - 3)
                cell[a] = bound + (0.3 if low else -0.3)
                v = np.array([0.11, -0.07, 0.05])
                v[a] = (-0.6 if low else 0.6) * DX / dt
                st = _one(cfg, cell * DX, v)
                x0 = st.x[0].copy()
                orc.step(cfg, st, dt, 1)
                assert st.x[0, a] == bound * DX, (window, a, low, st.x[0, a] / DX, bound)
                assert st.v[0, a] == 0.0
                for b in range(3):
                    if b != a:
                        assert abs(st.v[0, b] - v[b]) < 1e-12
                        assert abs(st.x[0, b] - (x0[b] + dt * st.v[0, b])) < 1e-15
                # the free particle one cell further in is not clamped
                cell[a] += 1.0 if low else -1.0
                st = _one(cfg, cell * DX, v)
                orc.step(cfg, st, dt, 1)
                assert st.v[0, a] != 0.0 and st.x[0, a] != bound * DX


def test_position_clamp_bounds_and_velocity():
    check_position_clamp()


# ---------------------------------------------------------------- stress clamp
def _p2g_momentum(cfg, F, vol=2e-4, dt=1e-4):
    st = _one(cfg, np.full(3, 8.3 * DX), F=F, vol=vol)
    gm, gp = orc.p2g(cfg, st, dt)
    return gp


def check_stress_clamp():
    cfg = _cfg(youngs=1.0e4, poisson=0.3)
    nu = float(np.float32(0.3))  # the scene holds nu in f32
    mu = 1.0e4 / (2 * (1 + nu))
    lam = 1.0e4 * nu / ((1 + nu) * (1 - 2 * nu))
    R, Q = _rot(1), _rot(2)
    pairs = [((0.05, 1.1, 0.9), (0.1, 1.1, 0.9)), ((0.02, 0.05, 1.3), (0.1, 0.1, 1.3)),
             ((12.0, 1.0, 0.8), (10.0, 1.0, 0.8)), ((1.0, 30.0, 0.01), (1.0, 10.0, 0.1))]
    for s_out, s_in in pairs:
        pa = _p2g_momentum(cfg, R @ np.diag(s_out) @ Q.T)
        pb = _p2g_momentum(cfg, R @ np.diag(s_in) @ Q.T)
        assert np.abs(pa - pb).max() <= 1e-9 * np.abs(pb).max(), (s_out, s_in)
    # just inside the clamp the stress differs
    pa = _p2g_momentum(cfg, R @ np.diag((0.15, 1.1, 0.9)) @ Q.T)
    pb = _p2g_momentum(cfg, R @ np.diag((0.1, 1.1, 0.9)) @ Q.T)
    assert np.abs(pa - pb).max() > 1e-3 * np.abs(pb).max()
    # diagonal F: the scatter's momentum is w_i (4/dx^2)(-dt vol) tau (x_i - x_p)
    # with the fixed-corotated Kirchhoff stress of the clamped singular values
    dt, vol = 1e-4, 2e-4
    xp = 8.3
    w = orc.quad_weights(xp - np.floor(xp - 0.5))
    base = int(np.floor(xp - 0.5))
    for s in ((0.05, 1.2, 0.9), (1.1, 0.95, 1.05), (11.0, 1.0, 0.5)):
        sc = np.clip(s, 0.1, 10.0)
        J = np.prod(sc)
        tau = 2 * mu * (sc - 1) * sc + lam * (J - 1) * J
        gp = _p2g_momentum(cfg, np.diag(s), vol, dt).reshape(G, G, G, 3)  # [z, y, x]
        for i, j, k in ((0, 0, 0), (2, 1, 0), (1, 2, 2), (0, 1, 2)):
            d = (np.array([i, j, k]) + base - xp) * DX
            wt = w[i] * w[j] * w[k]
            exp = wt * 4 / DX ** 2 * (-dt * vol) * tau * d
            got = gp[base + k, base + j, base + i]
            assert np.allclose(got, exp, rtol=1e-10, atol=1e-18), (s, i, j, k, got, exp)


def test_stress_sigma_clamp():
    check_stress_clamp()


# ---------------------------------------------------------------- plasticity clamp
def check_plastic_clamp():
    cfg = _cfg()
    from molla_amd.config import material_row
    out, jp = orc.plastic_project(cfg, material_row(Viscous(0.0, 0.0)), [0.01, 1.0, 30.0], 0.0)
    assert np.allclose(out, [0.05, 1.0, 20.0], rtol=0, atol=1e-15) and jp == 0.0
    R, Q = _rot(3), _rot(4)
    for s, exp in (((0.01, 1.0, 30.0), (0.05, 1.0, 20.0)), ((0.06, 0.5, 19.0), (0.06, 0.5, 19.0))):
        st = _one(cfg, np.full(3, 8.3 * DX), F=R @ np.diag(s) @ Q.T)  # vol 0: stress free
        orc.step(cfg, st, 1e-4, 1)
        assert np.abs(st.F[0] - R @ np.diag(exp) @ Q.T).max() < 1e-12, (s, st.F[0])


def test_plasticity_sigma_clamp():
    check_plastic_clamp()


# ---------------------------------------------------------------- fluid
def check_fluid():
    bulk = 5.0e3
    cfg = _cfg(rheology=Fluid(bulk))
    R, Q = _rot(5), _rot(6)
    for s, s_exp in (((0.8, 1.0, 1.25), 1.0), ((2.0, 1.0, 1.0), 2.0 ** (1 / 3)),
                     ((0.5, 0.5, 0.5), 0.5), ((1e-5, 1.0, 1.0), 0.05), ((40.0, 40.0, 40.0), 20.0)):
        st = _one(cfg, np.full(3, 8.3 * DX), F=R @ np.diag(s) @ Q.T)
        orc.step(cfg, st, 1e-4, 1)
        assert np.abs(st.F[0] - s_exp * np.eye(3)).max() < 1e-12, (s, st.F[0])
        if 0.05 < s_exp < 20.0:
            assert abs(np.linalg.det(st.F[0]) - np.prod(s)) < 1e-12
    # pressure of F = s I: tau = bulk (J - 1) J on every axis (mu = 0)
    dt, vol, xp = 1e-4, 2e-4, 8.3
    w = orc.quad_weights(xp - np.floor(xp - 0.5))
    base = int(np.floor(xp - 0.5))
    for s in (0.9, 1.07):
        J = s ** 3
        tau = bulk * (J - 1) * J
        gp = _p2g_momentum(cfg, s * np.eye(3), vol, dt).reshape(G, G, G, 3)
        for i, j, k in ((0, 0, 0), (2, 1, 0), (1, 2, 2)):
            d = (np.array([i, j, k]) + base - xp) * DX
            exp = w[i] * w[j] * w[k] * 4 / DX ** 2 * (-dt * vol) * tau * d
            assert np.allclose(gp[base + k, base + j, base + i], exp, rtol=1e-10, atol=1e-18), (s, i, j, k)


def test_fluid_reset_and_pressure():
    check_fluid()


# ---------------------------------------------------------------- planted mistakes
@pytest.mark.parametrize("mutation, check", [
    (1, check_wall_cells), (2, check_wall_cells), (3, check_wall_cells),
    (1, check_wall_particles), (2, check_wall_particles), (3, check_wall_particles),
    (4, check_position_clamp), (5, check_position_clamp),
    (6, check_stress_clamp), (9, check_stress_clamp),
    (7, check_plastic_clamp),
    (8, check_fluid), (9, check_fluid),
])
def test_planted_mistake_fails_its_check(mutation, check):
    with planted(mutation):
        with pytest.raises(AssertionError):
            check()
