"""Rigid bodies on the grid through the C ABI against the FP64 oracle
(rigid_imprint, the rigid mask, the legacy shell force and the wrench
readback, mpm.wgsl:927-1226), after the reference's GPU tests:

* prescribed solids (mpm_prescribed_solids_tests.rs:247-318): a rotating box
  in a 65-particle block moves the sand, every P2G mode agrees to 2e-5, and
  removing the solids empties the wrench list;
* legacy spheres (tests/mpm_active_tiles_gpu.rs:586-635): 129 particles of
  alternating sand and viscous material under a moving sphere and a distant
  one, four steps with the spheres removed in the last, explicit and
  implicit, compared per step with the oracle's particles and wrenches;
* graph replays with bodies equal eager stepping;
* validation: bad poses, implicit stepping, dt = 0, capacity, and a restore
  with bodies attached are refused and change nothing.

Bounds: positions, velocities and wrenches are compared with the FP64 oracle
at a multiple of the f32 oracle's own error, with floors.
"""
import math

import numpy as np
import pytest

import oracle as orc
from molla_amd import Implicit, MpmConfig, MpmError, MpmSimulation, Sand, Viscous
from molla_amd.rigid import MpmPrescribedSolid, SolidBox, SolidCylinder, SolidSphere, body_rows
from tests.gpu_util import log, rel_err

pytestmark = pytest.mark.gpu


def solid_config():
    # mpm_prescribed_solids_tests.rs:14-25
    return MpmConfig.sand().copy(grid_size=16, grid_origin=(-1.0, -1.0, -1.0), cell_size=0.125,
                                 max_particles=129, max_rigid=2, youngs=1000.0, gravity_y=-2.0)


def solid_points():
    return np.array([[-0.125 + (i % 5) * 0.0625, -0.125 + (i // 5 % 5) * 0.0625, (i // 25) * 0.0625]
                     for i in range(65)], np.float32)


BOX = MpmPrescribedSolid(shape=SolidBox((0.3, 0.3, 0.1)), linear_velocity=(0.01, 0.0, 0.02),
                         angular_velocity=(0.0, 0.25, 0.0))


def _oracle_run(cfg, x, v, steps, dt, substeps, mat=None):
    """steps: list of body-row arrays, one per step_n call; returns the f64 and
    f32 oracle states and wrenches after every call."""
    out = []
    for dtype in (np.float64, np.float32):
        n = len(x)
        st = orc.State(x, v, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume),
                       mat=mat, dtype=dtype)
        hist = []
        for rows in steps:
            w, _, _ = orc.step_rigid(cfg, st, dt, substeps, rows)
            hist.append((st.x.copy(), st.v.copy(), w.copy()))
        out.append(hist)
    return out


def _bound(e64, e32, floor):
    return max(8 * e32, floor)


def test_prescribed_solid_moves_sand_like_the_oracle_and_modes_agree():
    cfg = solid_config()
    x = solid_points()
    results = {}
    for mode in ("fused", "global_atomic", "deterministic"):
        sim = MpmSimulation(cfg)
        sim.set_p2g_mode(mode)
        sim.set_particles(x)
        sim.set_prescribed_solids([BOX])
        sim.step_n(1e-4, 8)
        w = sim.read_rigid_wrenches()
        xv = (sim.read_positions(), sim.read_velocities(), np.array(w[0][0] + w[0][1]))
        assert np.abs(xv[1]).max() > 1e-4
        sim.set_prescribed_solids([])
        sim.step_n(1e-4, 2)
        assert sim.read_rigid_wrenches() == []
        results[mode] = xv + (sim.read_positions(), sim.read_velocities())
        sim.close()
    (o64, o32) = _oracle_run(cfg, x, None, [body_rows(solids=[BOX])], 1e-4, 8)
    gx, gv, gw = results["fused"][:3]
    ox, ov, ow = o64[0]
    sx, sv, sw = o32[0]
    ex, ev = rel_err(gx, ox), rel_err(gv, ov)
    ew = float(np.abs(gw - ow[0]).max() / max(np.abs(ow[0]).max(), 1e-12))
    ew32 = float(np.abs(sw[0] - ow[0]).max() / max(np.abs(ow[0]).max(), 1e-12))
    log(f"prescribed box: x {ex:.2e} v {ev:.2e} wrench {ew:.2e} (f32 oracle {rel_err(sx, ox):.2e} "
        f"{rel_err(sv, ov):.2e} {ew32:.2e}), |F| {np.abs(ow[0][:3]).max():.3e}")
    assert ex <= _bound(ex, rel_err(sx, ox), 1e-6)
    assert ev <= _bound(ev, rel_err(sv, ov), 1e-4)
    assert ew <= _bound(ew, ew32, 1e-3)
    for mode in ("global_atomic", "deterministic"):
        for a, b in zip(results["fused"], results[mode]):
            assert np.abs(np.asarray(a) - np.asarray(b)).max() <= 2e-5 * max(1.0, float(np.abs(a).max())), mode


def _legacy_setup(size, implicit):
    # tests/mpm_active_tiles_gpu.rs:454-468, 470-481, 591-618
    cfg = MpmConfig.sand().copy(grid_size=size, grid_origin=(-1.0, -1.0, -1.0), cell_size=0.0625,
                                max_particles=512, max_rigid=4, youngs=3000.0,
                                time_step=Implicit(3) if implicit else MpmConfig.sand().time_step,
                                materials=[Sand(friction_angle=0.5), Viscous(0.1, 0.2)])
    i = np.arange(129)
    x = np.stack([0.24 + (i % 7).astype(np.float32) * np.float32(0.021) + np.float32(0.00013),
                  0.1 + (i // 7 % 7).astype(np.float32) * np.float32(0.021) + np.float32(0.00031),
                  -0.2 + (i // 49).astype(np.float32) * np.float32(0.021) + np.float32(0.00047)], 1)
    x = x.astype(np.float32)
    v = np.tile(np.array([0.3, -0.02, 0.1], np.float32), (len(x), 1))
    mat = (i % 2).astype(np.uint32)
    steps = []
    for step in range(4):
        steps.append([] if step == 3 else [((0.30 + 0.01 * step, 0.18, -0.16), 0.065, (0.1, -0.01, 0.0)),
                                           ((-0.85, -0.8, -0.85), 0.11, (0.0, 0.0, 0.0))])
    return cfg, x, v, mat, steps


@pytest.mark.parametrize("size,implicit,mode", [(32, False, "fused"), (35, True, "fused"),
                                                (32, True, "global_atomic")])
def test_legacy_spheres_match_oracle_over_moving_steps(size, implicit, mode):
    cfg, x, v, mat, steps = _legacy_setup(size, implicit)
    sim = MpmSimulation(cfg)
    sim.set_p2g_mode(mode)
    sim.set_particles(x, v)
    sim.set_particle_materials(mat)
    gpu = []
    for spheres in steps:
        sim.set_rigid_spheres(spheres)
        sim.step_n(1e-4, 2)
        w = sim.read_rigid_wrenches()
        gpu.append((sim.read_positions(), sim.read_velocities(),
                    np.array([list(f) + list(t) for f, t in w]).reshape(-1, 6)))
    sim.close()
    o64, o32 = _oracle_run(cfg, x, v, [body_rows(spheres=s) for s in steps], 1e-4, 2, mat=mat)
    for k, ((gx, gv, gw), (ox, ov, ow), (sx, sv, sw)) in enumerate(zip(gpu, o64, o32)):
        ex, ev = rel_err(gx, ox), rel_err(gv, ov)
        assert gw.shape[0] == len(steps[k])
        scale = max(np.abs(ow).max(), 1e-9) if ow.size else 1.0
        ew = float(np.abs(gw - ow).max() / scale) if ow.size else 0.0
        ew32 = float(np.abs(sw - ow).max() / scale) if ow.size else 0.0
        log(f"size={size} implicit={implicit} {mode} step {k}: x {ex:.2e} v {ev:.2e} wrench {ew:.2e} "
            f"(f32 oracle {rel_err(sx, ox):.2e} {rel_err(sv, ov):.2e} {ew32:.2e})")
        assert ex <= _bound(ex, rel_err(sx, ox), 2e-6)
        assert ev <= _bound(ev, rel_err(sv, ov), 1e-3)
        assert ew <= _bound(ew, ew32, 1e-3)
        if k < 3:
            assert np.abs(ow[0]).max() > 0 and np.all(ow[1] == 0)  # the far sphere touches no sand
            assert np.all(gw[1] == 0)


def test_graph_replay_with_bodies_equals_eager():
    cfg = solid_config()
    x = solid_points()
    sims = [MpmSimulation(cfg) for _ in range(2)]
    for s in sims:
        s.set_p2g_mode("deterministic")
        s.set_particles(x)
        s.set_prescribed_solids([BOX])
    for rep in range(3):
        sims[0].step_n(1e-4, 4)
        sims[1].step_n_graph(1e-4, 4)
        assert np.array_equal(sims[0].read_positions(), sims[1].read_positions()), rep
        # f64 atomics: the summation order of the reaction is not fixed
        assert np.allclose(np.array(sims[0].read_rigid_wrenches()), np.array(sims[1].read_rigid_wrenches()),
                           rtol=1e-6, atol=1e-9)
    for s in sims:
        s.close()


def test_rigid_validation_refuses_and_changes_nothing():
    cfg = solid_config()
    x = solid_points()
    sim = MpmSimulation(cfg)
    sim.set_p2g_mode("deterministic")
    sim.set_particles(x)
    sim.set_prescribed_solids([BOX])
    bad = [BOX.replace(orientation=(0.0, 0.0, 0.0, 0.0)), BOX.replace(angular_velocity=(0.0, math.nan, 0.0)),
           BOX.replace(center=(math.inf, 0.0, 0.0)), BOX.replace(shape=SolidCylinder(1.0, 0.0)),
           BOX.replace(shape=SolidBox((1.0, -1.0, 1.0))), BOX.replace(orientation=(0.0, 0.0, 0.0, 1.01))]
    for b in bad:
        with pytest.raises(MpmError):
            sim.set_prescribed_solids([b])
    with pytest.raises(MpmError):
        sim.set_prescribed_solids([BOX, BOX, BOX])  # max_rigid 2
    ref = MpmSimulation(cfg)
    ref.set_p2g_mode("deterministic")
    ref.set_particles(x)
    ref.set_prescribed_solids([BOX])
    for s in (sim, ref):
        s.step_n(1e-4, 3)
    assert np.array_equal(sim.read_positions(), ref.read_positions())
    with pytest.raises(MpmError):
        sim.step_n(0.0, 1)  # prescribed solids need dt > 0
    sim.set_time_step(Implicit(2))
    with pytest.raises(MpmError):
        sim.step_n(1e-4, 1)
    with pytest.raises(MpmError):
        sim.set_prescribed_solids([BOX])
    sim.set_time_step(MpmConfig.sand().time_step)
    # a restore needs no attached bodies (validate_soil_restore_origin)
    ck = sim.checkpoint_soil()
    with pytest.raises(MpmError):
        sim.restore_soil(ck)
    sim.set_prescribed_solids([])
    sim.restore_soil(ck)
    sim.close()
    ref.close()
