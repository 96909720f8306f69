"""The sparse grid (a pool of 4^3 tiles, mpm_create_sparse) through the C ABI,
after the reference's GPU tests (tests/mpm_sparse_grid_gpu.rs):

* sparse equals dense within 2e-5 in positions, velocities and rigid
  wrenches, with two materials, two legacy spheres, grid edges and five
  reseeded windows, explicit and implicit, fused and global-atomic P2G
  (:72-138);
* continuous transport through a 64-tile pool keeps the stored mass and the
  momentum (:140-201);
* an undersized pool latches the overflow, freezes the particles and the
  gather count until a reset (:203-238);
* growth keeps the particles' history (:298-...);
* the marked tile count equals the oracle's restatement of the reference's
  marking rule (oracle.marked_tiles), and large domains store only the pool.
"""
import numpy as np
import pytest

import oracle as orc
from molla_amd import Implicit, MpmConfig, MpmError, MpmSimulation, Sand, Viscous
from tests.gpu_util import log

pytestmark = pytest.mark.gpu


def config(size, implicit):
    # mpm_sparse_grid_gpu.rs:6-23
    vol = 0.021 ** 3
    cfg = MpmConfig.sand().copy(grid_size=size, grid_origin=(-1.0, -1.0, -1.0), cell_size=0.0625, max_particles=512,
                                max_rigid=4, youngs=3000.0, particle_volume=vol, particle_mass=vol * 1600.0)
    return cfg.copy(time_step=Implicit(3)) if implicit else cfg


def particles(n, offset):
    # mpm_sparse_grid_gpu.rs:25-35
    return np.array([[offset[0] + (i % 7) * 0.021 + 0.00013, offset[1] + (i // 7 % 7) * 0.021 + 0.00031,
                      offset[2] + (i // 49) * 0.021 + 0.00047] for i in range(n)], np.float32)


def compare(dense, sparse):
    for name in ("position", "velocity"):
        a, b = dense.read_field(name), sparse.read_field(name)
        assert a.shape == b.shape
        assert np.abs(a - b).max(initial=0.0) <= 2e-5, name
    wa, wb = dense.read_rigid_wrenches(), sparse.read_rigid_wrenches()
    assert len(wa) == len(wb)
    for (fa, ta), (fb, tb) in zip(wa, wb):
        for a, b in zip((*fa, *ta), (*fb, *tb)):
            assert abs(a - b) <= 2e-5 * max(abs(a), 1.0), (a, b)


@pytest.mark.parametrize("size,implicit,mode", [(35, True, "fused"), (32, False, "fused"),
                                                (64, True, "global_atomic")])
def test_sparse_pool_matches_dense_materials_rigids_edges_and_reseeding(size, implicit, mode):
    cfg = config(size, implicit)
    dense = MpmSimulation(cfg)
    sparse = MpmSimulation.new_sparse(cfg, 64)
    assert sparse.active_grid_enabled()
    with pytest.raises(MpmError):
        sparse.read_grid_mass_momentum()
    with pytest.raises(MpmError):
        sparse.set_p2g_mode("deterministic")
    windows = [(129, [0.24, 0.1, -0.2]), (97, [-0.75, -0.7, -0.65]), (1, [1.17, 1.17, 1.17]),
               (33, [-1.01, -0.95, -0.97]), (129, [0.24, 0.1, -0.2])]
    for window, (count, offset) in enumerate(windows):
        points = particles(count, offset)
        log(f"sparse vs dense: size {size} implicit {implicit} {mode} window {window} ({count} particles)")
        for sim in (dense, sparse):
            sim.set_p2g_mode(mode)
            sim.set_particles(np.zeros((0, 3), np.float32))
            sim.set_particles(points, np.tile(np.array([0.3, -0.02, 0.1], np.float32), (count, 1)))
            sim.set_materials([Sand(0.5), Viscous(0.1, 0.2)])
            sim.set_particle_materials((np.arange(count) % 2).astype(np.uint32))
            sim.set_rigid_spheres([((0.30 + 0.01 * window, 0.18, -0.16), 0.065, (0.1, -0.01, 0.0)),
                                   ((-0.85, -0.8, -0.85), 0.11, (0.0, 0.0, 0.0))])
            sim.step_n(1e-4, 2)
        compare(dense, sparse)
        st = sparse.read_active_grid_stats()
        assert st["completed_gathers"] == (window + 1) * 2, st
        assert 0 < st["active_tiles"] <= 64, st
    log(f"  grid storage: dense {dense.grid_storage_bytes()} B, sparse {sparse.grid_storage_bytes()} B")
    assert sparse.grid_storage_bytes() < dense.grid_storage_bytes()
    dense.close()
    sparse.close()


def test_marked_tiles_equal_the_reference_rule():
    # one substep per call: the pool's slots are the tiles marked by the
    # particles at the start of the substep and by the spheres
    cfg = config(64, False)
    rng = np.random.default_rng(3)
    x = np.concatenate([particles(129, [0.24, 0.1, -0.2]), rng.uniform(-0.9, 0.9, (40, 3)).astype(np.float32)])
    spheres = [((0.3, 0.18, -0.16), 0.065, (0.1, -0.01, 0.0)), ((-0.85, -0.8, -0.85), 0.11, (0.0, 0.0, 0.0))]
    for mode in ("fused", "global_atomic"):
        sim = MpmSimulation.new_sparse(cfg, 4096)
        sim.set_p2g_mode(mode)
        sim.set_particles(x)
        sim.set_rigid_spheres(spheres)
        for call in range(3):
            x0 = sim.read_positions()
            sim.step_n(1e-4, 1)
            want = orc.marked_tiles(cfg, x0, [(c, r) for c, r, _ in spheres])
            got = np.sort(sim.read_active_tile_ids())
            log(f"  {mode} call {call}: {len(got)} marked tiles, oracle {len(want)}")
            assert np.array_equal(got, want), (mode, call)
            assert sim.read_status()["active_tiles"] == len(want)
        sim.close()


def test_sparse_pool_recycles_slots_during_continuous_transport():
    # mpm_sparse_grid_gpu.rs:140-201
    for implicit in (False, True):
        cfg = config(64, implicit).copy(youngs=0.0, gravity_y=-0.5)
        points = particles(129, [-0.6, 0.1, 0.1])
        mass = len(points) * float(cfg.particle_mass)
        dense = MpmSimulation(cfg)
        sparse = MpmSimulation.new_sparse(cfg, 64)
        for sim in (dense, sparse):
            sim.set_particles(points, np.tile(np.array([4.0, 0.0, 0.0], np.float32), (len(points), 1)))
        for window in range(1, 9):
            for sim in (dense, sparse):
                sim.step_n(1.0 / 512.0, 16)
            compare(dense, sparse)
            gm, _ = sparse.read_grid_storage_mass_momentum()
            assert gm.shape == (64 * 64,)
            assert abs(float(np.sum(gm, dtype=np.float64)) - mass) <= 2e-5 * mass
            v = sparse.read_velocities().astype(np.float64)
            mom = v.sum(axis=0) * float(cfg.particle_mass)
            want = [mass * 4.0, mass * cfg.gravity_y * window * 16 / 512.0, 0.0]
            for a in range(3):
                assert abs(mom[a] - want[a]) <= 2e-5 * max(abs(want[a]), 1.0), (a, mom, want)
            assert sparse.read_active_grid_stats()["completed_gathers"] == window * 16
        assert np.all(sparse.read_positions()[:, 0] - points[:, 0] > 0.99)
        dense.close()
        sparse.close()


def test_sparse_pool_overflow_freezes_particles_until_explicit_reset():
    # mpm_sparse_grid_gpu.rs:203-238
    cfg = config(64, True)
    for cap in (0, 4097):
        with pytest.raises(MpmError):
            MpmSimulation.new_sparse(cfg, cap)
    points = np.array([[0.1] * 3, [1.1] * 3], np.float32)
    vel = np.array([[0.3, 0.0, 0.0]] * 2, np.float32)
    for mode in ("global_atomic", "fused"):
        sim = MpmSimulation.new_sparse(cfg, 8)
        sim.set_p2g_mode(mode)
        sim.set_particles(points, vel)
        for _ in range(2):
            sim.step_n(1e-4, 3)
            st = sim.read_status()
            assert st["overflow"] and st["completed_gathers"] == 0, st
            with pytest.raises(MpmError):
                sim.read_active_grid_stats()
            with pytest.raises(MpmError):
                sim.read_active_tile_ids()
            assert np.array_equal(sim.read_positions(), points)
            assert np.array_equal(sim.read_velocities(), vel)
        sim.set_particles(points[:1])
        sim.reset_status()
        sim.step_n(1e-4, 2)
        st = sim.read_active_grid_stats()
        assert st["completed_gathers"] == 2 and st["active_tiles"] == 8 and not st["overflow"], st
        sim.close()


def test_sparse_pool_growth_preserves_history():
    # mpm_sparse_grid_gpu.rs:298-...: an 8-tile pool overflows on the third
    # substep of two particles moving at 4 m/s; growing it keeps the two
    # completed substeps, and stepping on matches the dense grid
    cfg = config(64, False).copy(youngs=0.0, gravity_y=0.0)
    points = np.array([[0.1, 0.1, 0.1], [0.21, 0.1, 0.1]], np.float32)
    vel = np.array([[4.0, 0.0, 0.0]] * 2, np.float32)
    dense = MpmSimulation(cfg)
    sparse = MpmSimulation.new_sparse(cfg, 8)
    for sim in (dense, sparse):
        sim.set_particles(points, vel)
    dense.step_n(0.02, 2)
    sparse.step_n(0.02, 3)
    st = sparse.read_status()
    assert st["overflow"] and st["completed_gathers"] == 2, st
    compare(dense, sparse)  # the two completed substeps
    sparse.reserve_sparse_tiles(32)
    sparse.reset_status()
    for sim in (dense, sparse):
        sim.step_n(0.02, 3)
    compare(dense, sparse)
    assert sparse.read_active_grid_stats()["completed_gathers"] == 3
    assert 0 < sparse.read_status()["active_tiles"] <= 32
    dense.close()
    sparse.close()


@pytest.mark.parametrize("size", [32, 96, 512, 1536])
def test_sparse_pool_bounds_storage_in_large_domains(size):
    # mpm_sparse_grid_gpu.rs:240-296: 129 particles mark 8 tiles whatever the
    # domain; the storage is the pool's, and a sphere covering the domain
    # overflows it without moving the particles
    cfg = config(size, True)
    sim = MpmSimulation.new_sparse(cfg, 64)
    sim.set_particles(particles(129, [0.1, 0.1, 0.1]))
    sim.step_n(1e-4, 1)
    sim.step_n(1e-4, 4)
    st = sim.read_active_grid_stats()
    assert st["active_tiles"] == 8 and st["completed_gathers"] == 5, st
    W = (size + 3) // 4
    assert sim.grid_storage_bytes() == 64 * 64 * 32 + 64 * 64 * 16 + W ** 3 * 8 + 64 * 4
    if size == 1536:
        remote = particles(129, [93.0, 93.0, 93.0])
        sim.set_particles(remote)
        sim.step_n(1e-4, 2)
        st = sim.read_active_grid_stats()
        assert not st["overflow"] and 0 < st["active_tiles"] <= 64
        pos = sim.read_positions()
        assert np.all(np.isfinite(pos)) and np.all(pos > 92.0)
        v = sim.read_velocities()
        assert np.all(np.abs(v[:, 0]) <= 2e-5) and np.all(np.abs(v[:, 2]) <= 2e-5)
        assert np.all(np.abs(v[:, 1] - 2.0 * cfg.gravity_y * 1e-4) <= 2e-5)
        gm, _ = sim.read_grid_storage_mass_momentum()
        assert abs(float(np.sum(gm, dtype=np.float64)) - 129 * cfg.particle_mass) <= 2e-5 * 129 * cfg.particle_mass
        sim.set_rigid_spheres([((0.0, 0.0, 0.0), 1000.0, (0.0, 0.0, 0.0))])
        sim.step_n(1e-4, 1)
        f = sim.read_status()
        assert f["overflow"] and f["completed_gathers"] == st["completed_gathers"], f
        assert np.array_equal(sim.read_positions(), pos)
    sim.close()
