"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer work (tile keys, sorted ids) must be bit-exact.  Floating-point
state is compared with the FP64 oracle; the tolerance of each comparison is
stated where it is asserted.  For whole-step comparisons the bound is
``max(8 * e32, floor)`` where ``e32`` is the error of the oracle's own f32
restatement against FP64 on the same input: the HIP path must be as accurate
as a plain f32 implementation of the reference arithmetic.
"""
import math

import numpy as np
import pytest

import oracle as orc
from molla_amd import Fluid, Implicit, MpmConfig, MpmError, MpmSimulation, Sand, Snow, Viscous, workloads
from tests.fixtures import fill_box, order_config, points
from tests.gpu_util import log, oracle_state, rel_err, run_gpu

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- ordering
@pytest.mark.parametrize("size", [35, 1536])
def test_sort_keys_and_ids_bit_exact(size):
    # src/mpm_particle_order_tests.rs:66-150 with count in [513, 257, 1, 0, 63]
    cfg = order_config(size)
    for count in [513, 257, 1, 63]:
        pos = points(count)
        if count > 1:
            pos[0] = [-1.0625] * 3
            pos[1] = [size * cfg.cell_size] * 3
        log(f"sort keys size={size} count={count}")
        sim = MpmSimulation(cfg, max_particles=513)
        sim.set_particles(pos)
        keys, ids = sim.sort_keys_debug(capacity=513)
        ek, ei = orc.sorted_tile_keys(cfg, pos, capacity=513)
        assert np.array_equal(keys, ek)
        assert np.array_equal(ids, ei)
        sim.close()


def test_sort_keys_large_random_bit_exact():
    cfg = MpmConfig.sand().copy(grid_size=256, cell_size=1 / 256, grid_origin=(0, 0, 0))
    rng = np.random.default_rng(5)
    x = rng.random((300_000, 3)).astype(np.float32)
    x[:1000] = x[1000:2000]  # ties
    log("sort keys 300k random")
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x)
    keys, ids = sim.sort_keys_debug()
    ek, ei = orc.sorted_tile_keys(cfg, x)
    assert np.array_equal(keys, ek) and np.array_equal(ids, ei)


# ---------------------------------------------------------------- P2G / grid
def test_grid_mass_matches_particle_mass_and_oracle_p2g():
    # tests/mpm_gpu.rs:840-885 (rel < 1e-4, |p| < 1e-6), plus cellwise vs the oracle
    cfg = MpmConfig.sand().copy(gravity_y=0.0, grid_origin=(-1.6, -1.6, -1.6), grid_size=32,
                                cell_size=0.1, particle_mass=2.5e-3, boundary_friction=1.0)
    x = fill_box([-0.35] * 3, [0.35] * 3, 7, 19)
    log("grid mass")
    sim, _ = run_gpu(cfg, x, dt=1 / 240, substeps=1)
    m, p = sim.read_grid_mass_momentum()
    exp = cfg.particle_mass * len(x)
    assert abs(m.astype(np.float64).sum() - exp) / exp < 1e-4
    assert np.abs(p).max() < 1e-6
    gm, gp = orc.p2g(cfg, oracle_state(cfg, x), 1 / 240)
    # cellwise against FP64, bounded by 4x what a sequential f32 scatter (the
    # reference's accumulation) reaches on the same input
    m32, _ = orc.step(cfg, oracle_state(cfg, x, dtype=np.float32), 1 / 240, 1, want_grid=True)
    e, e32 = rel_err(m, gm), rel_err(m32, gm)
    log(f"  cell mass err {e:.2e} (f32 scatter {e32:.2e})")
    assert e < max(4 * e32, 1e-6)


def test_p2g_with_stress_matches_oracle():
    cfg = MpmConfig.sand().copy(youngs=2.0e4, rheology=Viscous(0, 0))
    rng = np.random.default_rng(7)
    x = fill_box([-0.2, 0.2, -0.2], [0.2, 0.6, 0.2], 12, 9)
    n = len(x)
    F = np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3))
    v = 0.3 * rng.standard_normal((n, 3))
    C = 0.5 * rng.standard_normal((n, 3, 3))
    for mode in ("fused", "global_atomic", "deterministic"):
        log(f"p2g vs oracle mode={mode}")
        sim = MpmSimulation(cfg, max_particles=n)
        sim.set_particles(x, v.astype(np.float32))
        sim.write_field("F", F.astype(np.float32))
        sim.write_field("C", C.astype(np.float32))
        sim.set_p2g_mode(mode)
        sim.step_n(1e-4, 1)
        m, p = sim.read_grid_mass_momentum()
        st = oracle_state(cfg, x, v.astype(np.float32), F=F.astype(np.float32))
        st.C[:] = C.astype(np.float32)
        gm, gp = orc.p2g(cfg, st, 1e-4)
        s32 = oracle_state(cfg, x, v.astype(np.float32), F=F.astype(np.float32), dtype=np.float32)
        s32.C[:] = C.astype(np.float32)
        m32, p32 = orc.step(cfg, s32, 1e-4, 1, want_grid=True)
        em, ep = rel_err(m, gm), rel_err(p, gp)
        em32, ep32 = rel_err(m32, gm), rel_err(p32, gp)
        log(f"  mass err {em:.2e} momentum err {ep:.2e} (f32 scatter {em32:.2e} {ep32:.2e})")
        assert em < max(4 * em32, 1e-6) and ep < max(4 * ep32, 1e-6)


def test_p2g_mixes_fast_and_resting_particles_within_one_chunk():
    # The fused scatter's momentum words are fixed point at 2^-23 of the
    # chunk's largest single contribution bound, so a cell fed only by slow
    # particles next to a fast one (in the same chunk) is rounded to that
    # quantum rather than to f32 of its own value.  The bound per cell is
    # N_contributions x 2^-23 x (the chunk's bound): here one particle at
    # 40 m/s among 1,000 at ~1 mm/s, all in one tile, compared with the FP64
    # P2G cell by cell.  The mass words (2^-40 of the largest particle mass)
    # stay at f32-scatter accuracy.
    cfg = MpmConfig.sand().copy(youngs=2.0e4, rheology=Viscous(0, 0), gravity_y=0.0, grid_origin=(0.0, 0.0, 0.0),
                                grid_size=32, cell_size=1 / 32)
    rng = np.random.default_rng(17)
    n = 1000
    x = (np.array([0.5, 0.5, 0.5]) + (rng.random((n, 3)) - 0.5) * (4 / 32)).astype(np.float32)  # one 4^3 tile
    v = (1e-3 * rng.standard_normal((n, 3))).astype(np.float32)
    v[0] = (40.0, 0.0, 0.0)
    sim = MpmSimulation(cfg, max_particles=n)
    sim.set_particles(x, v)
    sim.step_n(1e-5, 1)
    m, p = sim.read_grid_mass_momentum()
    st = sim.read_status()
    sim.close()
    assert st["escapes"] == 0 and st["fallback_chunks"] == 0
    gm, gp = orc.p2g(cfg, oracle_state(cfg, x, v), 1e-5)
    m32, p32 = orc.step(cfg, oracle_state(cfg, x, v, dtype=np.float32), 1e-5, 1, want_grid=True)
    em, em32 = rel_err(m, gm), rel_err(m32, gm)
    log(f"  cell mass err {em:.2e} (f32 scatter {em32:.2e})")
    assert em < max(4 * em32, 1e-6)
    # momentum: per cell within the fixed-point bound; the bound on any one
    # contribution is |m v| + 1.5 dx |A| <= 40 m (no stress, C = 0)
    bound = 40.0 * cfg.particle_mass
    quantum = bound * 2.0 ** -23
    err = np.abs(p.astype(np.float64) - gp).max()
    log(f"  momentum max abs err {err:.3e} = {err / quantum:.1f} quanta of 2^-23 x the chunk bound "
        f"(f32 scatter {np.abs(p32.astype(np.float64) - gp).max():.3e})")
    # The chunk's true quantum is 2^(ceil(log2 bound) - 23) <= 2 x quantum and
    # each contribution rounds to nearest (<= half of it, unbiased), so a node
    # of N contributions (here up to a few hundred) is off by N/2 true quanta
    # at worst and by about sqrt(N / 12) of them typically: 64 x quantum sits
    # between the two.
    assert err <= 64 * quantum
    # the slow cells' velocities (p / m) within that bound over their own mass
    slow = (gm > 1e-3 * gm.max()) & (np.abs(gp[:, 0]) < 1e-2 * bound)
    dv = np.abs(p[slow].astype(np.float64) / m[slow, None] - gp[slow] / gm[slow, None]).max(axis=1)
    log(f"  slow cells: {int(slow.sum())}, max velocity error {dv.max():.2e} m/s (fast particle 40 m/s)")
    assert slow.sum() > 50 and np.all(dv <= 64 * quantum / gm[slow])


# ---------------------------------------------------------------- whole steps
def test_single_particle_free_fall_is_exact():
    # tests/mpm_freefall_regression_gpu.rs:76-111
    cfg = MpmConfig.snow()
    sim = MpmSimulation(cfg, max_particles=1)
    sim.set_particles(np.array([[0.013, 0.2, 0.017]], np.float32))
    dt = 1.0 / 480.0
    for step in range(20):
        sim.step(dt)
        pos = sim.read_positions()
        vel = sim.read_velocities()
        vy = -9.81 * dt * (step + 1)
        assert abs(vel[0, 1] - vy) < 1e-4
        assert abs(vel[0, 0]) < 1e-5 and abs(vel[0, 2]) < 1e-5
        assert abs(pos[0, 0] - 0.013) < 1e-5 and abs(pos[0, 2] - 0.017) < 1e-5


def _compare_with_oracle(cfg, x, dt, substeps, v=None, mode="fused", fields=None, tag="", launch=None):
    log(f"whole-step parity {tag} mode={mode} n={len(x)} substeps={substeps}")
    _, g = run_gpu(cfg, x, v, dt, substeps, mode=mode, fields=fields, launch=launch)
    kw = {}
    if fields:
        kw = {k: fields[k] for k in ("mass", "volume", "material", "youngs") if k in fields}
    def ostate(dtype):
        return oracle_state(cfg, x, v, dtype=dtype, mass=kw.get("mass"), vol=kw.get("volume"),
                            mat=kw.get("material"), E=kw.get("youngs"))
    o64 = ostate(np.float64)
    o32 = ostate(np.float32)
    orc.step(cfg, o64, dt, substeps)
    orc.step(cfg, o32, dt, substeps)
    errs = {}
    for f, ref, low in (("x", o64.x, o32.x), ("v", o64.v, o32.v), ("F", o64.F, o32.F), ("C", o64.C, o32.C)):
        e = rel_err(g[f], ref)
        e32 = rel_err(low, ref)
        errs[f] = (e, e32)
    log("  " + " ".join(f"{k}: gpu {e:.2e} f32-oracle {e32:.2e}" for k, (e, e32) in errs.items()))
    assert g["status"]["fault_bits"] == 0
    return errs


def _assert_errs(errs, floors):
    for f, (e, e32) in errs.items():
        assert e <= max(8 * e32, floors[f]), (f, e, e32)


def test_elastic_cube_matches_fp64_oracle():
    cfg, x, meta = workloads.elastic_cube(8192, seed=1)
    errs = _compare_with_oracle(cfg, x, meta["dt"], 100, tag="elastic_cube")
    _assert_errs(errs, {"x": 1e-6, "v": 1e-3, "F": 1e-5, "C": 1e-3})


@pytest.mark.parametrize("rheo", ["sand", "snow", "viscous", "fluid"])
def test_rheologies_match_fp64_oracle(rheo):
    r = {"sand": Sand(), "snow": Snow(), "viscous": Viscous(4.0, 8.0), "fluid": Fluid(2.0e3)}[rheo]
    cfg = MpmConfig.sand().copy(rheology=r, youngs=2.0e3 if rheo != "snow" else 1.4e4,
                                poisson=0.3 if rheo != "snow" else 0.2)
    x = fill_box([-0.15, 0.05, -0.15], [0.15, 0.35, 0.15], 12, 31)
    errs = _compare_with_oracle(cfg, x, 1 / 960, 60, tag=rheo)
    _assert_errs(errs, {"x": 1e-5, "v": 5e-3, "F": 1e-4, "C": 5e-3})


def test_p2g_modes_agree_and_deterministic_is_bitwise_repeatable():
    cfg, x, meta = workloads.elastic_cube(8192, seed=2)
    res = {}
    for mode in ("fused", "global_atomic", "deterministic"):
        log(f"mode {mode}")
        _, res[mode] = run_gpu(cfg, x, None, meta["dt"], 50, mode=mode)
    for mode in ("global_atomic", "deterministic"):
        for f in ("x", "v", "F"):
            e = rel_err(res[mode][f], res["fused"][f])
            log(f"  {mode} vs fused {f}: {e:.2e}")
            assert e < {"x": 1e-6, "v": 1e-3, "F": 1e-5}[f]
    _, again = run_gpu(cfg, x, None, meta["dt"], 50, mode="deterministic")
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(again[f], res["deterministic"][f])


def test_sort_interval_and_write_vc_do_not_change_results_beyond_rounding():
    cfg, x, meta = workloads.elastic_cube(8192, seed=3)
    _, base = run_gpu(cfg, x, None, meta["dt"], 40, sort_every=0)
    for kw in ({"sort_every": 1}, {"sort_every": 7}, {"write_vc": False}):
        log(f"variant {kw}")
        _, r = run_gpu(cfg, x, None, meta["dt"], 40, **kw)
        for f in ("x", "v", "F"):
            assert rel_err(r[f], base[f]) < {"x": 1e-6, "v": 1e-3, "F": 1e-5}[f]
        if kw.get("sort_every") == 1:
            # The sort spreads each chunk's particles over cells and moves
            # the sorted tile keys with them; a key left behind would send
            # particles of a slow scene to the global-atomic fallback.
            assert r["status"]["sorts"] >= 2 and r["status"]["escapes"] == 0, r["status"]


def test_graph_replays_match_eager():
    cfg, x, meta = workloads.elastic_cube(4096, seed=4)
    dt = meta["dt"]
    eager = MpmSimulation(cfg, max_particles=len(x))
    graph = MpmSimulation(cfg, max_particles=len(x))
    for s in (eager, graph):
        s.set_particles(x)
        s.set_p2g_mode("deterministic")
    for rep in range(4):
        log(f"graph replay {rep}")
        eager.step_n(dt, 5)
        graph.step_n_graph(dt, 5)
        a, b = eager.read_positions(), graph.read_positions()
        assert np.array_equal(a, b)
        _ = eager.read_field("F")  # eager work between replays


def test_fused_graph_replays_with_sorts_inside_match_eager():
    # The batch number is measured with step_n_graph in the fused mode, whose
    # captured window bakes in where its sorts fall (mpm_command_batch.rs
    # :157-216 is the reference's prepared-step contract).  With a sort every
    # 3 substeps and 8 substeps per call, every replay holds two or three
    # sorts at offsets that change from call to call; the eager handle steps
    # the same calls.  Changing the sort interval and writing a per-particle
    # E between calls must not replay a graph captured before.
    cfg, x, meta = workloads.elastic_cube(4096, seed=4)
    dt = meta["dt"]
    eager = MpmSimulation(cfg, max_particles=len(x))
    graph = MpmSimulation(cfg, max_particles=len(x))
    for s in (eager, graph):
        s.set_particles(x)
        s.set_sort_interval(3)
    E = np.full(len(x), 1.5 * cfg.youngs, np.float32)
    for rep in range(9):
        if rep == 4:
            for s in (eager, graph):
                s.set_sort_interval(2)
        if rep == 6:
            for s in (eager, graph):
                s.write_field("youngs", E)
        log(f"fused graph replay {rep}")
        eager.step_n(dt, 8)
        graph.step_n_graph(dt, 8)
        for f in ("x", "v", "F"):
            a = eager.read_field({"x": "position", "v": "velocity", "F": "F"}[f])
            b = graph.read_field({"x": "position", "v": "velocity", "F": "F"}[f])
            assert np.array_equal(a, b), (rep, f, float(np.abs(a - b).max()))
        se, sg = eager.read_status(), graph.read_status()
        assert se["escapes"] == 0 and sg["sorts"] == se["sorts"], (se, sg)
    assert se["sorts"] >= 20
    eager.close()
    graph.close()


def test_batched_graph_replays_match_eager_and_oracle():
    # A sample of the 512-scene batch stepped by graph replays (fused mode,
    # sorts inside the window) equals eager stepping bit for bit, and each
    # scene matches its own FP64 oracle run.
    cfgs, xs, meta = workloads.batched_scenes(8, seed=13, n_lo=1000, n_hi=6000)
    dt, calls, k = meta["dt"], 4, 6
    log(f"batch of {len(cfgs)} scenes, {meta['particles']} particles, grids {meta['grid_min']}-{meta['grid_max']}")
    eager = MpmSimulation(cfgs)
    graph = MpmSimulation(cfgs)
    for s in (eager, graph):
        s.set_particles_batch(xs)
        s.set_sort_interval(4)
    for rep in range(calls):
        eager.step_n(dt, k)
        graph.step_n_graph(dt, k)
        a, b = eager.read_positions(), graph.read_positions()
        assert np.array_equal(a, b), rep
        assert np.array_equal(eager.read_field("F"), graph.read_field("F")), rep
    pos = graph.read_positions()
    st = graph.read_status()
    assert st["sorts"] >= calls * k // 4 and not st["faulted"], st
    off = 0
    for c, x in zip(cfgs, xs):
        o64 = oracle_state(c, x)
        o32 = oracle_state(c, x, dtype=np.float32)
        orc.step(c, o64, dt, calls * k)
        orc.step(c, o32, dt, calls * k)
        e = rel_err(pos[off:off + len(x)], o64.x)
        e32 = rel_err(o32.x, o64.x)
        log(f"scene of {len(x)}: x err {e:.2e} (f32 oracle {e32:.2e})")
        assert e <= max(8 * e32, 1e-6)
        off += len(x)
    eager.close()
    graph.close()


def test_fault_latch_freezes_state():
    cfg = MpmConfig.sand()
    x = fill_box([-0.1, 0.3, -0.1], [0.1, 0.5, 0.1], 4, 2)
    v = np.zeros_like(x)
    v[3, 0] = np.nan
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, v)
    sim.step_n(1 / 240, 3)
    st = sim.read_status()
    assert st["numerical_fault"] and st["faulted"]
    assert np.array_equal(sim.read_positions(), x)
    sim.reset_status()
    assert not sim.read_status()["faulted"]


def test_closed_system_momentum_and_dropped_block():
    # tests/mpm_gpu.rs:781-838 and tests/mpm_freefall_regression_gpu.rs:113-128
    cfg = MpmConfig.sand().copy(gravity_y=0.0, grid_origin=(-1.6, -1.6, -1.6), grid_size=32,
                                cell_size=0.1, boundary_friction=1.0)
    x = fill_box([-0.4] * 3, [0.4] * 3, 8, 7)
    _, r = run_gpu(cfg, x, None, 1 / 240, 60)
    p = np.linalg.norm(r["v"].astype(np.float64).sum(0)) * cfg.particle_mass
    assert p < 0.01 * cfg.particle_mass * len(x)
    cfg = MpmConfig.snow()
    x = fill_box([-0.075, 0.12, -0.075], [0.075, 0.27, 0.075], 6, 22)
    for mode in ("fused", "global_atomic"):
        _, r = run_gpu(cfg, x, None, 1 / 480, 300, mode=mode)
        assert r["x"][:, 1].max() < 0.35 and np.abs(r["x"][:, 0]).max() < 0.25


def test_batched_scenes_match_oracle_per_scene():
    cfgs, xs, meta = workloads.batched_scenes(6, seed=11, n_lo=500, n_hi=4000, g_lo=32, g_hi=48)
    log(f"batch of {len(cfgs)} scenes, {meta['particles']} particles")
    sim = MpmSimulation(cfgs, max_particles=None)
    sim.set_particles_batch(xs)
    dt, steps = meta["dt"], 30
    sim.step_n(dt, steps)
    pos = sim.read_positions()
    off = 0
    for c, x in zip(cfgs, xs):
        o64 = oracle_state(c, x)
        o32 = oracle_state(c, x, dtype=np.float32)
        orc.step(c, o64, dt, steps)
        orc.step(c, o32, dt, steps)
        e = rel_err(pos[off:off + len(x)], o64.x)
        e32 = rel_err(o32.x, o64.x)
        log(f"  scene n={len(x)} G={c.grid_size}: x err {e:.2e} (f32 oracle {e32:.2e})")
        assert e <= max(8 * e32, 1e-6)
        off += len(x)


def test_cpu_diff_model_forward_matches_oracle():
    cfg, d, meta = workloads.diff_scene(4096, seed=6, grid=32)
    x, v = d["x"], d["v"]
    log("cpudiff forward")
    _, g = run_gpu(cfg, x, v, meta["dt"], 16, fields={"youngs": d["E"]})
    st = oracle_state(cfg, x, v, E=d["E"].astype(np.float64))
    orc.diff_forward(cfg, st, meta["dt"], 16, record=False)
    ex, ev, eF = rel_err(g["x"], st.x), rel_err(g["v"], st.v), rel_err(g["F"], st.F)
    log(f"  x {ex:.2e} v {ev:.2e} F {eF:.2e}")
    assert ex < 1e-6 and ev < 1e-3 and eF < 1e-4


def test_cpu_diff_floor_row_on_a_rounding_tie():
    # dx = 1/80 and floor_y = 3 dx: node row 3 sits at 0.0375 in f32 exactly
    # where floor_y rounds to, but below it in the f64 of the same f32 inputs,
    # the oracle's arithmetic.  The floor row must be decided the same way,
    # else a whole layer of nodes changes its boundary rule.
    cfg, d, meta = workloads.diff_scene(2048, seed=7, grid=80)
    x, v = d["x"], d["v"]
    _, g = run_gpu(cfg, x, v, meta["dt"], 32, fields={"youngs": d["E"]})
    st = oracle_state(cfg, x, v, E=d["E"].astype(np.float64))
    orc.diff_forward(cfg, st, meta["dt"], 32, record=False)
    ex, ev = rel_err(g["x"], st.x), rel_err(g["v"], st.v)
    log(f"  floor tie: x {ex:.2e} v {ev:.2e}")
    assert ex < 1e-5 and ev < 1e-3


def test_sparse_scene_with_more_chunks_than_partial_slots_matches_oracle():
    # A few thousand particles scattered over a 128^3 grid: most segments hold
    # a particle or two, so the chunk list outgrows the partial buffer
    # (cap/256 + cap/64 + scenes + 64 slots) and the chunks beyond it scatter
    # with global atomics (status.fallback_chunks).  Results must still match
    # the oracle.
    rng = np.random.default_rng(61)
    n = 3000
    x = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    cfg = MpmConfig.sand().copy(grid_size=128, cell_size=1.0 / 128, grid_origin=(0.0, 0.0, 0.0),
                                rheology=Viscous(0, 0), youngs=1.0e3, max_particles=n)
    dt, steps = 1e-4, 20
    _, g = run_gpu(cfg, x, None, dt, steps)
    st = g["status"]
    log(f"sparse scene: {st}")
    assert st["fallback_chunks"] > 0 and st["fault_bits"] == 0
    o64 = oracle_state(cfg, x)
    o32 = oracle_state(cfg, x, dtype=np.float32)
    orc.step(cfg, o64, dt, steps)
    orc.step(cfg, o32, dt, steps)
    for f, floor in (("x", 1e-6), ("v", 1e-3), ("F", 1e-5)):
        e, e32 = rel_err(g[f], getattr(o64, f)), rel_err(getattr(o32, f), getattr(o64, f))
        log(f"  {f}: gpu {e:.2e} f32-oracle {e32:.2e}")
        assert e <= max(8 * e32, floor), (f, e, e32)


def test_chunk_boxes_beyond_lds_fall_back_to_global_atomics():
    # Lines of particles along x, one per tile row: greedy chunks span six
    # tiles (27 cells with the stencil), and particles drifting in y and z
    # within their sort neighbourhood grow a chunk's box past the 1600-cell
    # LDS box before the next sort, so whole chunks scatter with global
    # atomics (status.fallback_chunks).  Results must still match the oracle.
    rng = np.random.default_rng(29)
    dx = 1.0 / 64
    rows = [(ty, tz) for ty in (5, 7, 9) for tz in (6, 9)]
    xs = np.arange(0.2, 0.8, dx / 2)
    x = np.array([[xx, (4 * ty + 2) * dx, (4 * tz + 2) * dx] for ty, tz in rows for xx in xs], np.float32)
    v = np.zeros_like(x)
    v[:, 1:] = rng.uniform(-3.0, 3.0, (len(x), 2)).astype(np.float32)
    cfg = MpmConfig.sand().copy(gravity_y=0.0, rheology=Viscous(0, 0), youngs=1.0e3, grid_size=64, cell_size=dx,
                                grid_origin=(0.0, 0.0, 0.0), max_particles=len(x))
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, v)
    sim.set_launch_config(particles_per_thread=2, segment_tiles=6, sort_every=40)
    sim.step_n(1e-3, 40)
    st = sim.read_status()
    xg = sim.read_positions()
    sim.close()
    log(f"lines: {len(x)} particles, status {st}")
    assert st["fallback_chunks"] > 0 and st["fault_bits"] == 0
    o64 = oracle_state(cfg, x, v)
    o32 = oracle_state(cfg, x, v, dtype=np.float32)
    orc.step(cfg, o64, 1e-3, 40)
    orc.step(cfg, o32, 1e-3, 40)
    e, e32 = rel_err(xg, o64.x), rel_err(o32.x, o64.x)
    log(f"  x err {e:.2e} (f32 oracle {e32:.2e})")
    assert e <= max(8 * e32, 1e-6)


def test_escaped_particles_fall_back_to_global_atomics():
    # Fast particles and a long sort interval: stencils leave their sorted
    # 3x3x3 tile neighbourhood, so those particle-substeps take the global
    # atomic path.  Results must still match the oracle.
    cfg = MpmConfig.sand().copy(gravity_y=0.0, rheology=Viscous(0, 0), youngs=1.0e3,
                                boundary_friction=1.0, grid_size=48)
    x = fill_box([-0.3, 0.5, -0.3], [-0.15, 0.65, -0.15], 10, 13)
    v = np.tile(np.array([3.0, 1.5, 2.0], np.float32), (len(x), 1))
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, v)
    sim.set_sort_interval(64)
    sim.step_n(2e-3, 60)
    st = sim.read_status()
    log(f"escapes {st['escapes']} sorts {st['sorts']}")
    assert st["escapes"] > 0 and st["fault_bits"] == 0
    o64 = oracle_state(cfg, x, v)
    o32 = oracle_state(cfg, x, v, dtype=np.float32)
    orc.step(cfg, o64, 2e-3, 60)
    orc.step(cfg, o32, 2e-3, 60)
    e, e32 = rel_err(sim.read_positions(), o64.x), rel_err(o32.x, o64.x)
    log(f"  x err {e:.2e} (f32 oracle {e32:.2e})")
    assert e <= max(8 * e32, 1e-6)


@pytest.mark.parametrize("ppt,seg", [(1, 4), (1, 8), (2, 4), (2, 6), (2, 8)])
def test_launch_configurations_match_fp64_oracle_and_repeat_bitwise(ppt, seg):
    # Particles per thread and the chunk rule change the chunks (256 or 512
    # particles; greedy rows of at most 4 or 6 tiles, or aligned 4- and
    # 8-tile segments) and with them each chunk's fixed-point scale, so results
    # differ across configurations at f32 rounding only; one configuration
    # run twice is bitwise identical.  The cube holds 16 particles per cell,
    # so 8-tile segments and full 512-particle chunks occur, and with 11
    # sorts over 60 substeps the sorted chunk lists are rebuilt under the
    # configuration.
    cfg, x, meta = workloads.elastic_cube(8192, seed=6)
    launch = {"particles_per_thread": ppt, "segment_tiles": seg, "sort_every": 6}
    errs = _compare_with_oracle(cfg, x, meta["dt"], 60, tag=f"ppt{ppt} seg{seg}", launch=launch)
    _assert_errs(errs, {"x": 1e-6, "v": 1e-3, "F": 1e-5, "C": 1e-3})
    _, a = run_gpu(cfg, x, None, meta["dt"], 60, launch=launch)
    _, b = run_gpu(cfg, x, None, meta["dt"], 60, launch=launch)
    assert a["status"]["escapes"] == 0 and a["status"]["sorts"] >= 10
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(a[f], b[f]), f


def test_grid_update_launch_shapes_agree_and_repeat():
    # The grid update runs one wave per tile, or for small scenes (at most 64
    # particles per grid-update block) the block's four waves over four
    # ranges of each tile's candidate chunks.  Both shapes match the oracle
    # and each other to f32 rounding, and each repeats bitwise.  512 blocks
    # put the 8,192-particle cube in the split shape, 64 in the other.
    cfg, x, meta = workloads.elastic_cube(8192, seed=8)
    res = {}
    for gb in (64, 512):
        launch = {"grid_blocks": gb, "sort_every": 8}
        _, res[gb] = run_gpu(cfg, x, None, meta["dt"], 40, launch=launch)
        _, again = run_gpu(cfg, x, None, meta["dt"], 40, launch=launch)
        for f in ("x", "v", "C", "F"):
            assert np.array_equal(res[gb][f], again[f]), (gb, f)
    for f in ("x", "v", "F"):
        e = rel_err(res[64][f], res[512][f])
        log(f"  split vs one wave per tile {f}: {e:.2e}")
        assert e < {"x": 1e-6, "v": 1e-3, "F": 1e-5}[f]
    errs = _compare_with_oracle(cfg, x, meta["dt"], 40, tag="split grid update", launch={"grid_blocks": 512})
    _assert_errs(errs, {"x": 1e-6, "v": 1e-3, "F": 1e-5, "C": 1e-3})


def test_launch_configuration_round_trips_and_rejects_bad_values():
    cfg, x, meta = workloads.elastic_cube(1024, seed=7)
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, None)
    sim.set_launch_config(particle_blocks=64, grid_blocks=128, sort_every=9, particles_per_thread=1, segment_tiles=4)
    assert sim.launch_config() == {"particle_blocks": 64, "grid_blocks": 128, "sort_every": 9,
                                   "particles_per_thread": 1, "segment_tiles": 4}
    for bad in ({"particles_per_thread": 0}, {"particles_per_thread": 3}, {"segment_tiles": 5}):
        with pytest.raises(MpmError):
            sim.set_launch_config(**bad)
    assert sim.launch_config()["particles_per_thread"] == 1
    # a change of chunking between steps re-sorts and keeps stepping exact
    # against a handle that ran that configuration throughout
    sim.step_n(meta["dt"], 5)
    sim.set_launch_config(particles_per_thread=2, segment_tiles=8)
    sim.step_n(meta["dt"], 5)
    ref = MpmSimulation(cfg, max_particles=len(x))
    ref.set_particles(x, None)
    ref.set_launch_config(particle_blocks=64, grid_blocks=128, sort_every=9, particles_per_thread=1, segment_tiles=4)
    ref.step_n(meta["dt"], 5)
    ref.set_launch_config(particles_per_thread=2, segment_tiles=8)
    ref.step_n(meta["dt"], 5)
    assert np.array_equal(sim.read_positions(), ref.read_positions())
    assert sim.read_status()["sorts"] == 2
    sim.close()
    ref.close()


def _model_chunk_count(tile_counts: np.ndarray, W: int, cap: int, segment_tiles: int) -> int:
    """Chunks the sorted layout's chunk rule makes for one scene whose tiles
    (x fastest) hold tile_counts particles.  segment_tiles 4 or 6: greedy
    rows (a chunk runs for cap particles or until it would span more than
    that many tiles); 8: aligned segments of eight tiles holding at least
    2 cap particles, else of four, cut every cap particles."""
    n = 0
    for row in range(W * W):
        c = tile_counts[row * W:(row + 1) * W]
        if not c.any():
            continue
        lb = np.concatenate([[0], np.cumsum(c)])
        if segment_tiles in (4, 6):
            s, tt = 0, 0
            while s < lb[W]:
                n += 1
                while lb[tt + 1] <= s:
                    tt += 1
                s = min(s + cap, lb[min(tt + segment_tiles, W)])
        else:
            for g8 in range(0, W, 8):
                n8 = min(8, W - g8)
                tot8 = lb[g8 + n8] - lb[g8]
                if tot8 >= 2 * cap:
                    n += -(-tot8 // cap)
                else:
                    for g4 in range(g8, g8 + n8, 4):
                        n += -(-(lb[min(g4 + 4, W)] - lb[g4]) // cap)
    return n


@pytest.mark.parametrize("ppt", [1, 2])
def test_chunk_rules_make_the_modelled_chunk_count(ppt):
    # The sort's chunk list, counted after the first (sorting) substep, equals
    # a row-by-row model of each rule on the oracle's tile keys: one scene and
    # a ten-scene batch of mixed sizes.
    cfgs, xs, meta = workloads.batched_scenes(10, seed=11)
    cube_cfg, cube_x, cube_meta = workloads.elastic_cube(8192, seed=3)
    for case, (cs, ps, dt) in {"cube": ([cube_cfg], [cube_x], cube_meta["dt"]),
                               "batch": (cfgs, xs, meta["dt"])}.items():
        for seg in (4, 6, 8):
            want = 0
            for c, x in zip(cs, ps):
                W = (c.grid_size + 3) // 4
                keys, _ = orc.tile_keys(c, x)
                want += _model_chunk_count(np.bincount(keys.astype(np.int64), minlength=W ** 3), W, 256 * ppt, seg)
            sim = MpmSimulation(cs) if len(cs) > 1 else MpmSimulation(cs[0], max_particles=len(ps[0]))
            if len(cs) > 1:
                sim.set_particles_batch(ps)
            else:
                sim.set_particles(ps[0], None)
            sim.set_launch_config(particles_per_thread=ppt, segment_tiles=seg)
            sim.step_n(dt, 1)
            st = sim.read_status()
            sim.close()
            log(f"  {case} ppt {ppt} segment_tiles {seg}: {st['chunks']} chunks (model {want})")
            assert st["sorts"] == 1 and st["chunks"] == want


def test_autotune_restores_state_and_keeps_results():
    cfg, x, meta = workloads.elastic_cube(8192, seed=5)
    dt = meta["dt"]
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, None)
    x0, F0 = sim.read_positions(), sim.read_field("F")
    tuned = sim.autotune(dt, 4)
    log(f"autotune -> {tuned}")
    assert tuned["particles_per_thread"] in (1, 2) and tuned["segment_tiles"] in (4, 6, 8)
    assert sim.launch_config() == tuned
    assert np.array_equal(sim.read_positions(), x0) and np.array_equal(sim.read_field("F"), F0)
    st = sim.read_status()
    assert st["fault_bits"] == 0 and st["completed_gathers"] == 0 and st["sorts"] == 0
    sim.step_n(dt, 30)
    _, ref = run_gpu(cfg, x, None, dt, 30)
    e = rel_err(sim.read_positions(), ref["x"])
    log(f"  tuned vs default x: {e:.2e}")
    assert e < 1e-6
    assert sim.read_status()["completed_gathers"] == 30


# ------------------------------------------------------------ semi-implicit
@pytest.mark.parametrize("iters", [1, 2, 3])
def test_implicit_step_matches_fp64_oracle(iters):
    # MpmTimeStep::Implicit (mpm.wgsl:1060-1153): the Jacobi sweeps on the
    # active tiles against the oracle's dense-grid sweeps.
    cfg = MpmConfig.sand_calibrated().copy(time_step=Implicit(iters), rheology=Viscous(0, 0))
    x = fill_box([-0.2, 0.3, -0.2], [0.2, 0.6, 0.2], 10, 41)
    errs = _compare_with_oracle(cfg, x, 1e-3, 40, tag=f"implicit{iters}")
    _assert_errs(errs, {"x": 1e-6, "v": 1e-5, "F": 1e-5, "C": 1e-4})


def test_implicit_vs_explicit_settle_like_the_reference():
    # tests/mpm_v3_gpu.rs:160-230: a sand pile settles under both time steps;
    # per-particle paths stay within 0.20 m and the pile has dropped.
    x = fill_box([-0.10, 0.30, -0.10], [0.10, 0.50, 0.10], 6, 1234)
    dt = 1.0 / 480.0
    out = {}
    for name, ts in (("explicit", None), ("implicit", Implicit(2))):
        cfg = MpmConfig.sand().copy(youngs=2.0e4)
        if ts is not None:
            cfg = cfg.copy(time_step=ts)
        sim = MpmSimulation(cfg, max_particles=len(x))
        sim.set_particles(x, None)
        for _ in range(480):
            sim.step_n(dt, 1)
        out[name] = sim.read_positions()
        assert sim.read_status()["fault_bits"] == 0
        sim.close()
    diff = np.linalg.norm(out["explicit"] - out["implicit"], axis=1)
    y0 = x[:, 1].mean()
    log(f"implicit vs explicit: max diff {diff.max():.4f} m, mean y {y0:.3f} -> "
        f"{out['explicit'][:, 1].mean():.3f} / {out['implicit'][:, 1].mean():.3f}")
    assert out["explicit"][:, 1].mean() < y0 - 0.05
    assert diff.max() < 0.20


def test_implicit_survives_larger_dt_like_the_reference():
    # tests/mpm_v3_gpu.rs:234-312: at 3x the explicit CFL limit the implicit
    # step stays finite and bounded (< 100 m/s).
    x = fill_box([-0.10, 0.60, -0.10], [0.10, 0.80, 0.10], 6, 555)
    v = np.tile(np.array([0.0, -3.0, 0.0], np.float32), (len(x), 1))
    cfg = MpmConfig.sand().copy(youngs=5.0e5, time_step=Implicit(5))
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, v)
    for _ in range(40):
        sim.step_n(1.0 / 133.0, 1)
    pos, vel = sim.read_positions(), sim.read_velocities()
    st = sim.read_status()
    sim.close()
    speed = float(np.linalg.norm(vel, axis=1).max())
    log(f"implicit at dt=1/133: max |v| {speed:.2f} m/s, fault bits {st['fault_bits']}")
    assert st["fault_bits"] == 0 and np.all(np.isfinite(pos))
    assert speed < 100.0
