"""FP32 error of the GPU path against the FP64 oracle on the five benchmark
workloads, as relative errors max|gpu - oracle| / max|oracle| per field.

  C1  elastic cube, 8,192 particles, 32^3, 1000 substeps (the whole workload)
  C2  the 2M-particle large scene, cut to 100,000 particles on the same 256^3
      grid, 20 substeps; and at its full density (one elastic block whole and
      the central column of the fluid pool, ~500,000 particles) over all 200
      substeps (bench.py reports the full scene's first substeps)
  C3  four scenes of the 512-scene batch (the smallest, the largest and two
      between) stepped as one batch, 100 substeps, each against its own run
  C4  the 65,536-particle diff scene, 64 substeps, checkpoint 16: x_T, and
      dJ/dv0 and dJ/dE_p of the centroid loss
  C5  two scenes of the 128-scene diff batch (the smallest and the largest)
      as one batch, 32 substeps, checkpoint 16: x_T, dJ/dx0, dJ/dv0, dJ/dE_p

Bound: 1e-4 on every forward field (the target of the f32 path), and, where
the f32 oracle itself comes close to it, 2x the f32 oracle's error
(threaded f32 port, another summation order than the GPU's).

Gradients: 1e-4 of the largest entry, as in test_gpu_adjoint.py, or twice
the oracle's own change when F0 = I is moved off the repeated-singular-value
set by 1e-9 N(0,1), whichever is larger.  At F0 = I the reference's SVD
adjoint is discontinuous (its Tikhonov term zeroes the 1/(s_i^2 - s_j^2)
couplings whose true limit is finite); an f64 state stays on that set while
the f32 state is rounded off it, so the oracle's gradient there is one side
of a jump, and the report states the jump beside the GPU's error.
"""
import numpy as np
import pytest

import oracle as orc
from molla_amd import MpmSimulation, workloads

from gpu_util import log, oracle_state, rel_err

pytestmark = pytest.mark.gpu

TARGET = 1e-4


def _check(tag, errs, e32=None):
    line = " ".join(f"{k} {e:.2e}" for k, e in errs.items())
    if e32:
        line += " | f32 port " + " ".join(f"{k} {e:.2e}" for k, e in e32.items())
    log(f"ERROR_REPORT {tag}: {line}")
    for k, e in errs.items():
        bound = max(TARGET, 2 * e32[k]) if e32 and k in e32 else TARGET
        assert e <= bound, (tag, k, e, bound)


def _shader_run(cfg, x, dt, steps, fields=None):
    sim = MpmSimulation(cfg, max_particles=len(x))
    sim.set_particles(x, None)
    for k, val in (fields or {}).items():
        sim.write_field(k, val)
    sim.step_n(dt, steps)
    out = {"x": sim.read_positions(), "v": sim.read_velocities(), "F": sim.read_field("F").reshape(-1, 3, 3)}
    assert sim.read_status()["fault_bits"] == 0
    sim.close()
    return out


def _oracle_pair(cfg, x, dt, steps, **kw):
    o64 = oracle_state(cfg, x, **kw)
    o32 = oracle_state(cfg, x, dtype=np.float32, **kw)
    orc.step(cfg, o64, dt, steps)
    orc.step_threads(cfg, o32, dt, steps)
    return o64, o32


def _errs(g, o64):
    return {f: rel_err(g[f], getattr(o64, f)) for f in ("x", "v", "F")}


def test_c1_elastic_cube_full_run():
    cfg, x, meta = workloads.elastic_cube(8192, seed=1)
    g = _shader_run(cfg, x, meta["dt"], 1000)
    o64, o32 = _oracle_pair(cfg, x, meta["dt"], 1000)
    _check("C1 elastic_cube n=8192 substeps=1000", _errs(g, o64), _errs({f: getattr(o32, f) for f in "xvF"}, o64))


def test_c2_large_scene_cut():
    cfg, d, meta = workloads.large_scene(n=100_000)
    fields = {"mass": d["mass"], "volume": d["volume"], "material": d["material"]}
    sim = MpmSimulation(cfg, max_particles=len(d["x"]))
    sim.set_particles(d["x"], d["v"])
    for k, val in fields.items():
        sim.write_field(k, val)
    sim.step_n(meta["dt"], 20)
    g = {"x": sim.read_positions(), "v": sim.read_velocities(), "F": sim.read_field("F").reshape(-1, 3, 3)}
    assert sim.read_status()["fault_bits"] == 0
    sim.close()
    kw = {"v": d["v"], "mass": d["mass"], "vol": d["volume"], "mat": d["material"]}
    o64, o32 = _oracle_pair(cfg, d["x"], meta["dt"], 20, **kw)
    _check("C2 large_scene n=100000 G=256 substeps=20", _errs(g, o64),
           _errs({f: getattr(o32, f) for f in "xvF"}, o64))


def test_c2_full_density_cut_200_substeps():
    # The large scene at its real density over its whole 200 substeps: one of
    # the four elastic blocks whole (250,000 particles) and the fluid pool cut
    # to its central 48 x 48-cell column (about 250,000 particles), on the
    # same 256^3 grid.  FP64 reference: the oracle's threaded f64 step.
    cfg, d, meta = workloads.large_scene()
    x = d["x"]
    nf = int(np.sum(d["material"] == 0))
    ne = (len(x) - nf) // 4
    dx = cfg.cell_size
    fl = np.arange(nf)[(np.abs(x[:nf, 0] - 128 * dx) < 24 * dx) & (np.abs(x[:nf, 2] - 128 * dx) < 24 * dx)]
    sel = np.concatenate([fl, nf + np.arange(ne)])
    sub = {k: d[k][sel] for k in ("x", "v", "mass", "volume", "material")}
    steps = meta["substeps"]
    log(f"C2 full-density cut: {len(fl)} fluid + {ne} elastic particles, {steps} substeps")
    sim = MpmSimulation(cfg, max_particles=len(sel))
    sim.set_particles(sub["x"], sub["v"])
    for k in ("mass", "volume", "material"):
        sim.write_field(k, sub[k])
    sim.step_n(meta["dt"], steps)
    g = {"x": sim.read_positions(), "v": sim.read_velocities(), "F": sim.read_field("F").reshape(-1, 3, 3)}
    st = sim.read_status()
    sim.close()
    assert st["fault_bits"] == 0, st
    kw = {"v": sub["v"], "mass": sub["mass"], "vol": sub["volume"], "mat": sub["material"]}
    o64 = oracle_state(cfg, sub["x"], **kw)
    o32 = oracle_state(cfg, sub["x"], dtype=np.float32, **kw)
    orc.step_threads(cfg, o64, meta["dt"], steps)
    orc.step_threads(cfg, o32, meta["dt"], steps)
    _check(f"C2 full-density cut n={len(sel)} G=256 substeps={steps}", _errs(g, o64),
           _errs({f: getattr(o32, f) for f in "xvF"}, o64))


def test_c3_batch_sample():
    cfgs, xs, meta = workloads.batched_scenes(512)
    counts = np.array([len(x) for x in xs])
    order = np.argsort(counts)
    pick = [int(order[0]), int(order[len(order) // 3]), int(order[2 * len(order) // 3]), int(order[-1])]
    cfgs = [cfgs[i] for i in pick]
    xs = [xs[i] for i in pick]
    sim = MpmSimulation(cfgs)
    sim.set_particles_batch(xs)
    dt, steps = meta["dt"], 100
    sim.step_n(dt, steps)
    g = {"x": sim.read_positions(), "v": sim.read_velocities(), "F": sim.read_field("F").reshape(-1, 3, 3)}
    assert sim.read_status()["fault_bits"] == 0
    sim.close()
    off = 0
    for i, c, x in zip(pick, cfgs, xs):
        sl = slice(off, off + len(x))
        o64, o32 = _oracle_pair(c, x, dt, steps)
        _check(f"C3 scene {i} n={len(x)} G={c.grid_size} substeps={steps}",
               _errs({f: g[f][sl] for f in "xvF"}, o64), _errs({f: getattr(o32, f) for f in "xvF"}, o64))
        off += len(x)


def _centroid(n, target):
    def loss_grad(xT):
        r = np.asarray(xT, np.float64).mean(axis=0) - target
        return np.tile((2.0 / n) * r, (n, 1))
    return loss_grad


def _diff_oracle(cfg, x, v, E, dt, steps, grad, F0=None):
    st = oracle_state(cfg, x, v, E=E.astype(np.float64), F=F0)
    st0 = st.copy()
    tape = orc.diff_forward(cfg, st, dt, steps)
    o = orc.diff_backward(cfg, st0, dt, steps, tape, grad(st.x), np.zeros((len(x), 3)))
    o["xT"] = st.x
    return o


def _near_identity(n, seed):
    rng = np.random.default_rng(seed)
    return np.eye(3)[None] + 1e-9 * rng.standard_normal((n, 3, 3))


def _check_grads(tag, errs, jumps):
    log(f"ERROR_REPORT {tag}: " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()) +
        " | oracle jump at F0 = I: " + " ".join(f"{k} {e:.2e}" for k, e in jumps.items()))
    for k, e in errs.items():
        bound = max(TARGET, 2 * jumps.get(k, 0.0))
        assert e <= bound, (tag, k, e, bound)


def test_c4_diff_scene_full_run():
    cfg, d, meta = workloads.diff_scene(65536)
    n, dt, steps, every = len(d["x"]), meta["dt"], meta["substeps"], meta["checkpoint_every"]
    grad = _centroid(n, np.asarray(meta["target"]))
    sim = MpmSimulation(cfg, max_particles=n)
    sim.set_particles(d["x"], d["v"])
    sim.write_field("youngs", d["E"])
    sim.diff_forward(dt, steps, every)
    xT = sim.read_positions()
    gx = grad(xT).astype(np.float32)
    g = sim.diff_backward(gx)
    sim.close()
    o = _diff_oracle(cfg, d["x"], d["v"], d["E"], dt, steps, grad)
    p = _diff_oracle(cfg, d["x"], d["v"], d["E"], dt, steps, grad, F0=_near_identity(n, 52))
    _check_grads(f"C4 diff_scene n={n} substeps={steps} checkpoint={every}",
                 {"x_T": rel_err(xT, o["xT"]), "dJ/dv0": rel_err(g["v"], o["v"]), "dJ/dE": rel_err(g["E"], o["E"])},
                 {"dJ/dv0": rel_err(p["v"], o["v"]), "dJ/dE": rel_err(p["E"], o["E"])})


def test_c5_diff_batch_sample():
    cfgs, d, meta = workloads.batched_diff_scenes(128)
    counts = np.asarray(meta["counts"])
    pick = [int(np.argmin(counts)), int(np.argmax(counts))]
    cfgs = [cfgs[i] for i in pick]
    xs = [d["x"][i] for i in pick]
    vs = [d["v"][i] for i in pick]
    dt, steps, every = meta["dt"], meta["substeps"], meta["checkpoint_every"]
    n = sum(len(x) for x in xs)
    rng = np.random.default_rng(51)
    gx = rng.standard_normal((n, 3)).astype(np.float32)
    Es = [np.full(len(x), meta["e0"][i], np.float32) for i, x in zip(pick, xs)]
    sim = MpmSimulation(cfgs)
    sim.set_particles_batch(xs, vs)
    sim.write_field("youngs", np.concatenate(Es))
    sim.diff_forward(dt, steps, every)
    xT = sim.read_positions()
    g = sim.diff_backward(gx)
    sim.close()
    off = 0
    for i, c, x, v, E in zip(pick, cfgs, xs, vs, Es):
        sl = slice(off, off + len(x))
        fixed = gx[sl].astype(np.float64)
        o = _diff_oracle(c, x, v, E, dt, steps, lambda xT: fixed)
        p = _diff_oracle(c, x, v, E, dt, steps, lambda xT: fixed, F0=_near_identity(len(x), 53 + i))
        gxs = gx[sl]
        _check_grads(f"C5 scene {i} n={len(x)} G={c.grid_size} substeps={steps} checkpoint={every}",
                     {"x_T": rel_err(xT[sl], o["xT"]), "dJ/dx0": rel_err(g["x"][sl] - gxs, o["x"] - gxs),
                      "dJ/dv0": rel_err(g["v"][sl], o["v"]), "dJ/dE": rel_err(g["E"][sl], o["E"])},
                     {"dJ/dx0": rel_err(p["x"] - gxs, o["x"] - gxs), "dJ/dv0": rel_err(p["v"], o["v"]),
                      "dJ/dE": rel_err(p["E"], o["E"])})
        off += len(x)
