#!/usr/bin/env python3
"""MPM substep benchmark on one MI355X.

Workload (N=1): the 2M-particle single scene on a 256^3 grid (half
weakly-compressible fluid, half fixed-corotated elastic blocks, dt = 5e-5),
built by molla_amd.workloads.large_scene.  A "step" is one MPM substep over
all particles; K substeps run as one step_n call with inputs resident in HBM.

Prints ONE JSON line.  Besides the headline value it carries:
  roofline      the dominant kernel (fused G2P->P2G) against the HBM peak, with
                algorithmic bytes = 192 B per particle-substep (f32 x, v, C, F
                read once and written once) and its duration measured with HIP
                events recorded on the library's stream around every launch
                of a second run of the same number of substeps right after the
                timed one (the events slow a substep by about 5 %, so `value`
                is taken without them); `traffic` (HBM bytes per launch) is
                taken from the FETCH_SIZE/WRITE_SIZE counter summary in
                profiles/, scaled by the same counters on a copy of known size;
                `roofline.alu` puts the same launches against the VALU issue
                rate (instructions and flops per launch from the counter
                summary in profiles/, duration live), the roof that binds
  cpu_baseline  the FP64 single-core oracle (restatement of the reference CPU
                path) on a bounded sample of the same workload, with the
                GPU's relative error against it on that sample
  cpu_baseline_fair
                the f32 threaded (OpenMP) port of the oracle on every host
                thread, same sample
  batched       the 512-scene batch (one launch per substep, hipGraph replay)
  adjoint, batched_adjoint
                diff_forward with checkpoints, then diff_backward: the
                adjoint/forward ratio from HIP events around each call's device
                work with the gradients resident in HBM, and beside it the
                wall-clock ratio that includes the gradients' host transfers
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MPM particle-substeps/s, plain vs batched, and % HBM roofline; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
BYTES_FWD = 192.0
STEP_STAGES = ("sort", "p2g", "grid", "g2p2g", "g2p")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_large(sim_cls, scale):
    from molla_amd import workloads

    n = int(2_000_000 * scale)
    cfg, d, meta = workloads.large_scene(n=n)
    sim = sim_cls(cfg, max_particles=len(d["x"]))
    sim.set_particles(d["x"], d["v"])
    sim.write_field("mass", d["mass"])
    sim.write_field("volume", d["volume"])
    sim.write_field("material", d["material"])
    return sim, cfg, d, meta


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_particles_g2p2g_large_scene_traffic.json")
ALU_FILE = os.path.join(ROOT, "profiles", "k_particles_g2p2g_large_scene_alu.json")
# one wave64 VALU instruction per CU per cycle (4 SIMD16 x 4 cycles), 256 CUs
# at 2.4 GHz; f32 non-packed and f64 FMA issue at this rate
VALU_ISSUE_PEAK = 256 * 2.4e9


def calibrate_copy(mib):
    """Device-to-device copy of `mib` MiB (one blit kernel per call), for
    scaling the HBM byte counters: run under rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE next to the MPM kernels."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = int(mib) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {nbytes} B device-to-device")


def traffic_summary(n):
    """The committed counter and kernel-trace summary of the G2P2G launches,
    if it was taken on the same particle count."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    return t if t.get("particles") == n else {}


def alu_roofline(n, avg_s):
    """The fused pass against the VALU issue rate: instructions and flops per
    launch from the committed counter summary (same particle count), over the
    launch duration measured live."""
    try:
        with open(ALU_FILE) as f:
            a = json.load(f)
    except (OSError, ValueError):
        return None
    if a.get("particles") != n or not avg_s:
        return None
    ach = a["valu_instructions_per_launch"] / avg_s
    return {"bound": "valu_issue", "achieved": ach, "peak": VALU_ISSUE_PEAK,
            "unit": "wave64 VALU instructions/s", "frac": ach / VALU_ISSUE_PEAK,
            "valu_instructions_per_launch": a["valu_instructions_per_launch"],
            "flops_per_particle_substep": {"f32": a["flops_f32_per_launch"] / n, "f64": a["flops_f64_per_launch"] / n},
            "tflops": {"f32": a["flops_f32_per_launch"] / avg_s / 1e12, "f64": a["flops_f64_per_launch"] / avg_s / 1e12},
            "lds_bank_conflict_fraction": a["lds_bank_conflict_cycles"] / a["lds_indexed_active_cycles"],
            "source": a["source"]}


def time_steps(sim, dt, k, graph):
    sim.synchronize()
    t0 = time.perf_counter()
    if graph:
        sim.step_n_graph(dt, k)
    else:
        sim.step_n(dt, k)
    sim.synchronize()
    return time.perf_counter() - t0


def time_sort(sim, dt, graph):
    """Wall time of one layout sort: a substep forced to sort (interval 1,
    right after a substep) minus a substep without one, best of three."""
    best = []
    for force in (True, False):
        ts = []
        for _ in range(3):
            sim.set_sort_interval(1 if force else 1 << 20)
            if force:
                sim.step_n(dt, 1)  # since_sort = 1 >= 1: the next substep sorts
            ts.append(time_steps(sim, dt, 1, False))
        best.append(min(ts))
    return max(best[0] - best[1], 0.0)


def cpu_baseline(cfg, d, dt, substeps, gpu_state):
    """The CPU legs on one bounded sample (the first `substeps` substeps of the
    workload): the FP64 oracle on one core (faithful), the f32 threaded port
    on every host thread (fair), and, with the FP64 result as the checker,
    the relative error of the GPU's state after the same substeps."""
    import oracle as orc

    n = len(d["x"])

    def state(dtype):
        return orc.State(d["x"], d["v"], mass=d["mass"], vol=d["volume"], mat=d["material"], dtype=dtype)

    st = state(np.float64)
    t0 = time.perf_counter()
    orc.step(cfg, st, dt, substeps)
    el = time.perf_counter() - t0
    s32 = state(np.float32)
    t0 = time.perf_counter()
    threads = orc.step_threads(cfg, s32, dt, substeps)
    el_fair = time.perf_counter() - t0

    def rel(a, b):
        return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))

    err = {f: rel(gpu_state[f], getattr(st, f)) for f in ("x", "v", "F")}
    err32 = {f: rel(getattr(s32, f), getattr(st, f)) for f in ("x", "v", "F")}
    return {"value": n * substeps / el, "seconds": el, "fair_value": n * substeps / el_fair,
            "fair_seconds": el_fair, "threads": threads, "gpu_error": err, "fair_error": err32}


def cpu_legs(scale, out):
    """CPU timings of the other workloads on this host, beside the GPU's:
    the elastic cube in full (1,000 substeps) with the FP64 oracle on one core
    and the f32 threaded port on every host thread; the differentiable scene
    and a sample of the differentiable batch through the oracle's CPUDIFF
    forward and reverse (mpm_cpu_diff.rs step_once_record and
    mpm_adjoint.rs backward restated; single thread, f64, as the reference)."""
    import oracle as orc
    from molla_amd import workloads

    legs = {}
    # elastic cube, whole workload
    n = max(1024, int(8192 * min(scale, 1.0)))
    cfg, x, meta = workloads.elastic_cube(n, seed=1)
    steps = meta["substeps"]
    gpu = out.get("plain", {}).get("value")
    for dtype, name in ((np.float64, "elastic_cube_fp64"), (np.float32, "elastic_cube_f32_threads")):
        st = orc.State(x, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume), dtype=dtype)
        t0 = time.perf_counter()
        if dtype == np.float64:
            orc.step(cfg, st, meta["dt"], steps)
            threads = 1
        else:
            threads = orc.step_threads(cfg, st, meta["dt"], steps)
        el = time.perf_counter() - t0
        v = n * steps / el
        legs[name] = {"value": v, "unit": "particle-substeps/s", "cores": threads,
                      "kind": "port", "sample": f"whole workload: {n} particles x {steps} substeps, {el:.1f} s",
                      "gpu_over_cpu": gpu / v if gpu else None}
        log(f"cpu leg {name}: {legs[name]}")

    def diff_leg(cfgs, states, dt, steps, grads):
        tf = tb = 0.0
        for c, st, g in zip(cfgs, states, grads):
            st0 = st.copy()
            t0 = time.perf_counter()
            tape = orc.diff_forward(c, st, dt, steps)
            tf += time.perf_counter() - t0
            gx = g(st)
            t0 = time.perf_counter()
            orc.diff_backward(c, st0, dt, steps, tape, gx)
            tb += time.perf_counter() - t0
        return tf, tb

    # differentiable scene: the first 16 of its 64 substeps (one checkpoint
    # segment), forward with tape and reverse
    n = max(4096, int(65536 * scale))
    cfg, d, meta = workloads.diff_scene(n)
    k = 16
    target = np.asarray(meta["target"])
    st = orc.State(d["x"], d["v"], mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume),
                   E=d["E"])
    tf, tb = diff_leg([cfg], [st], meta["dt"], k,
                      [lambda s: np.tile((2.0 / n) * (s.x.mean(axis=0) - target), (n, 1))])
    ga = out.get("adjoint", {})
    legs["diff_scene"] = {"forward_value": n * k / tf, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
                          "adjoint_fwd_ratio": tb / tf,
                          "sample": f"first {k} of {meta['substeps']} substeps, {n} particles: forward {tf:.2f} s, "
                                    f"reverse {tb:.2f} s",
                          "gpu_over_cpu_forward": (n * meta["substeps"] / (1e-3 * ga["forward_ms"])) / (n * k / tf)
                          if "forward_ms" in ga else None}
    log(f"cpu leg diff_scene: {legs['diff_scene']}")
    # differentiable batch: its first 2 scenes, all 32 substeps
    cfgs, d, meta = workloads.batched_diff_scenes(max(4, int(128 * scale)))
    m = 2
    states = [orc.State(d["x"][i], d["v"][i], mass=np.full(len(d["x"][i]), cfgs[i].particle_mass),
                        vol=np.full(len(d["x"][i]), cfgs[i].particle_volume),
                        E=np.full(len(d["x"][i]), meta["e0"][i])) for i in range(m)]
    ns = sum(len(d["x"][i]) for i in range(m))
    tf, tb = diff_leg(cfgs[:m], states, meta["dt"], meta["substeps"],
                      [lambda s: (2.0 / s.n) * s.x] * m)
    gb = out.get("batched_adjoint", {})
    legs["batched_diff_scenes"] = {
        "forward_value": ns * meta["substeps"] / tf, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
        "adjoint_fwd_ratio": tb / tf,
        "sample": f"first {m} scenes ({ns} particles) x {meta['substeps']} substeps: forward {tf:.2f} s, "
                  f"reverse {tb:.2f} s",
        "gpu_over_cpu_forward": gb["forward_value"] / (ns * meta["substeps"] / tf) if "forward_value" in gb else None}
    log(f"cpu leg batched_diff_scenes: {legs['batched_diff_scenes']}")
    return legs


def run_batched(steps, warmup, scale, autotune=True):
    from molla_amd import MpmSimulation, workloads

    n_scenes = max(8, int(512 * scale))
    cfgs, xs, meta = workloads.batched_scenes(n_scenes)
    sim = MpmSimulation(cfgs)
    sim.set_particles_batch(xs)
    dt = meta["dt"]
    tuned = sim.autotune(dt, 10) if autotune else None
    if warmup:
        sim.step_n_graph(dt, steps)
    el = time_steps(sim, dt, steps, graph=True)
    n = meta["particles"]
    sim.close()
    return {"value": n * steps / el, "unit": "particle-substeps/s", "scenes": n_scenes,
            "particles": n, "substeps": steps, "ms_per_step": 1e3 * el / steps,
            "hbm_frac": n * steps / el * BYTES_FWD / (HBM_PEAK_GBS * 1e9),
            "grids": [meta["grid_min"], meta["grid_max"]], "graph": True, "autotuned": tuned}


def run_plain_per_scene(steps, scale):
    """The same scenes as the batch, each stepped through its own
    single-scene handle (the reference's one MpmSimulation per scene,
    mpm.rs:2252-2353): every handle's step_n of `steps` substeps is issued in
    turn and all are waited for at the end.  Launch sizes are the library
    defaults (the autotuner's trials would cost more than the run)."""
    from molla_amd import MpmSimulation, workloads

    n_scenes = max(8, int(512 * scale))
    cfgs, xs, meta = workloads.batched_scenes(n_scenes)
    dt = meta["dt"]
    sims = []
    try:
        for c, x in zip(cfgs, xs):
            s = MpmSimulation(c, max_particles=len(x))
            s.set_particles(x)
            sims.append(s)
        for s in sims:  # warm-up: first sort, grid allocation
            s.step_n(dt, 2)
        for s in sims:
            s.synchronize()
        t0 = time.perf_counter()
        for s in sims:
            s.step_n(dt, steps)
        for s in sims:
            s.synchronize()
        el = time.perf_counter() - t0
    finally:
        for s in sims:
            s.close()
    n = meta["particles"]
    return {"workload": "batched_scenes, one handle per scene", "scenes": n_scenes, "particles": n,
            "substeps": steps, "value": n * steps / el, "unit": "particle-substeps/s",
            "ms_per_step": 1e3 * el / steps}


def run_plain(scale):
    """The 8,192-particle elastic cube on a 32^3 grid, one scene, 1000
    substeps in one step_n call (the reference's plain single-simulation
    configuration)."""
    from molla_amd import MpmSimulation, workloads

    n = max(1024, int(8192 * min(scale, 1.0)))
    cfg, x, meta = workloads.elastic_cube(n, seed=1)
    sim = MpmSimulation(cfg, max_particles=n)
    sim.set_particles(x, None)
    # the autotuner's launch sizes, as for the batch; eager stepping (a graph
    # replay of this two-kernel substep is slower on this stack)
    tuned = sim.autotune(meta["dt"], 20)
    sim.step_n(meta["dt"], 20)
    el = time_steps(sim, meta["dt"], 1000, graph=False)
    sim.close()
    return {"workload": "elastic_cube", "particles": n, "grid": cfg.grid_size, "substeps": 1000,
            "value": n * 1000 / el, "unit": "particle-substeps/s", "ms_per_step": 1e3 * el / 1000,
            "autotuned": tuned}


def timed_diff(sim, dt, steps, every, loss_grad):
    """One timed diff_forward + diff_backward.  Wall-clock times include the
    host side of each call (the backward's upload of dJ/dx_T and download of
    the initial-state gradients); the device times are HIP events on the
    library's stream around each call's device work, with the gradients
    resident in HBM (the backward from its uploaded input to its last reverse
    kernel)."""
    sim.synchronize()
    sim.set_profiling(True, stages=False)
    sim.reset_profile()
    t0 = time.perf_counter()
    sim.diff_forward(dt, steps, every)
    t_fwd = time.perf_counter() - t0
    J, gx = loss_grad()
    t0 = time.perf_counter()
    g = sim.diff_backward(gx)
    t_bwd = time.perf_counter() - t0
    prof = sim.read_profile()
    sim.set_profiling(False)
    f_dev, b_dev = prof["diff_forward"][0], prof["diff_backward"][0]
    assert prof["diff_forward"][1] == 1 and prof["diff_backward"][1] == 1
    return J, g, {"forward_ms": f_dev, "backward_ms": b_dev, "adjoint_fwd_ratio": b_dev / f_dev,
                  "forward_wall_ms": 1e3 * t_fwd, "backward_wall_ms": 1e3 * t_bwd,
                  "adjoint_fwd_ratio_wall": t_bwd / t_fwd}


def device_wall(sim, dt, steps, every, times, reset, loss_grad):
    """Adds the wall-clock ratio with the gradients kept in HBM
    (diff_backward_device) to `times`; the host-transfer variant's wall ratio
    is kept as adjoint_fwd_ratio_wall_host.  A failure is reported, not
    hidden."""
    times = dict(times)
    times["adjoint_fwd_ratio_wall_host"] = times.pop("adjoint_fwd_ratio_wall")
    times["forward_wall_ms_host"] = times.pop("forward_wall_ms")
    times["backward_wall_ms_host"] = times.pop("backward_wall_ms")
    try:
        reset()
        timed_diff_device(sim, dt, steps, every, loss_grad)  # warm-up
        reset()
        times.update(timed_diff_device(sim, dt, steps, every, loss_grad))
    except Exception as e:  # reported, never silently replaced
        times["device_wall_error"] = str(e)
    return times


def timed_diff_device(sim, dt, steps, every, loss_grad):
    """One timed diff_forward + diff_backward_device: the reverse pass reads
    dJ/dx_T from device memory and writes the initial-state gradients to
    device memory, so nothing crosses PCIe inside it.  dJ/dx_T is placed in
    HBM before the timed call (a caller whose loss runs on the GPU has it
    there).  Wall-clock times of the two calls, each up to the end of its
    device work."""
    from molla_amd.devmem import DeviceArray

    n = sim.n
    outs = [DeviceArray((n, 3)), DeviceArray((n, 3)), DeviceArray(n), DeviceArray(2, np.float64)]
    sim.synchronize()
    t0 = time.perf_counter()
    sim.diff_forward(dt, steps, every)
    sim.synchronize()
    t_fwd = time.perf_counter() - t0
    _, gx = loss_grad()
    gx_dev = DeviceArray.from_host(gx)
    t0 = time.perf_counter()
    sim.diff_backward_device(gx_dev.address, 0, *(o.address for o in outs))
    sim.synchronize()
    t_bwd = time.perf_counter() - t0
    for o in (*outs, gx_dev):
        o.free()
    return {"forward_wall_ms": 1e3 * t_fwd, "backward_wall_ms": 1e3 * t_bwd,
            "adjoint_fwd_ratio_wall": t_bwd / t_fwd}


def run_adjoint(scale):
    """Forward with checkpoints + reverse pass of the 65,536-particle diff scene
    (64 substeps, checkpoint every 16, COM loss); reports the reverse/forward
    time ratio."""
    from molla_amd import MpmSimulation, workloads

    n = max(4096, int(65536 * scale))
    cfg, d, meta = workloads.diff_scene(n)
    steps, every, dt = meta["substeps"], meta["checkpoint_every"], meta["dt"]
    target = np.asarray(meta["target"])
    sim = MpmSimulation(cfg, max_particles=n)

    def reset():
        sim.set_particles(d["x"], d["v"])
        sim.write_field("youngs", d["E"])

    def loss_grad():
        xT = sim.read_positions().astype(np.float64)
        r = xT.mean(axis=0) - target
        return float(r @ r), np.tile((2.0 / n) * r, (n, 1)).astype(np.float32)

    reset()
    sim.diff_forward(dt, steps, every)
    sim.diff_backward(loss_grad()[1])
    reset()
    J, g, times = timed_diff(sim, dt, steps, every, loss_grad)
    reset()
    times = device_wall(sim, dt, steps, every, times, reset, loss_grad)
    reset()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step_n(dt, steps)
    sim.synchronize()
    t_plain = time.perf_counter() - t0
    sim.close()
    return {"workload": "diff_scene", "particles": n, "grid": meta["grid"], "substeps": steps,
            "checkpoint_every": every, **times, "plain_step_n_ms": 1e3 * t_plain,
            "loss": J, "dJ_dyoungs_sum": float(np.sum(g["E"], dtype=np.float64))}


def run_batched_adjoint(scale):
    """Batched system identification: 128 CPU-diff scenes (~1.8M particles,
    48^3..80^3 grids), 32 substeps, checkpoint every 16; loss = sum over
    scenes of the mean squared distance of x_T to the ground-truth run's."""
    from molla_amd import MpmSimulation, workloads

    n_scenes = max(4, int(128 * scale))
    cfgs, d, meta = workloads.batched_diff_scenes(n_scenes)
    counts, dt, steps, every = meta["counts"], meta["dt"], meta["substeps"], meta["checkpoint_every"]
    n = int(sum(counts))
    sim = MpmSimulation(cfgs)

    def start(e_scene):
        sim.set_particles_batch(d["x"], d["v"])
        sim.write_field("youngs", np.concatenate([np.full(c, e, np.float32) for c, e in zip(counts, e_scene)]))

    start(meta["e_true"])
    sim.step_n(dt, steps)
    target = sim.read_positions()
    scale_p = np.repeat(1.0 / np.asarray(counts, np.float64), counts)[:, None]

    def loss_grad():
        r = sim.read_positions().astype(np.float64) - target
        return float(np.sum(scale_p * r * r)), (2.0 * scale_p * r).astype(np.float32)

    start(meta["e0"])
    sim.diff_forward(dt, steps, every)
    sim.diff_backward(loss_grad()[1])
    start(meta["e0"])
    J, g, times = timed_diff(sim, dt, steps, every, loss_grad)
    times = device_wall(sim, dt, steps, every, times, lambda: start(meta["e0"]), loss_grad)
    sim.close()
    ge = g["youngs_per_scene"]
    # the loss is minimised at E_true: dJ/dE should point from E0 towards it
    agree = float(np.mean(np.sign(ge) == np.sign(np.asarray(meta["e0"]) - np.asarray(meta["e_true"]))))
    return {"workload": "batched_diff_scenes", "scenes": n_scenes, "particles": n, "substeps": steps,
            "checkpoint_every": every, "grids": [meta["grid_min"], meta["grid_max"]], **times,
            "forward_value": n * steps / (1e-3 * times["forward_ms"]), "unit": "particle-substeps/s", "loss": J,
            "dJ_dE_sign_agrees_with_E0_minus_Etrue": agree}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload size")
    ap.add_argument("--cpu-substeps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay step_n as a hipGraph")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner's choice")
    ap.add_argument("--launch", type=int, nargs=3, metavar=("PARTICLE_BLOCKS", "GRID_BLOCKS", "SORT_EVERY"),
                    help="fixed launch sizes and sort interval instead of the autotuner's (for profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device copy of MIB MiB (for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the batched workload would shard by scene with no communication)", file=sys.stderr)
        sys.exit(2)

    from molla_amd import MpmSimulation

    sim, cfg, d, meta = build_large(MpmSimulation, args.scale)
    n = len(d["x"])
    dt = meta["dt"]
    log(f"large_scene: {n} particles, grid {meta['grid']}^3, dt {dt}")
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
    # launch sizes and sort interval from the library's autotuner (trial
    # substeps, state restored), as for the batch
    if args.launch:
        sim.set_launch(args.launch[0], args.launch[1])
        sim.set_sort_interval(args.launch[2])
        tuned = {"particle_blocks": args.launch[0], "grid_blocks": args.launch[1], "sort_every": args.launch[2],
                 "fixed": True}
    else:
        tuned = None if args.no_tune else sim.autotune(dt, 10)
    log(f"autotuned: {tuned}")
    if args.warmup:
        time_steps(sim, dt, args.warmup, args.graph)
    sorts0 = sim.read_status()["sorts"]
    el_raw = time_steps(sim, dt, args.steps, args.graph)
    sorts_in = sim.read_status()["sorts"] - sorts0
    # The layout sort runs every `sort_every` substeps; a timed window shorter
    # than that interval may hold none.  The window's time is adjusted to carry
    # exactly steps / sort_every sorts, each at the time of a forced sort
    # measured right after.
    every = int((tuned or {}).get("sort_every", 16))  # the library's default interval is 16
    t_sort = time_sort(sim, dt, args.graph)
    amort = args.steps / every if every else 0.0
    el = el_raw + (amort - sorts_in) * t_sort
    sim.set_sort_interval(every)
    value = n * args.steps / el
    log(f"timed {args.steps} substeps: {el_raw * 1e3:.2f} ms with {sorts_in} sorts; amortised sorts "
        f"({amort:.3f} x {t_sort * 1e3:.3f} ms) -> {el * 1e3:.2f} ms -> {value:.3e} particle-substeps/s")
    # The same number of substeps again with HIP events on the library's
    # stream around every launch, for the per-kernel durations (the events
    # cost about 5 % of a substep, so `value` comes from the run above; graph
    # replays are one opaque launch, so per-kernel events need the eager path)
    sim.set_profiling(not args.graph, calls=False)
    sim.reset_profile()
    el_prof = time_steps(sim, dt, args.steps, args.graph)
    prof = sim.read_profile()
    sim.set_profiling(False)
    log(f"profiled {args.steps} substeps: {el_prof * 1e3:.2f} ms")
    st = sim.read_status()
    ms, cnt = prof["g2p2g"]
    avg = ms / max(cnt, 1) * 1e-3
    achieved = BYTES_FWD * n / avg / 1e9 if avg > 0 else None
    summary = traffic_summary(n)
    traffic = summary.get("bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": "k_particles<G2P2G>", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": traffic, "avg_launch_ms": avg * 1e3, "launches": cnt,
                "algorithmic_bytes_per_launch": BYTES_FWD * n,
                "profiled_ms_per_step": el_prof * 1e3 / args.steps,
                # rocprofv3 kernel-trace mean of the same launches (profiles/): the
                # event bracket also holds the dispatch delay its markers add
                "kernel_trace_avg_ms": summary.get("kernel_trace_avg_ms"),
                "stage_ms_per_substep": {k: (v[0] / args.steps) for k, v in prof.items() if k in STEP_STAGES},
                "stage_launches": {k: v[1] for k, v in prof.items() if k in STEP_STAGES},
                "alu": alu_roofline(n, avg)}
    log(f"profile: {prof}")
    log(f"status: {st}")

    out = {"metric": METRIC, "value": value, "unit": "particle-substeps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps,
           "sort_amortised": {"sort_every": every, "sort_ms": t_sort * 1e3, "sorts_in_window": sorts_in,
                              "sorts_carried": amort, "window_ms": el_raw * 1e3,
                              "value_window_only": n * args.steps / el_raw},
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "large_scene", "particles": n, "grid": meta["grid"], "dt": dt,
                      "rho": meta["rho"], "seed": meta["seed"], "fluid_fraction": 0.5,
                      "graph": bool(args.graph), "autotuned": tuned, "fault_bits": st["fault_bits"]},
           "roofline": roofline}

    if not args.no_batched:
        try:
            out["batched"] = run_batched(min(args.steps, 500), args.warmup, args.scale)
            log(f"batched: {out['batched']}")
        except Exception as e:  # reported, never silently replaced
            out["batched"] = {"error": str(e)}

    try:
        if args.no_plain:
            raise RuntimeError("skipped (--no-plain)")
        out["plain"] = run_plain(args.scale)
        log(f"plain: {out['plain']}")
        out["plain_per_scene"] = run_plain_per_scene(min(args.steps, 100), args.scale)
        log(f"plain per scene: {out['plain_per_scene']}")
        if "batched" in out and "value" in out["batched"]:
            # the batch against the same scenes stepped one handle each
            out["batched"]["vs_plain"] = out["batched"]["value"] / out["plain_per_scene"]["value"]
            out["batched"]["vs_plain_cube"] = out["batched"]["value"] / out["plain"]["value"]
    except Exception as e:  # reported, never silently replaced
        out.setdefault("plain", {"error": str(e)})
        out.setdefault("plain_per_scene", {"error": str(e)})

    if not args.no_adjoint:
        try:
            out["adjoint"] = run_adjoint(args.scale)
            log(f"adjoint: {out['adjoint']}")
        except Exception as e:  # reported, never silently replaced
            out["adjoint"] = {"error": str(e)}
        try:
            out["batched_adjoint"] = run_batched_adjoint(args.scale)
            log(f"batched adjoint: {out['batched_adjoint']}")
        except Exception as e:  # reported, never silently replaced
            out["batched_adjoint"] = {"error": str(e)}

    if not args.no_cpu:
        sim.close()
        k = args.cpu_substeps
        # the GPU's state after the sample's substeps, from the same start
        from molla_amd import MpmSimulation
        gsim, _, _, _ = build_large(MpmSimulation, args.scale)
        gsim.step_n(dt, k)
        gpu_state = {"x": gsim.read_positions(), "v": gsim.read_velocities(),
                     "F": gsim.read_field("F").reshape(-1, 3, 3)}
        gsim.close()
        cb = cpu_baseline(cfg, d, dt, k, gpu_state)
        sample = f"first {k} substeps of large_scene ({n} particles)"
        out["cpu_baseline"] = {"value": cb["value"], "unit": "particle-substeps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{sample}, FP64 oracle, {cb['seconds']:.1f} s",
                               "gpu_rel_error_vs_fp64": cb["gpu_error"]}
        out["cpu_baseline_fair"] = {"value": cb["fair_value"], "unit": "particle-substeps/s",
                                    "cores": cb["threads"], "kind": "port",
                                    "sample": f"{sample}, f32 threaded port of the oracle, "
                                              f"{cb['fair_seconds']:.2f} s",
                                    "rel_error_vs_fp64": cb["fair_error"]}
        try:
            out["cpu_legs"] = cpu_legs(args.scale, out)
        except Exception as e:  # reported, never silently replaced
            out["cpu_legs"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
