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
                events on the library's stream over the timed substeps
  cpu_baseline  the FP64 single-core oracle (restatement of the reference CPU
                path) on a bounded sample of the same workload
  batched       the 512-scene batch (one launch per substep, hipGraph replay)
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


def time_steps(sim, dt, k, graph):
    sim.synchronize()
    t0 = time.perf_counter()
    if graph:
        sim.step_n_graph(dt, k)
    else:
        sim.step_n(dt, k)
    sim.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(cfg, d, dt, substeps):
    import oracle as orc

    n = len(d["x"])
    st = orc.State(d["x"], d["v"], mass=d["mass"], vol=d["volume"], mat=d["material"], dtype=np.float64)
    t0 = time.perf_counter()
    orc.step(cfg, st, dt, substeps)
    el = time.perf_counter() - t0
    return n * substeps / el, el


def run_batched(steps, warmup, scale):
    from molla_amd import MpmSimulation, workloads

    n_scenes = max(8, int(512 * scale))
    cfgs, xs, meta = workloads.batched_scenes(n_scenes)
    sim = MpmSimulation(cfgs)
    sim.set_particles_batch(xs)
    dt = meta["dt"]
    if warmup:
        sim.step_n_graph(dt, steps)
    el = time_steps(sim, dt, steps, graph=True)
    n = meta["particles"]
    sim.close()
    return {"value": n * steps / el, "unit": "particle-substeps/s", "scenes": n_scenes,
            "particles": n, "substeps": steps, "ms_per_step": 1e3 * el / steps,
            "grids": [meta["grid_min"], meta["grid_max"]], "graph": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload size")
    ap.add_argument("--cpu-substeps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay step_n as a hipGraph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(scenes shard across GPUs without communication, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from molla_amd import MpmSimulation

    sim, cfg, d, meta = build_large(MpmSimulation, args.scale)
    n = len(d["x"])
    dt = meta["dt"]
    log(f"large_scene: {n} particles, grid {meta['grid']}^3, dt {dt}")
    if args.warmup:
        time_steps(sim, dt, args.warmup, args.graph)
    el = time_steps(sim, dt, args.steps, args.graph)
    value = n * args.steps / el
    log(f"timed {args.steps} substeps: {el * 1e3:.2f} ms -> {value:.3e} particle-substeps/s")

    # per-kernel durations over a second run of the same K substeps
    sim.set_profiling(True)
    sim.reset_profile()
    sim.step_n(dt, args.steps)
    sim.synchronize()
    prof = sim.read_profile()
    sim.set_profiling(False)
    st = sim.read_status()
    ms, cnt = prof["g2p2g"]
    avg = ms / max(cnt, 1) * 1e-3
    achieved = BYTES_FWD * n / avg / 1e9 if avg > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": "k_particles<G2P2G>", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": None, "avg_launch_ms": avg * 1e3,
                "algorithmic_bytes_per_launch": BYTES_FWD * n,
                "stage_ms_per_substep": {k: (v[0] / max(v[1], 1)) for k, v in prof.items()}}
    log(f"profile: {prof}")
    log(f"status: {st}")

    out = {"metric": METRIC, "value": value, "unit": "particle-substeps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "large_scene", "particles": n, "grid": meta["grid"], "dt": dt,
                      "rho": meta["rho"], "seed": meta["seed"], "fluid_fraction": 0.5,
                      "graph": bool(args.graph), "fault_bits": st["fault_bits"]},
           "roofline": roofline}

    if not args.no_batched:
        try:
            out["batched"] = run_batched(min(args.steps, 500), args.warmup, args.scale)
            log(f"batched: {out['batched']}")
        except Exception as e:  # reported, never silently replaced
            out["batched"] = {"error": str(e)}

    if not args.no_cpu:
        sim.close()
        k = args.cpu_substeps
        rate, el_cpu = cpu_baseline(cfg, d, dt, k)
        out["cpu_baseline"] = {"value": rate, "unit": "particle-substeps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first {k} substeps of large_scene ({n} particles), FP64 oracle, "
                                         f"{el_cpu:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
