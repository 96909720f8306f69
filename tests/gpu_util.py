"""Shared helpers for the GPU parity tests (HIP path vs the CPU oracle)."""
from __future__ import annotations

import sys

import numpy as np

import oracle as orc
from molla_amd import MpmSimulation


def rel_err(a, b, floor=1e-12) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, floor)
    return float(np.abs(a - b).max()) / scale if a.size else 0.0


def log(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_state(cfg, x, v=None, dtype=np.float64, mass=None, vol=None, mat=None, E=None, F=None):
    n = len(x)
    return orc.State(x, v, F=F, mass=np.full(n, cfg.particle_mass) if mass is None else mass,
                     vol=np.full(n, cfg.particle_volume) if vol is None else vol, mat=mat, E=E,
                     dtype=dtype)


def run_gpu(cfg, x, v=None, dt=1e-4, substeps=10, mode="fused", fields=None, sort_every=None,
            graph=False, write_vc=True, launch=None):
    sim = MpmSimulation(cfg, max_particles=max(len(x), 1))
    sim.set_particles(x, v)
    for name, val in (fields or {}).items():
        sim.write_field(name, val)
    sim.set_p2g_mode(mode)
    if launch:
        sim.set_launch_config(**launch)
    if sort_every is not None:
        sim.set_sort_interval(sort_every)
    sim.set_write_vc(write_vc)
    if graph:
        sim.step_n_graph(dt, substeps)
    else:
        sim.step_n(dt, substeps)
    out = {"x": sim.read_positions(), "v": sim.read_velocities(), "C": sim.read_field("C"),
           "F": sim.read_field("F"), "jp": sim.read_field("jp"), "status": sim.read_status()}
    return sim, out
