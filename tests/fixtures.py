"""Procedural fixtures restated from the reference's own tests.

The reference commits no golden vectors; its tests generate inputs
procedurally.  These generators reproduce them in f32 arithmetic:

* ``fill_box`` -- crates/molla-solvers/tests/mpm_gpu.rs:174-200 (and the copy
  in tests/mpm_freefall_regression_gpu.rs:25-48): a jittered lattice driven by
  the 64-bit LCG (6364136223846793005, 1442695040888963407).
* ``points`` -- crates/molla-solvers/src/mpm_particle_order_tests.rs:20-32.
* ``order_config`` -- the ``config(size)`` of the same test file (:7-18).
"""
from __future__ import annotations

import numpy as np

from molla_amd.config import MpmConfig

_MASK = (1 << 64) - 1
f32 = np.float32


def lcg(seed: int):
    s = seed & _MASK

    def nxt():
        nonlocal s
        s = (s * 6364136223846793005 + 1442695040888963407) & _MASK
        return f32(f32(s >> 32) / f32(4294967295.0)) * f32(2.0) - f32(1.0)

    return nxt


def fill_box(lo, hi, per_axis: int, seed: int) -> np.ndarray:
    nxt = lcg(seed)
    lo = [f32(t) for t in lo]
    hi = [f32(t) for t in hi]
    d = [f32((hi[a] - lo[a]) / f32(per_axis)) for a in range(3)]
    out = []
    j = f32(0.05)
    for ix in range(per_axis):
        for iy in range(per_axis):
            for iz in range(per_axis):
                p = []
                for a, i in enumerate((ix, iy, iz)):
                    base = lo[a] + (f32(i) + f32(0.5)) * d[a]
                    p.append(f32(base + (j * d[a]) * nxt()))
                out.append(p)
    return np.asarray(out, dtype=np.float32).reshape(-1, 3)


def points(count: int) -> np.ndarray:
    out = np.zeros((count, 3), np.float32)
    for pid in range(count):
        i = (pid * 131) % count
        out[pid, 0] = f32(-0.4) + f32(i % 7) * f32(0.021) + f32(i % 2) * f32(0.55)
        out[pid, 1] = f32(-0.2) + f32(i // 7 % 7) * f32(0.021)
        out[pid, 2] = f32(-0.3) + f32(i // 49) * f32(0.021)
    return out


def order_config(size: int) -> MpmConfig:
    cfg = MpmConfig.sand()
    cfg.grid_size = size
    cfg.grid_origin = (-1.0, -1.0, -1.0)
    cfg.cell_size = 0.0625
    cfg.max_particles = 513
    cfg.max_rigid = 2
    cfg.youngs = 3000.0
    cfg.particle_volume = float(f32(0.021) ** 3)
    cfg.particle_mass = float(f32(cfg.particle_volume) * f32(1600.0))
    return cfg


def reference_key(cfg, p) -> int:
    """Closed-form key() of mpm_particle_order_tests.rs:34-43 (divides by dx)."""
    t = []
    for a in range(3):
        c = np.floor(f32(f32(p[a]) - f32(cfg.grid_origin[a])) / f32(cfg.cell_size))
        c = min(max(c, 0.0), float(cfg.grid_size - 1))
        t.append(int(c) // 4)
    w = (cfg.grid_size + 3) // 4
    return t[0] + w * (t[1] + w * t[2])


def soil_fixture(count: int):
    """The checkpoint fixture of mpm_checkpoint_tests.rs:3-60: MpmConfig::sand()
    with max_particles 129, max_rigid 2, E 1000, a sand/snow/viscous material
    table, and ``count`` particles whose nine streams follow the particle index.
    Returns (MpmConfig, materials, words as uint32)."""
    from molla_amd.config import Sand, Snow, Viscous

    cfg = MpmConfig.sand().copy(max_particles=129, max_rigid=2, youngs=1000.0)
    materials = [Sand(friction_angle=0.5), Snow(theta_c=0.025, theta_s=0.0075, hardening=10.0),
                 Viscous(velocity_damping=3.0, affine_damping=5.0)]
    vol, mass = f32(cfg.particle_volume), f32(cfg.particle_mass)
    streams = [[] for _ in range(9)]
    for i in range(count):
        f = f32(i)
        streams[0].append([f32(-0.13) + f32(i % 7) * f32(0.021), f32(0.3) + f32(i // 7 % 7) * f32(0.021),
                           f32(-0.07) + f32(i // 49) * f32(0.021), f32(-0.0)])
        streams[1].append([f32(0.013) + f * f32(0.0001), f32(-0.007), f32(0.002), f32(0.0)])
        streams[2].append([0.01, 0.002, 0.0, 0.0, 0.0, -0.02, 0.003, 0.0, 0.004, 0.0, 0.008, 0.0])
        streams[3].append([0.97, 0.001, 0.0, 0.0, 0.0, 1.01, 0.002, 0.0, 0.001, 0.0, 0.98, 0.0])
        streams[4].append([vol * (f32(1.0) + f * f32(0.002))])
        streams[5].append([mass * (f32(1.0) + f * f32(0.001))])
        streams[6].append([f32(-0.1) - f * f32(0.001)])
        streams[8].append([f32(0.005) + f * f32(0.00001), f32(0.01) + f * f32(0.001), f32(0.02), f32(3.0)])
    parts = []
    for s in range(9):
        if s == 7:
            parts.append(np.arange(count, dtype=np.uint32) % 3)
        else:
            parts.append(np.asarray(streams[s], np.float32).reshape(-1).view(np.uint32))
    return cfg, materials, np.concatenate(parts).astype(np.uint32)
