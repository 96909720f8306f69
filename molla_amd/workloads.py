"""Seeded synthetic scenes for the benchmark and tests.

Each builder returns configs plus particle arrays.  All unspecified physical
values are chosen here so that the explicit step is stable (wave CFL
``dt * c_p / dx <= 0.3`` with ``c_p = sqrt((lambda + 2 mu) / rho)``) and are
recorded in the returned ``meta`` dict.

* ``elastic_cube``   -- 8,192 particles, 32^3, dx = 1/32, E = 1e4, nu = 0.2,
  rho = 1, dt = 1e-4, gravity -9.8, sticky walls/floor, 1,000 substeps.
* ``large_scene``    -- 2M particles on a 256^3 grid: half weakly-compressible
  fluid (bulk 1e4), half fixed-corotated elastic blocks (E = 5e4), dt = 5e-5.
* ``batched_scenes`` -- 512 scenes, N ~ logU[1e3, 6.4e4], grid 32^3..96^3
  from each scene's extent (block side x U[2.5, 4.5] + 8 cells),
  E ~ logU[1e3, 1e5], nu ~ U[0.1, 0.4].
* ``diff_scene``     -- 65,536 elastic particles on 64^3 for the adjoint.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

from .config import MODEL_CPUDIFF, Fluid, MpmConfig, Viscous

ELASTIC = Viscous(0.0, 0.0)  # reference's closest elastic row (mpm.wgsl:1412-1419)


def _cp(E: float, nu: float, rho: float) -> float:
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return math.sqrt((lam + 2 * mu) / rho)


def elastic_cube(n: int = 8192, seed: int = 1, grid: int = 32) -> Tuple[MpmConfig, np.ndarray, Dict]:
    rng = np.random.default_rng(seed)
    lo = np.array([0.375, 0.40, 0.375])
    hi = np.array([0.625, 0.65, 0.625])
    x = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    vol = 0.25 ** 3 / n
    rho = 1.0
    cfg = MpmConfig(grid_size=grid, cell_size=1.0 / grid, grid_origin=(0.0, 0.0, 0.0), gravity_y=-9.8,
                    youngs=1.0e4, poisson=0.2, boundary_friction=0.0, particle_mass=rho * vol,
                    particle_volume=vol, max_particles=n, rheology=ELASTIC)
    dt = 1e-4
    meta = {"workload": "elastic_cube", "particles": n, "grid": grid, "dt": dt, "substeps": 1000,
            "rho": rho, "seed": seed, "box": [lo.tolist(), hi.tolist()],
            "cfl": dt * _cp(1e4, 0.2, rho) * grid}
    return cfg, x, meta


def large_scene(n: int = 2_000_000, seed: int = 2, grid: int = 256) -> Tuple[MpmConfig, Dict[str, np.ndarray], Dict]:
    """Half fluid (one pool on the floor), half elastic (four falling blocks)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / grid
    rho = 120.0
    E_el, nu_el, bulk = 5.0e4, 0.2, 1.0e4
    nf = n // 2
    ne = n - nf
    # fluid pool: 96 x 24 x 96 cells above the 3-cell wall layer
    f_lo = np.array([80 * dx, 3 * dx, 80 * dx])
    f_hi = f_lo + np.array([96, 24, 96]) * dx
    xf = f_lo + (f_hi - f_lo) * rng.random((nf, 3))
    vol_f = float(np.prod(f_hi - f_lo) / nf)
    # four elastic blocks of 32^3 cells at seeded positions above the pool
    xs, vs = [], []
    per = [ne // 4 + (1 if i < ne % 4 else 0) for i in range(4)]
    side = 32 * dx
    for i, k in enumerate(per):
        c = np.array([rng.uniform(70, 154) * dx, rng.uniform(40, 60) * dx, rng.uniform(70, 154) * dx])
        xs.append(c + side * rng.random((k, 3)))
        vs.append(np.broadcast_to(np.array([rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)]), (k, 3)))
    xe = np.concatenate(xs)
    ve = np.concatenate(vs)
    vol_e = side ** 3 / per[0]
    x = np.concatenate([xf, xe]).astype(np.float32)
    v = np.concatenate([np.zeros((nf, 3)), ve]).astype(np.float32)
    vol = np.concatenate([np.full(nf, vol_f), np.full(ne, vol_e)]).astype(np.float32)
    mass = (rho * vol).astype(np.float32)
    mat = np.concatenate([np.zeros(nf, np.uint32), np.ones(ne, np.uint32)])
    cfg = MpmConfig(grid_size=grid, cell_size=dx, grid_origin=(0.0, 0.0, 0.0), gravity_y=-9.8,
                    youngs=E_el, poisson=nu_el, boundary_friction=0.0, particle_mass=float(mass[0]),
                    particle_volume=float(vol[0]), max_particles=n,
                    materials=[Fluid(bulk), ELASTIC])
    dt = 5e-5
    meta = {"workload": "large_scene", "particles": n, "grid": grid, "dt": dt, "substeps": 200,
            "rho": rho, "seed": seed, "fluid_bulk": bulk, "elastic_E": E_el, "elastic_nu": nu_el,
            "cfl": dt * max(_cp(E_el, nu_el, rho), math.sqrt(bulk / rho)) / dx}
    return cfg, {"x": x, "v": v, "mass": mass, "volume": vol, "material": mat}, meta


def batched_scenes(n_scenes: int = 512, seed: int = 3, n_lo: float = 1e3, n_hi: float = 6.4e4,
                   g_lo: int = 32, g_hi: int = 96, rho: float = 1000.0,
                   dt: float = 1e-4) -> Tuple[List[MpmConfig], List[np.ndarray], Dict]:
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_scenes))).astype(int)
    # Each scene's domain spans `extent` block sides plus an 8-cell margin,
    # extent ~ U[2.5, 4.5] from a stream of its own (the other draws keep
    # their values), so the grids cover 32^3..96^3.
    extent = np.random.default_rng(seed + 7919).uniform(2.5, 4.5, n_scenes)
    cfgs, xs = [], []
    cfl = 0.0
    for i, N in enumerate(counts):
        a = (N / 8.0) ** (1.0 / 3.0)  # cube side in cells at 8 particles per cell
        G = int(min(max(4 * math.ceil((extent[i] * a + 8) / 4), g_lo), g_hi))
        dx = 1.0 / G
        side = a * dx
        E = float(math.exp(rng.uniform(math.log(1e3), math.log(1e5))))
        nu = float(rng.uniform(0.1, 0.4))
        lo = np.array([0.5 - side / 2, 3 * dx + 0.25 * side, 0.5 - side / 2])
        x = (lo + side * rng.random((N, 3))).astype(np.float32)
        vol = side ** 3 / N
        cfgs.append(MpmConfig(grid_size=G, cell_size=dx, grid_origin=(0.0, 0.0, 0.0), gravity_y=-9.8,
                              youngs=E, poisson=nu, boundary_friction=0.0, particle_mass=rho * vol,
                              particle_volume=vol, max_particles=int(N), rheology=ELASTIC))
        xs.append(x)
        cfl = max(cfl, dt * _cp(E, nu, rho) / dx)
    meta = {"workload": "batched_scenes", "scenes": n_scenes, "particles": int(counts.sum()),
            "dt": dt, "substeps": 500, "rho": rho, "seed": seed, "cfl_max": cfl,
            "extent": extent.tolist(), "grids": [c.grid_size for c in cfgs],
            "grid_min": int(min(c.grid_size for c in cfgs)), "grid_max": int(max(c.grid_size for c in cfgs))}
    return cfgs, xs, meta


def diff_scene(n: int = 65_536, seed: int = 4, grid: int = 64) -> Tuple[MpmConfig, Dict[str, np.ndarray], Dict]:
    """Elastic block dropped onto the floor plane (CPU-diff semantics) so the
    COM loss depends on E through the contact (mpm_adjoint.rs:698-703)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / grid
    rho = 1000.0
    side = 16 * dx
    lo = np.array([0.5 - side / 2, 3.5 * dx, 0.5 - side / 2])
    x = (lo + side * rng.random((n, 3))).astype(np.float32)
    v = np.tile(np.array([0.2, -1.5, 0.1], np.float32), (n, 1))
    vol = side ** 3 / n
    E = np.exp(rng.uniform(math.log(5e3), math.log(2e4), n)).astype(np.float32)
    cfg = MpmConfig(grid_size=grid, cell_size=dx, grid_origin=(0.0, 0.0, 0.0), gravity_y=-9.8,
                    youngs=1.0e4, poisson=0.3, particle_mass=rho * vol, particle_volume=vol,
                    max_particles=n, model=MODEL_CPUDIFF, floor_y=3 * dx, floor_friction=0.5,
                    rheology=ELASTIC)
    dt = 5e-4
    target = np.array([0.52, 0.05, 0.49])
    meta = {"workload": "diff_scene", "particles": n, "grid": grid, "dt": dt, "substeps": 64,
            "checkpoint_every": 16, "rho": rho, "seed": seed, "target": target.tolist(),
            "cfl": dt * _cp(2e4, 0.3, rho) / dx}
    return cfg, {"x": x, "v": v, "E": E, "target": target}, meta


def batched_diff_scenes(n_scenes: int = 128, seed: int = 5, n_lo: float = 4e3, n_hi: float = 3.2e4,
                        g_lo: int = 48, g_hi: int = 80, rho: float = 1000.0,
                        dt: float = 5e-4) -> Tuple[List[MpmConfig], Dict[str, List[np.ndarray]], Dict]:
    """Batched system identification (CPU-diff semantics): per scene an
    elastic block of N ~ logU[4k, 32k] particles dropped onto the floor of a
    48^3..80^3 grid; the ground-truth Young's modulus E_true ~ logU[5e3, 2e4]
    per scene and the starting guess E_0 = E_true * exp(U[-0.3, 0.3])."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_scenes))).astype(int)
    t = (np.log(counts) - math.log(n_lo)) / (math.log(n_hi) - math.log(n_lo))
    cfgs, xs, vs, e_true, e0 = [], [], [], [], []
    cfl = 0.0
    for i, N in enumerate(counts):
        G = int(4 * round((g_lo + (g_hi - g_lo) * t[i]) / 4))
        dx = 1.0 / G
        side = (N / 8.0) ** (1.0 / 3.0) * dx  # 8 particles per cell
        lo = np.array([0.5 - side / 2, 3.5 * dx, 0.5 - side / 2])
        x = (lo + side * rng.random((N, 3))).astype(np.float32)
        v = np.tile(np.array([rng.uniform(-0.3, 0.3), -1.0, rng.uniform(-0.3, 0.3)], np.float32), (N, 1))
        vol = side ** 3 / N
        et = float(math.exp(rng.uniform(math.log(5e3), math.log(2e4))))
        cfgs.append(MpmConfig(grid_size=G, cell_size=dx, grid_origin=(0.0, 0.0, 0.0), gravity_y=-9.8,
                              youngs=et, poisson=0.3, particle_mass=rho * vol, particle_volume=vol,
                              max_particles=int(N), model=MODEL_CPUDIFF, floor_y=3 * dx, floor_friction=0.5,
                              rheology=ELASTIC))
        xs.append(x)
        vs.append(v)
        e_true.append(et)
        e0.append(et * math.exp(rng.uniform(-0.3, 0.3)))
        cfl = max(cfl, dt * _cp(2e4, 0.3, rho) / dx)
    meta = {"workload": "batched_diff_scenes", "scenes": n_scenes, "particles": int(counts.sum()),
            "counts": counts.tolist(), "dt": dt, "substeps": 32, "checkpoint_every": 16, "rho": rho,
            "seed": seed, "cfl_max": cfl, "e_true": e_true, "e0": e0,
            "grid_min": int(min(c.grid_size for c in cfgs)), "grid_max": int(max(c.grid_size for c in cfgs))}
    return cfgs, {"x": xs, "v": vs}, meta
