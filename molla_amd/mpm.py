"""Host-side mirror of molla's ``MpmSimulation`` over the C ABI.

Method names, argument meaning and error behaviour follow
``crates/molla-solvers/src/mpm.rs`` (``new`` :834, ``set_particles`` :1451,
``set_particles_with_velocity`` :1458-1519, ``step``/``step_n`` :2232-2272,
``read_positions_with`` :2718, ``read_velocities_with`` :2760,
``read_grid_mass_momentum`` :2792, ``read_active_grid_status_with`` :1418,
``set_p2g_mode`` :1642, ``set_materials`` :1570) so that tests read like the
reference's own.  Errors surface as ``MpmError`` carrying the library's
message (the reference returns ``Error::Solver`` / ``Error::Gpu``).

The compute runs in ``libmolla_mpm.so`` (HIP, gfx950).  There is no CPU
fallback: if the library is missing or no device is present, construction
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, List, Optional, Sequence

import numpy as np

from .config import MODEL_SHADER, MpmConfig, material_row

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmolla_mpm.so")

MAX_MATERIALS = 8

P2G_FUSED = 0
P2G_GLOBAL_ATOMIC = 1
P2G_DETERMINISTIC = 2
_P2G_NAMES = {"fused": P2G_FUSED, "two_pass": P2G_FUSED, "global_atomic": P2G_GLOBAL_ATOMIC,
              "legacy": P2G_GLOBAL_ATOMIC, "deterministic": P2G_DETERMINISTIC}

FIELDS = {"position": (0, 3), "velocity": (1, 3), "C": (2, 9), "F": (3, 9), "jp": (4, 1),
          "mass": (5, 1), "volume": (6, 1), "youngs": (7, 1), "material": (8, 1)}

STAGES = ["sort", "p2g", "grid", "g2p2g", "g2p", "diff_forward", "diff_backward"]
PROFILE_STAGES, PROFILE_CALLS = 1, 2


class MpmError(RuntimeError):
    """Raised when a library call fails (status code + last-error message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Material(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("rheology", "dp_alpha", "theta_c", "theta_s", "hardening",
                                         "velocity_damping", "affine_damping", "bulk")]


class SceneConfig(C.Structure):
    _fields_ = [
        ("grid_size", C.c_uint32),
        ("cell_size", C.c_float),
        ("grid_origin", C.c_float * 3),
        ("grid_window_cells", C.c_int32 * 3),
        ("gravity_y", C.c_float),
        ("youngs", C.c_float),
        ("poisson", C.c_float),
        ("boundary_friction", C.c_float),
        ("particle_mass", C.c_float),
        ("particle_volume", C.c_float),
        ("max_particles", C.c_uint32),
        ("model", C.c_uint32),
        ("floor_y", C.c_float),
        ("floor_friction", C.c_float),
        ("n_materials", C.c_uint32),
        ("materials", Material * MAX_MATERIALS),
    ]


class LaunchConfig(C.Structure):
    """mpm_launch_config."""
    _fields_ = [("particle_blocks", C.c_uint32), ("grid_blocks", C.c_uint32), ("sort_every", C.c_uint32),
                ("particles_per_thread", C.c_uint32), ("segment_tiles", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class StatusInfo(C.Structure):
    _fields_ = [("fault_bits", C.c_uint32), ("completed_gathers", C.c_uint32),
                ("active_tiles", C.c_uint32), ("sorts", C.c_uint32), ("escapes", C.c_uint32),
                ("fallback_chunks", C.c_uint32), ("chunks", C.c_uint32)]


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "mpm_abi_version": (C.c_int, []),
    "mpm_create": (C.c_int, [C.POINTER(SceneConfig), C.POINTER(C.c_void_p)]),
    "mpm_create_batch": (C.c_int, [C.POINTER(SceneConfig), C.c_uint32, C.POINTER(C.c_void_p)]),
    "mpm_destroy": (C.c_int, [C.c_void_p]),
    "mpm_last_error": (C.c_char_p, [C.c_void_p]),
    "mpm_set_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_set_particles_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm_write_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]),
    "mpm_read_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]),
    "mpm_read_positions": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_read_velocities": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_set_materials": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Material), C.c_uint32]),
    "mpm_set_scene_params": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(SceneConfig)]),
    "mpm_set_p2g_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_set_sort_interval": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mpm_set_time_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mpm_set_write_vc": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_set_launch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "mpm_set_launch_config": (C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    "mpm_get_launch_config": (C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    "mpm_autotune_config": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32, C.POINTER(LaunchConfig)]),
    "mpm_step_n": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32]),
    "mpm_step_n_graph": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32]),
    "mpm_synchronize": (C.c_int, [C.c_void_p]),
    "mpm_read_grid_mass_momentum": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mpm_read_status": (C.c_int, [C.c_void_p, C.POINTER(StatusInfo)]),
    "mpm_reset_status": (C.c_int, [C.c_void_p]),
    "mpm_sort_keys_debug": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mpm_sort_keys_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mpm_read_field_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]),
    "mpm_create_sparse": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "mpm_reserve_sparse_tiles": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mpm_grid_storage_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "mpm_read_grid_storage_mass_momentum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_read_active_tile_ids": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "mpm_diff_backward_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "mpm_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_read_profile": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm_reset_profile": (C.c_int, [C.c_void_p]),
    "mpm_diff_forward": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32, C.c_uint32]),
    "mpm_diff_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "mpm_diff_read_adjoint": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]),
    "mpm_autotune": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
}


def lib():
    """Load libmolla_mpm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpmError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        from .checkpoint import SIGS as _CKPT_SIGS
        from .rigid import SIGS as _RIGID_SIGS

        for name, (res, args) in {**_SIGS, **_CKPT_SIGS, **_RIGID_SIGS}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> List[str]:
    from .checkpoint import SIGS as _CKPT_SIGS
    from .rigid import SIGS as _RIGID_SIGS

    return list(_SIGS) + list(_CKPT_SIGS) + list(_RIGID_SIGS)


def scene_config(cfg: MpmConfig, max_particles: Optional[int] = None) -> SceneConfig:
    s = SceneConfig()
    s.grid_size = int(cfg.grid_size)
    s.cell_size = float(cfg.cell_size)
    for a in range(3):
        s.grid_origin[a] = float(cfg.grid_origin[a])
        s.grid_window_cells[a] = int(cfg.grid_window_cells[a])
    s.gravity_y = float(cfg.gravity_y)
    s.youngs = float(cfg.youngs)
    s.poisson = float(cfg.poisson)
    s.boundary_friction = float(cfg.boundary_friction)
    s.particle_mass = float(cfg.particle_mass)
    s.particle_volume = float(cfg.particle_volume)
    s.max_particles = int(cfg.max_particles if max_particles is None else max_particles)
    s.model = int(cfg.model)
    s.floor_y = float(cfg.floor_y)
    s.floor_friction = float(cfg.floor_friction)
    rows = cfg.material_rows()
    s.n_materials = len(rows)
    for i, r in enumerate(rows):
        for name, val in zip(Material._fields_, r):
            setattr(s.materials[i], name[0], float(val))
    return s


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class MpmSimulation:
    """One MPM simulation (or a batch of independent scenes) on the GPU."""

    def __init__(self, cfg, max_particles: Optional[int] = None, tile_capacity: Optional[int] = None):
        cfgs = list(cfg) if isinstance(cfg, (list, tuple)) else [cfg]
        if max_particles is not None:  # the capacity is cfg.max_particles, as in the reference
            cfgs = [c.copy(max_particles=int(max_particles)) for c in cfgs]
        self.cfgs: List[MpmConfig] = cfgs
        self.cfg = cfgs[0]
        arr = (SceneConfig * len(cfgs))(*[scene_config(c, max_particles) for c in cfgs])
        h = C.c_void_p()
        L = lib()
        self.tile_capacity = tile_capacity
        if tile_capacity is not None:
            if len(cfgs) != 1:
                raise MpmError(1, "a sparse tile pool holds one scene")
            rc = L.mpm_create_sparse(arr, int(tile_capacity), C.byref(h))
        else:
            rc = L.mpm_create_batch(arr, len(cfgs), C.byref(h))
        self._h = h
        if rc != 0:
            msg = L.mpm_last_error(h).decode() if h.value else "create failed"
            if h.value:
                L.mpm_destroy(h)
            self._h = None
            raise MpmError(rc, msg)
        self.counts = [0] * len(cfgs)
        iters = {int(getattr(c.time_step, "jacobi_iters", 0) or 0) for c in cfgs}
        if len(iters) != 1:
            self.close()
            raise MpmError(1, "all scenes of a batch must share one time step")
        if iters != {0}:
            self.set_time_step(cfgs[0].time_step)

    # -- construction helpers mirroring the reference names --
    @classmethod
    def new(cls, cfg: MpmConfig) -> "MpmSimulation":
        return cls(cfg)

    @classmethod
    def batch(cls, cfgs: Sequence[MpmConfig]) -> "MpmSimulation":
        return cls(list(cfgs))

    @classmethod
    def new_sparse(cls, cfg: MpmConfig, tile_capacity: int) -> "MpmSimulation":
        """``MpmSimulation::new_sparse`` (mpm.rs:848): grid storage for
        `tile_capacity` 4^3 tiles."""
        return cls(cfg, tile_capacity=tile_capacity)

    # -- sparse tile pool --
    def active_grid_enabled(self) -> bool:
        return self.tile_capacity is not None

    def reserve_sparse_tiles(self, tile_capacity: int):
        self._check(lib().mpm_reserve_sparse_tiles(self._h, int(tile_capacity)))
        self.tile_capacity = max(self.tile_capacity or 0, int(tile_capacity))

    def grid_storage_bytes(self) -> int:
        b = C.c_uint64(0)
        self._check(lib().mpm_grid_storage_bytes(self._h, C.byref(b)))
        return int(b.value)

    def read_grid_storage_mass_momentum(self):
        """Physical grid storage of the last P2G: (mass (cells,), momentum (cells, 3))."""
        tiles = self.tile_capacity if self.tile_capacity is not None else sum(
            ((c.grid_size + 3) // 4) ** 3 for c in self.cfgs)
        cells = 64 * tiles
        mass = np.zeros(cells, np.float32)
        mom = np.zeros((cells, 3), np.float32)
        self._check(lib().mpm_read_grid_storage_mass_momentum(self._h, _ptr(mass), _ptr(mom), cells))
        return mass, mom

    def read_active_tile_ids(self) -> np.ndarray:
        """World tile ids of the pool's slots, in slot order."""
        n = C.c_uint32(0)
        ids = np.zeros(max(int(self.tile_capacity or 0), 1), np.uint32)
        self._check(lib().mpm_read_active_tile_ids(self._h, _ptr(ids), len(ids), C.byref(n)))
        return ids[: n.value].copy()

    read_active_tile_ids_with = read_active_tile_ids

    def read_active_grid_stats(self) -> dict:
        """``read_active_grid_stats_with`` (mpm.rs:1397): the status of the
        last substep; refuses while a fault is latched."""
        st = self.read_status()
        if st["faulted"]:
            raise MpmError(1, f"MPM grid fault after {st['completed_gathers']} gathers; restore or reset required")
        return st

    read_active_grid_stats_with = read_active_grid_stats

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().mpm_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != 0:
            raise MpmError(rc, lib().mpm_last_error(self._h).decode())

    @property
    def n(self) -> int:
        return int(sum(self.counts))

    # -- particles --
    def set_particles(self, positions, velocities=None):
        """``set_particles`` / ``set_particles_with_velocity`` (mpm.rs:1451-1519)."""
        pos = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        vel = None
        if velocities is not None and len(velocities) > 0:
            vel = np.ascontiguousarray(np.asarray(velocities, np.float32).reshape(-1, 3))
            if vel.shape != pos.shape:
                raise MpmError(1, f"velocities len ({len(vel)}) != positions len ({len(pos)})")
        if len(self.cfgs) != 1:
            raise MpmError(1, "use set_particles_batch on a batch")
        self._check(lib().mpm_set_particles(self._h, _ptr(pos), _ptr(vel), pos.shape[0]))
        self.counts = [pos.shape[0]]

    set_particles_with_velocity = set_particles

    def set_particles_batch(self, positions: Sequence[np.ndarray], velocities=None):
        pos = [np.asarray(p, np.float32).reshape(-1, 3) for p in positions]
        counts = np.asarray([p.shape[0] for p in pos], np.uint32)
        allp = np.ascontiguousarray(np.concatenate(pos) if pos else np.zeros((0, 3), np.float32))
        allv = None
        if velocities is not None:
            allv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float32).reshape(-1, 3)
                                                        for v in velocities]))
        self._check(lib().mpm_set_particles_batch(self._h, _ptr(counts), _ptr(allp), _ptr(allv)))
        self.counts = counts.tolist()

    def write_field(self, name: str, data):
        fid, nc = FIELDS[name]
        dt = np.uint32 if name == "material" else np.float32
        a = np.ascontiguousarray(np.asarray(data, dt).reshape(self.n, nc) if nc > 1
                                 else np.asarray(data, dt).reshape(self.n))
        self._check(lib().mpm_write_field(self._h, fid, _ptr(a), self.n))

    def read_field(self, name: str) -> np.ndarray:
        fid, nc = FIELDS[name]
        dt = np.uint32 if name == "material" else np.float32
        out = np.zeros((self.n, nc) if nc > 1 else (self.n,), dt)
        self._check(lib().mpm_read_field(self._h, fid, _ptr(out), self.n))
        if nc == 9:
            return out.reshape(self.n, 3, 3)
        return out

    def read_field_device(self, name: str, out_dev: int):
        """The field into a caller-owned device buffer (address `out_dev`, n x
        components words, original order); ordered on the handle's stream."""
        fid, _ = FIELDS[name]
        self._check(lib().mpm_read_field_device(self._h, fid, C.c_void_p(out_dev), self.n))

    def read_positions(self) -> np.ndarray:
        return self.read_field("position")

    def read_velocities(self) -> np.ndarray:
        return self.read_field("velocity")

    read_positions_with = read_positions
    read_velocities_with = read_velocities

    # -- materials / parameters --
    def set_materials(self, rheologies: Iterable, scene: int = 0):
        rheologies = list(rheologies)
        rows = [material_row(r) for r in rheologies]
        arr = (Material * len(rows))(*[Material(*[float(t) for t in r]) for r in rows])
        self._check(lib().mpm_set_materials(self._h, scene, arr, len(rows)))
        self.cfgs[scene] = self.cfgs[scene].copy(materials=rheologies)
        self.cfg = self.cfgs[0]

    def set_particle_materials(self, ids):
        self.write_field("material", ids)

    def set_scene_params(self, cfg: MpmConfig, scene: int = 0):
        sc = scene_config(cfg, self.cfgs[scene].max_particles)
        self._check(lib().mpm_set_scene_params(self._h, scene, C.byref(sc)))
        self.cfgs[scene] = cfg

    def set_p2g_mode(self, mode):
        if isinstance(mode, bool):  # reference: set_p2g_mode(two_pass: bool)
            mode = P2G_FUSED if mode else P2G_GLOBAL_ATOMIC
        if isinstance(mode, str):
            mode = _P2G_NAMES[mode]
        self._check(lib().mpm_set_p2g_mode(self._h, int(mode)))

    def set_deterministic_p2g(self, enabled: bool):
        self.set_p2g_mode(P2G_DETERMINISTIC if enabled else P2G_FUSED)

    def set_time_step(self, ts):
        """``set_time_step`` (mpm.rs:1632): ``Explicit()`` or ``Implicit(jacobi_iters)``."""
        iters = int(getattr(ts, "jacobi_iters", 0) or 0)
        self._check(lib().mpm_set_time_step(self._h, iters))
        for i, c in enumerate(self.cfgs):
            self.cfgs[i] = c.copy(time_step=ts)
        self.cfg = self.cfgs[0]

    def set_sort_interval(self, every: int):
        self._check(lib().mpm_set_sort_interval(self._h, int(every)))

    def set_write_vc(self, every_substep: bool):
        self._check(lib().mpm_set_write_vc(self._h, 1 if every_substep else 0))

    def set_launch(self, particle_blocks: int = 0, grid_blocks: int = 0):
        self._check(lib().mpm_set_launch(self._h, int(particle_blocks), int(grid_blocks)))

    def launch_config(self) -> dict:
        c = LaunchConfig()
        self._check(lib().mpm_get_launch_config(self._h, C.byref(c)))
        return c.as_dict()

    def set_launch_config(self, **kw):
        """Change any of particle_blocks, grid_blocks, sort_every,
        particles_per_thread (1, 2) and segment_tiles (the chunk rule: 4 or 6
        greedy rows of at most that many tiles, 8 aligned eight- or four-tile
        segments)."""
        cur = self.launch_config()
        unknown = set(kw) - set(cur)
        if unknown:
            raise TypeError(f"unknown launch settings {sorted(unknown)}")
        cur.update(kw)
        self._check(lib().mpm_set_launch_config(self._h, C.byref(LaunchConfig(**{k: int(v) for k, v in cur.items()}))))

    # -- stepping --
    def step(self, dt: float):
        self.step_n(dt, 1)

    def step_n(self, dt: float, substeps: int):
        self._check(lib().mpm_step_n(self._h, float(dt), int(substeps)))

    # -- rigid bodies (mpm.rs:1521-1566, mpm_prescribed_solids.rs:91-140) --
    def set_rigid_spheres(self, spheres):
        """``set_rigid_spheres``: [(center xyz, radius, velocity xyz), ...]."""
        spheres = list(spheres)
        if len(spheres) > self.cfg.max_rigid:
            raise MpmError(2, f"set_rigid_spheres: n={len(spheres)} > max_rigid={self.cfg.max_rigid}")
        a = np.asarray([[*c, r, *v] for c, r, v in spheres], np.float32).reshape(-1, 7)
        a = np.ascontiguousarray(a)
        self._check(lib().mpm_set_rigid_spheres(self._h, _ptr(a), len(spheres)))
        self.rigid_spheres, self.prescribed_solids = spheres, []

    def set_prescribed_solids(self, solids):
        """``set_prescribed_solids``: replace all prescribed solids (explicit stepping only)."""
        from .rigid import PrescribedSolid, solid_struct

        solids = list(solids)
        if len(solids) > self.cfg.max_rigid:
            raise MpmError(1, "MPM prescribed solid capacity exceeded")
        arr = (PrescribedSolid * max(len(solids), 1))(*[solid_struct(b) for b in solids])
        self._check(lib().mpm_set_prescribed_solids(self._h, arr, len(solids)))
        self.rigid_spheres, self.prescribed_solids = [], solids

    def read_rigid_wrenches(self):
        """``read_rigid_wrenches_with``: [(force xyz, torque xyz), ...] of the last substep."""
        n = len(getattr(self, "rigid_spheres", [])) + len(getattr(self, "prescribed_solids", []))
        out = np.zeros((n, 6), np.float32)
        self._check(lib().mpm_read_rigid_wrenches(self._h, _ptr(out), n))
        return [(tuple(map(float, w[:3])), tuple(map(float, w[3:]))) for w in out]

    read_rigid_wrenches_with = read_rigid_wrenches

    # -- soil checkpoint (mpm_checkpoint.rs:388-541) --
    def checkpoint_soil(self):
        """``checkpoint_soil_with``: an ``MpmSoilCheckpoint`` of the particle state."""
        from .checkpoint import checkpoint_soil

        return checkpoint_soil(self)

    def restore_soil(self, checkpoint):
        """``restore_soil_with``: replace the particles with a checkpoint's."""
        from .checkpoint import restore_soil

        restore_soil(self, checkpoint)

    checkpoint_soil_with = checkpoint_soil
    restore_soil_with = restore_soil

    def step_n_graph(self, dt: float, substeps: int):
        self._check(lib().mpm_step_n_graph(self._h, float(dt), int(substeps)))

    def synchronize(self):
        self._check(lib().mpm_synchronize(self._h))

    # -- reads --
    def read_grid_mass_momentum(self, scene: int = 0):
        G = int(self.cfgs[scene].grid_size)
        m = np.zeros(G ** 3, np.float32)
        p = np.zeros((G ** 3, 3), np.float32)
        self._check(lib().mpm_read_grid_mass_momentum(self._h, scene, _ptr(m), _ptr(p)))
        return m, p

    def read_status(self) -> dict:
        st = StatusInfo()
        self._check(lib().mpm_read_status(self._h, C.byref(st)))
        return {"fault_bits": st.fault_bits, "overflow": bool(st.fault_bits & 1),
                "numerical_fault": bool(st.fault_bits & 2), "completed_gathers": st.completed_gathers,
                "active_tiles": st.active_tiles, "sorts": st.sorts, "escapes": st.escapes,
                "fallback_chunks": st.fallback_chunks, "chunks": st.chunks,
                "faulted": st.fault_bits != 0}

    read_active_grid_status = read_status

    def reset_status(self):
        self._check(lib().mpm_reset_status(self._h))

    def sort_keys_debug(self, scene: int = 0, capacity: Optional[int] = None):
        cnt = self.counts[scene]
        cap = cnt if capacity is None else int(capacity)
        keys = np.zeros(cap, np.uint32)
        ids = np.zeros(cap, np.uint32)
        self._check(lib().mpm_sort_keys_debug(self._h, scene, cap, _ptr(keys), _ptr(ids)))
        return keys, ids

    def sort_keys_noalloc(self, keys_dev: int, ids_dev: int, capacity: int, scene: int = 0):
        """The key query into caller-owned device buffers (addresses of
        `capacity` u32 each), with no allocation; ordered on the handle's
        stream (``synchronize`` waits for it)."""
        self._check(lib().mpm_sort_keys_noalloc(self._h, scene, int(capacity), C.c_void_p(keys_dev),
                                                 C.c_void_p(ids_dev)))

    # -- profiling --
    def set_profiling(self, enabled: bool, stages: bool = True, calls: bool = True):
        """HIP events per launch group of the step stages (`stages`) and around
        whole diff_forward / diff_backward calls (`calls`)."""
        flags = (PROFILE_STAGES if stages else 0) | (PROFILE_CALLS if calls else 0)
        self._check(lib().mpm_set_profiling(self._h, flags if enabled else 0))

    def read_profile(self) -> dict:
        ms = np.zeros(len(STAGES), np.float32)
        n = np.zeros(len(STAGES), np.uint32)
        self._check(lib().mpm_read_profile(self._h, _ptr(ms), _ptr(n)))
        return {s: (float(ms[i]), int(n[i])) for i, s in enumerate(STAGES)}

    def reset_profile(self):
        self._check(lib().mpm_reset_profile(self._h))

    def autotune(self, dt: float, trial_substeps: int = 10) -> dict:
        """Pick the fastest launch configuration (particles per thread and
        segment length, launch sizes, sort interval) for this scene mix; the
        particle state is restored afterwards."""
        c = LaunchConfig()
        self._check(lib().mpm_autotune_config(self._h, float(dt), int(trial_substeps), C.byref(c)))
        return c.as_dict()

    # -- differentiable path --
    def diff_forward(self, dt: float, steps: int, checkpoint_every: int = 16):
        self._check(lib().mpm_diff_forward(self._h, float(dt), int(steps), int(checkpoint_every)))

    def diff_backward(self, grad_x_final, grad_v_final=None):
        n = self.n
        gx = np.ascontiguousarray(np.asarray(grad_x_final, np.float32).reshape(n, 3))
        gv = None if grad_v_final is None else np.ascontiguousarray(
            np.asarray(grad_v_final, np.float32).reshape(n, 3))
        ox = np.zeros((n, 3), np.float32)
        ov = np.zeros((n, 3), np.float32)
        oe = np.zeros(n, np.float32)
        ml = np.zeros(2, np.float64)
        self._check(lib().mpm_diff_backward(self._h, _ptr(gx), _ptr(gv), _ptr(ox), _ptr(ov),
                                            _ptr(oe), _ptr(ml)))
        nu = float(self.cfgs[0].poisson)
        g_youngs = ml[0] / (2 * (1 + nu)) + ml[1] * nu / ((1 + nu) * (1 - 2 * nu))
        # per-scene dJ/dE (scene-wide E): the sum of the scene's per-particle dJ/dE_p
        bounds = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        e_scene = np.array([np.sum(oe[bounds[i]:bounds[i + 1]], dtype=np.float64)
                            for i in range(len(self.counts))])
        return {"x": ox, "v": ov, "E": oe, "mu": float(ml[0]), "lambda": float(ml[1]),
                "youngs": float(g_youngs), "youngs_per_scene": e_scene}

    def diff_backward_device(self, grad_x_final_dev: int, grad_v_final_dev: int = 0, grad_x0_dev: int = 0,
                             grad_v0_dev: int = 0, grad_E_dev: int = 0, grad_mu_lambda_dev: int = 0):
        """diff_backward with device addresses for every array (0 = none):
        the gradients stay in HBM; ordered on the handle's stream."""
        def ptr(v):
            return C.c_void_p(v) if v else None
        self._check(lib().mpm_diff_backward_device(self._h, ptr(grad_x_final_dev), ptr(grad_v_final_dev),
                                                   ptr(grad_x0_dev), ptr(grad_v0_dev), ptr(grad_E_dev),
                                                   ptr(grad_mu_lambda_dev)))

    def read_adjoint(self, name: str):
        """dJ/d(initial field) of the last diff_backward, original particle order."""
        fid, width = FIELDS[name]
        out = np.zeros((self.n, width) if width > 1 else self.n, np.float32)
        self._check(lib().mpm_diff_read_adjoint(self._h, fid, _ptr(out), self.n))
        return out
