"""Soil particle checkpoints in the reference's "MOLLSOIL" byte format.

Mirrors ``MpmSoilCheckpoint`` and ``MpmSimulation::checkpoint_soil_with`` /
``restore_soil_with`` (crates/molla-solvers/src/mpm_checkpoint.rs:137-541).
The byte codec and its BLAKE3 digest run in ``libmolla_mpm.so``
(``mpm_soil_*``); the capture and the restore are device calls
(``mpm_checkpoint_capture`` / ``mpm_checkpoint_restore``).  This module turns
an ``MpmConfig`` into the 21 configuration words (``config_words``,
mpm_checkpoint.rs:594-619) and back, and checks configuration compatibility on
restore, which the reference does on the host side too.

Streams of ``words`` (``STRIDES``, mpm_checkpoint.rs:82), 40 u32 per particle,
stream-major, particles in slot (original) order: 0 position + w, 1 velocity
+ w, 2 C and 3 F as three padded columns (word 4j+i = M_ij), 4 reference
volume, 5 mass, 6 log-Jp, 7 material id (u32), 8 tire compaction.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Sequence

import numpy as np

from .config import (MAX_MPM_MATERIALS, MODEL_SHADER, Explicit, Implicit, MpmConfig, Sand, Snow,
                     Viscous, material_row)

HEADER_BYTES = 44
WORDS_PER_PARTICLE = 40
CONFIG_WORDS = 21
STRIDES = (4, 4, 12, 12, 1, 1, 1, 1, 4)


class SoilMeta(C.Structure):
    _fields_ = [("grid_window_cells", C.c_int32 * 3), ("config", C.c_uint32 * CONFIG_WORDS),
                ("n_materials", C.c_uint32), ("materials", (C.c_uint32 * 4) * MAX_MPM_MATERIALS),
                ("count", C.c_uint32)]


SIGS = {
    "mpm_soil_encoded_bytes": (C.c_uint64, [C.POINTER(SoilMeta)]),
    "mpm_soil_encode": (C.c_int, [C.POINTER(SoilMeta), C.c_void_p, C.c_void_p, C.c_uint64]),
    "mpm_soil_decode_meta": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(SoilMeta)]),
    "mpm_soil_decode": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(SoilMeta), C.c_void_p,
                                  C.c_uint64]),
    "mpm_soil_validate": (C.c_int, [C.POINTER(SoilMeta), C.c_void_p]),
    "mpm_soil_validate_window": (C.c_int, [C.POINTER(SoilMeta), C.c_void_p]),
    "mpm_blake3": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "mpm_soil_last_error": (C.c_char_p, []),
    "mpm_checkpoint_capture": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_checkpoint_restore": (C.c_int, [C.c_void_p, C.POINTER(SoilMeta), C.c_void_p, C.c_void_p,
                                         C.c_uint32]),
}


def _lib():
    from .mpm import lib

    return lib()


def _bits(v) -> int:
    return int(np.float32(v).view(np.uint32))


def _f32(w: int) -> float:
    return float(np.uint32(w).view(np.float32))


def rheology_words(r) -> List[int]:
    """``rheology_words`` (mpm_checkpoint.rs:544-557)."""
    if isinstance(r, Sand):
        return [0, _bits(r.friction_angle), 0, 0]
    if isinstance(r, Snow):
        return [1, _bits(r.theta_c), _bits(r.theta_s), _bits(r.hardening)]
    if isinstance(r, Viscous):
        return [2, _bits(r.velocity_damping), _bits(r.affine_damping), 0]
    raise CheckpointError(f"rheology {r!r} has no checkpoint encoding")


def rheology_from_words(w: Sequence[int]):
    """Inverse of ``rheology_words`` on a canonical encoding (``Words::rheology``, :638-660)."""
    kind, a, b, c = (int(v) for v in w)
    if kind == 0:
        return Sand(friction_angle=_f32(a))
    if kind == 1:
        return Snow(theta_c=_f32(a), theta_s=_f32(b), hardening=_f32(c))
    if kind == 2:
        return Viscous(velocity_damping=_f32(a), affine_damping=_f32(b))
    raise CheckpointError("unsupported or noncanonical rheology encoding")


def config_words(cfg: MpmConfig) -> List[int]:
    """``config_words`` (mpm_checkpoint.rs:594-619)."""
    iters = getattr(cfg.time_step, "jacobi_iters", None)
    ts = [0, 0] if iters is None else [1, int(iters)]
    floats = (cfg.cell_size, *cfg.grid_origin, cfg.gravity_y, cfg.youngs, cfg.poisson, cfg.friction_angle,
              cfg.cohesion, cfg.boundary_friction, cfg.particle_mass, cfg.particle_volume)
    return ([int(cfg.grid_size) & 0xFFFFFFFF, int(cfg.max_particles) & 0xFFFFFFFF,
             int(cfg.max_rigid) & 0xFFFFFFFF] + [_bits(v) for v in floats] + rheology_words(cfg.rheology) + ts)


def config_from_words(w: Sequence[int]) -> MpmConfig:
    """``Words::config`` (mpm_checkpoint.rs:662-699)."""
    w = [int(v) for v in w]
    f = [_f32(v) for v in w[3:15]]
    ts = Explicit() if w[19] == 0 else Implicit(jacobi_iters=w[20])
    return MpmConfig(grid_size=w[0], max_particles=w[1], max_rigid=w[2], cell_size=f[0],
                     grid_origin=(f[1], f[2], f[3]), gravity_y=f[4], youngs=f[5], poisson=f[6],
                     friction_angle=f[7], cohesion=f[8], boundary_friction=f[9], particle_mass=f[10],
                     particle_volume=f[11], rheology=rheology_from_words(w[15:19]), time_step=ts)


class CheckpointError(RuntimeError):
    """``Error::Solver("MPM soil checkpoint: ...")``."""


def _codec_check(rc: int):
    if rc != 0:
        raise CheckpointError(_lib().mpm_soil_last_error().decode())


@dataclasses.dataclass
class MpmSoilCheckpoint:
    """Lossless particle state (mpm_checkpoint.rs:134-147): configuration,
    integer grid window, material table (empty: the configuration's rheology is
    authoritative), particle count and the 40-word-per-particle streams."""

    cfg: MpmConfig
    grid_window: tuple
    material_table: list
    count: int
    words: np.ndarray

    # -- accessors named as in the reference --
    def config(self) -> MpmConfig:
        return self.cfg

    def grid_window_cells(self) -> tuple:
        return tuple(self.grid_window)

    def particle_count(self) -> int:
        return int(self.count)

    def materials(self) -> list:
        return list(self.material_table)

    def particle_bytes(self) -> int:
        return int(self.words.size) * 4

    def stream(self, s: int) -> np.ndarray:
        """Words of stream ``s`` (``stream``, :269-272)."""
        off = sum(STRIDES[:s]) * self.count
        return self.words[off:off + STRIDES[s] * self.count]

    def meta(self) -> SoilMeta:
        m = SoilMeta()
        for a in range(3):
            m.grid_window_cells[a] = int(self.grid_window[a])
        for k, v in enumerate(config_words(self.cfg)):
            m.config[k] = v
        if len(self.material_table) > MAX_MPM_MATERIALS:
            raise CheckpointError("material table exceeds capacity")
        m.n_materials = len(self.material_table)
        for i, r in enumerate(self.material_table):
            for k, v in enumerate(rheology_words(r)):
                m.materials[i][k] = v
        m.count = int(self.count)
        return m

    def _words(self) -> np.ndarray:
        w = np.ascontiguousarray(self.words, dtype=np.uint32)
        if w.size != int(self.count) * WORDS_PER_PARTICLE:
            raise CheckpointError("invalid configuration or state length")
        return w

    def validate(self):
        """``validate`` (:281-330)."""
        w = self._words()
        _codec_check(_lib().mpm_soil_validate(C.byref(self.meta()), w.ctypes.data_as(C.c_void_p)))

    def to_bytes(self) -> bytes:
        """``to_bytes`` (:185-214): version 1, or 2 with a non-zero window."""
        m = self.meta()
        w = np.ascontiguousarray(self.words, dtype=np.uint32)
        if w.size < int(self.count) * WORDS_PER_PARTICLE:
            raise CheckpointError("particle words shorter than the count")
        n = _lib().mpm_soil_encoded_bytes(C.byref(m))
        out = (C.c_uint8 * n)()
        _codec_check(_lib().mpm_soil_encode(C.byref(m), w.ctypes.data_as(C.c_void_p), out, n))
        return bytes(out)

    @classmethod
    def from_bytes(cls, data: bytes, max_particles: int) -> "MpmSoilCheckpoint":
        """``from_bytes`` (:216-267): rejects corrupt, non-finite, unsupported,
        oversized and trailing data."""
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
        m = SoilMeta()
        L = _lib()
        _codec_check(L.mpm_soil_decode_meta(buf, len(data), int(max_particles), C.byref(m)))
        words = np.zeros(int(m.count) * WORDS_PER_PARTICLE, np.uint32)
        _codec_check(L.mpm_soil_decode(buf, len(data), int(max_particles), C.byref(m),
                                       words.ctypes.data_as(C.c_void_p), words.size))
        return cls._from_meta(m, words)

    @classmethod
    def _from_meta(cls, m: SoilMeta, words: np.ndarray) -> "MpmSoilCheckpoint":
        cfg = config_from_words(list(m.config))
        mats = [rheology_from_words(list(m.materials[i])) for i in range(m.n_materials)]
        return cls(cfg=cfg, grid_window=tuple(int(v) for v in m.grid_window_cells), material_table=mats,
                   count=int(m.count), words=words)


def blake3(data: bytes) -> bytes:
    """The checkpoint's digest primitive (``blake3::hash``)."""
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
    out = (C.c_uint8 * 32)()
    _codec_check(_lib().mpm_blake3(buf, len(data), out))
    return bytes(out)


def _rows(materials: list, cfg: MpmConfig):
    from .mpm import Material

    rows = [material_row(r) for r in materials] if materials else [material_row(cfg.rheology)]
    return (Material * len(rows))(*[Material(*[float(t) for t in r]) for r in rows]), len(rows)


def checkpoint_soil(sim) -> MpmSoilCheckpoint:
    """``checkpoint_soil_with`` (:388-431): blocking archive of the particle
    state; fails while a grid fault is latched."""
    from .mpm import MpmError

    if len(sim.cfgs) != 1 or sim.cfg.model != MODEL_SHADER:
        raise MpmError(1, "the soil checkpoint is defined for single-scene shader-model simulations")
    n = sim.n
    words = np.zeros(n * WORDS_PER_PARTICLE, np.uint32)
    sim._check(_lib().mpm_checkpoint_capture(sim._h, words.ctypes.data_as(C.c_void_p), n))
    return MpmSoilCheckpoint(cfg=_soil_config(sim.cfg), grid_window=tuple(sim.cfg.grid_window_cells),
                             material_table=list(sim.cfg.materials), count=n, words=words)


def _soil_config(cfg: MpmConfig) -> MpmConfig:
    """The ``MpmConfig`` part of this package's configuration (the fields the
    format carries; the window and the table travel beside it)."""
    return config_from_words(config_words(cfg))


def restore_soil(sim, ckpt: MpmSoilCheckpoint):
    """``restore_soil_with`` (:436-541): replace the particle set, material
    table and grid window; nothing changes when the checkpoint is rejected."""
    from .mpm import MpmError

    if len(sim.cfgs) != 1 or sim.cfg.model != MODEL_SHADER:
        raise MpmError(1, "the soil checkpoint is defined for single-scene shader-model simulations")
    w = ckpt._words()
    m = ckpt.meta()
    # validate_soil_restore_origin (:457-490): every configuration word but the
    # capacities must match this simulation's
    mine = config_words(sim.cfg)
    theirs = config_words(ckpt.cfg.copy(max_particles=sim.cfg.max_particles, max_rigid=sim.cfg.max_rigid))
    if mine != theirs:
        raise MpmError(1, "MPM soil checkpoint: incompatible configuration, capacity or attached rigid bodies")
    rows, n_rows = _rows(ckpt.material_table, ckpt.cfg)
    sim._check(_lib().mpm_checkpoint_restore(sim._h, C.byref(m), w.ctypes.data_as(C.c_void_p), rows, n_rows))
    sim.counts = [int(ckpt.count)]
    sim.cfgs[0] = sim.cfg.copy(materials=list(ckpt.material_table), grid_window_cells=tuple(ckpt.grid_window))
    sim.cfg = sim.cfgs[0]
