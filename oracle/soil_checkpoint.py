"""MOLLSOIL checkpoint codec in pure Python -- TEST INFRASTRUCTURE ONLY.

Restates ``MpmSoilCheckpoint`` of crates/molla-solvers/src/mpm_checkpoint.rs
on raw words so the product codec (libmolla_mpm.so, ``mpm_soil_*``) can be
checked byte for byte:

  constants        :78-85   (MAGIC, VERSION 1, LATTICE_VERSION 2, HEADER_BYTES
                             44, STRIDES, WORDS_PER_PARTICLE 40, size bound)
  digest           :87-94   (BLAKE3 of [version bytes, v2 only] || payload)
  encode           :185-214 (to_bytes)
  decode           :216-267 (from_bytes, Words reader :621-700)
  validate         :281-330 (with validate_rheology :559-592)

A checkpoint here is a dict: window (3 ints), config (21 u32), materials
(list of 4-word rheology encodings), count, words (list or array of u32).
Pinned by the reference's own assertions (mpm_checkpoint_tests.rs:62-153,
mpm_lattice_checkpoint_tests.rs:10-67) and by the BLAKE3 known answers; the
reference ships no encoded checkpoint file to compare bytes against.
"""
from __future__ import annotations

import math
import struct

import numpy as np

from .blake3_ref import blake3

MAGIC = b"MOLLSOIL"
VERSION = 1
LATTICE_VERSION = 2
HEADER_BYTES = 44
STRIDES = (4, 4, 12, 12, 1, 1, 1, 1, 4)
WORDS_PER_PARTICLE = 40
MAX_MATERIALS = 8
FRAC_PI_2_F32 = float(np.float32(math.pi / 2))


class Invalid(ValueError):
    pass


def max_encoded_bytes(max_particles: int) -> int:
    return HEADER_BYTES + (23 + 4 * MAX_MATERIALS + max_particles * WORDS_PER_PARTICLE) * 4


def digest(version: int, payload: bytes) -> bytes:
    prefix = struct.pack("<I", version) if version == LATTICE_VERSION else b""
    return blake3(prefix + payload)


def f32(w: int) -> float:
    return float(np.uint32(w).view(np.float32))


def encode(ck: dict) -> bytes:
    window = [int(v) for v in ck["window"]]
    version = VERSION if window == [0, 0, 0] else LATTICE_VERSION
    words = []
    if version == LATTICE_VERSION:
        words += [v & 0xFFFFFFFF for v in window]
    words += [int(v) & 0xFFFFFFFF for v in ck["config"]]
    words.append(len(ck["materials"]))
    for m in ck["materials"]:
        words += [int(v) for v in m]
    words.append(int(ck["count"]) & 0xFFFFFFFF)
    words += [int(v) for v in ck["words"]]
    payload = struct.pack(f"<{len(words)}I", *words)
    header = MAGIC + struct.pack("<I", version) + digest(version, payload)
    return header + payload


def _rheology(w, where):
    kind, a, b, c = w
    if kind == 0 and b == 0 and c == 0:
        return
    if kind == 1:
        return
    if kind == 2 and c == 0:
        return
    raise Invalid(f"unsupported or noncanonical rheology encoding ({where})")


def decode(data: bytes, max_particles: int) -> dict:
    if len(data) < HEADER_BYTES:
        raise Invalid("truncated checkpoint header")
    version = struct.unpack_from("<I", data, 8)[0]
    if (version not in (VERSION, LATTICE_VERSION)
            or len(data) > max_encoded_bytes(max_particles) + (12 if version == LATTICE_VERSION else 0)
            or data[:8] != MAGIC or (len(data) - HEADER_BYTES) % 4 != 0
            or data[12:HEADER_BYTES] != digest(version, data[HEADER_BYTES:])):
        raise Invalid("invalid format, digest or allocation limit")
    words = list(struct.unpack(f"<{(len(data) - HEADER_BYTES) // 4}I", data[HEADER_BYTES:]))
    pos = 0

    def take(k=1):
        nonlocal pos
        if pos + k > len(words):
            raise Invalid("truncated metadata")
        out = words[pos:pos + k]
        pos += k
        return out

    window = [0, 0, 0]
    if version == LATTICE_VERSION:
        window = [int(np.uint32(v).view(np.int32)) for v in take(3)]
        if window == [0, 0, 0]:
            raise Invalid("zero-offset checkpoints use version 1")
    config = take(19)
    _rheology(config[15:19], "config")
    config += take(2)
    if not ((config[19] == 0 and config[20] == 0) or config[19] == 1):
        raise Invalid("unsupported time step encoding")
    rows = take()[0]
    if rows > MAX_MATERIALS:
        raise Invalid("material table exceeds capacity")
    materials = []
    for i in range(rows):
        m = take(4)
        _rheology(m, f"row {i}")
        materials.append(m)
    count = take()[0]
    rest = words[pos:]
    if count > max_particles or len(rest) != count * WORDS_PER_PARTICLE:
        raise Invalid("particle count or payload length mismatch")
    ck = dict(window=window, config=config, materials=materials, count=count, words=rest)
    validate(ck)
    return ck


def _rheology_valid(w) -> bool:
    kind, a, b, c = w[0], f32(w[1]), f32(w[2]), f32(w[3])
    fin = math.isfinite
    if kind == 0:
        return fin(a) and 0.0 <= a < FRAC_PI_2_F32
    if kind == 1:
        return fin(a) and 0.0 <= a < 1.0 and fin(b) and b >= 0.0 and fin(c) and c >= 0.0
    if kind == 2:
        return fin(a) and a >= 0.0 and fin(b) and b >= 0.0
    return False


def validate(ck: dict):
    c = [int(v) for v in ck["config"]]
    g = c[0]
    h, ox, oy, oz, grav, E, nu, phi, coh, wall, pm, pv = (f32(v) for v in c[3:15])
    fin = math.isfinite
    count = int(ck["count"])
    if (g == 0 or g ** 3 > 0xFFFFFFFF or c[1] == 0 or c[2] == 0 or count > c[1]
            or not fin(h) or h <= 0 or not all(fin(o) for o in (ox, oy, oz)) or not fin(grav)
            or not fin(E) or E < 0 or not fin(nu) or nu <= -1 or nu >= 0.5 or not fin(phi)
            or not fin(coh) or coh < 0 or not fin(wall) or not (0.0 <= wall <= 1.0)
            or not fin(pm) or pm <= 0 or not fin(pv) or pv <= 0
            or len(ck["materials"]) > MAX_MATERIALS
            or len(ck["words"]) != count * WORDS_PER_PARTICLE):
        raise Invalid("invalid configuration or state length")
    if not _rheology_valid(c[15:19]) or not all(_rheology_valid(m) for m in ck["materials"]):
        raise Invalid("invalid material parameters")
    w = np.asarray(ck["words"], dtype=np.uint32)
    rows = max(len(ck["materials"]), 1)
    off = 0
    for s, stride in enumerate(STRIDES):
        seg = w[off:off + stride * count]
        off += stride * count
        if s == 7:
            if (seg >= rows).any():
                raise Invalid("material ID outside authoritative table")
            continue
        v = seg.view(np.float32)
        if not np.isfinite(v).all() or (s in (4, 5) and (v <= 0).any()):
            raise Invalid("non-finite history or nonpositive mass/reference volume")
