"""Soil checkpoint format ("MOLLSOIL", mpm_checkpoint.rs) on the CPU.

The codec is host code in libmolla_mpm.so; these tests need no GPU.

* BLAKE3 (the `blake3` crate 1.8.5 is absent): the oracle and the product
  against the published known answers in tests/golden/blake3_vectors.json,
  and against each other over tree shapes up to several levels and over the
  threaded path of the product (inputs of MiBs).
* The byte format: the product codec against the oracle restatement on the
  reference's fixture (mpm_checkpoint_tests.rs:3-60) for every rheology and
  time step, and the reference's own assertions: lossless round trip,
  allocation limit, every truncation and bit flip rejected, trailing words
  rejected, semantic validation behind a valid digest (:62-153) and the
  lattice (v2) header rules (mpm_lattice_checkpoint_tests.rs:10-67).
"""
import json
import os
import struct

import numpy as np
import pytest

from molla_amd.checkpoint import (HEADER_BYTES, CheckpointError, MpmSoilCheckpoint, blake3, config_words,
                                  rheology_words)
from molla_amd.config import Explicit, Implicit, MpmConfig, Sand, Snow, Viscous
from oracle import soil_checkpoint as ock
from oracle.blake3_ref import blake3 as blake3_oracle
from tests.fixtures import soil_fixture

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blake3_vectors.json")


def _vectors():
    g = json.load(open(GOLDEN))
    out = [(bytes(i % 251 for i in range(c["input_len"])), c["hash"]) for c in g["cases"]]
    out += [(c["input"].encode(), c["hash"]) for c in g["ascii"]]
    return out


@pytest.mark.parametrize("fn", [blake3_oracle, blake3], ids=["oracle", "product"])
def test_blake3_published_vectors(fn):
    for data, want in _vectors():
        assert fn(data).hex() == want, len(data)


def test_blake3_product_matches_oracle_over_tree_shapes():
    rng = np.random.default_rng(3)
    lens = [2, 63, 64, 65, 127, 128, 129, 1023, 3073, 4097, 5120, 5121, 6145, 7169, 8193, 16385, 33 * 1024 + 7]
    for n in lens:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert blake3(data) == blake3_oracle(data), n


def test_blake3_threaded_path_matches_oracle():
    # inputs of >= 1 MiB hash their subtrees on host threads; 2 MiB + 1 byte
    # spawns at three levels of the tree
    data = np.random.default_rng(9).integers(0, 256, (2 << 20) + 1, dtype=np.uint8).tobytes()
    assert blake3(data) == blake3_oracle(data)


def _ck(count, rheology=None, step=None, window=(0, 0, 0)):
    cfg, mats, words = soil_fixture(count)
    if rheology is not None:
        cfg = cfg.copy(rheology=rheology)
    if step is not None:
        cfg = cfg.copy(time_step=step)
    return MpmSoilCheckpoint(cfg=cfg, grid_window=window, material_table=mats, count=count, words=words)


def _odict(ck: MpmSoilCheckpoint):
    return dict(window=list(ck.grid_window), config=config_words(ck.cfg),
                materials=[rheology_words(r) for r in ck.material_table], count=ck.count,
                words=[int(w) for w in ck.words])


def test_config_words_layout():
    # config_words (mpm_checkpoint.rs:594-619) of MpmConfig::sand() (mpm.rs:540-560)
    b = lambda v: int(np.float32(v).view(np.uint32))  # noqa: E731
    w = config_words(MpmConfig.sand())
    assert w == [32, 16384, 8, b(0.05), b(-0.8), b(0.0), b(-0.8), b(-9.81), b(5e3), b(0.3),
                 b(np.float32(30.0) * (np.float32(np.pi) / np.float32(180.0))), b(0.0), b(0.5), b(1e-3),
                 b(1e-6), 0, w[16], 0, 0, 0, 0]
    assert w[16] == w[10]
    assert config_words(MpmConfig.sand().copy(time_step=Implicit(3)))[19:] == [1, 3]
    # the reference's corruption test writes 99 at words 15 and 19: the
    # rheology kind and the time-step kind
    assert rheology_words(Snow(0.025, 0.0075, 10.0))[0] == 1
    assert rheology_words(Viscous(3.0, 5.0))[3] == 0


@pytest.mark.parametrize("count", [0, 1, 129])
def test_format_is_lossless_bounded_and_matches_oracle(count):
    for rheology in (Sand(0.5), Snow(0.025, 0.0075, 10.0), Viscous(3.0, 5.0)):
        for step in (Explicit(), Implicit(3)):
            ck = _ck(count, rheology, step)
            data = ck.to_bytes()
            assert data == ock.encode(_odict(ck))
            assert data[:8] == b"MOLLSOIL" and data[8:12] == struct.pack("<I", 1)
            assert len(data) == HEADER_BYTES + 4 * (21 + 1 + 12 + 1) + 160 * count
            back = MpmSoilCheckpoint.from_bytes(data, count)
            assert back.to_bytes() == data
            assert config_words(back.config()) == config_words(ck.cfg)
            assert [rheology_words(r) for r in back.materials()] == [rheology_words(r) for r in ck.materials()]
            assert back.particle_bytes() == count * 160
            assert np.array_equal(back.words, ck.words)
            ock.decode(data, count)
            if count > 0:
                with pytest.raises(CheckpointError):
                    MpmSoilCheckpoint.from_bytes(data, count - 1)
                with pytest.raises(ock.Invalid):
                    ock.decode(data, count - 1)


def test_every_truncation_and_bit_flip_is_rejected():
    data = _ck(1).to_bytes()
    for n in range(len(data)):
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(data[:n], 129)
    for i in range(len(data)):
        bad = bytearray(data)
        bad[i] ^= 1
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(bytes(bad), 129)
    # trailing word with a recomputed digest: payload length mismatch
    trailing = bytearray(data + b"\0" * 4)
    trailing[12:HEADER_BYTES] = blake3_oracle(bytes(trailing[HEADER_BYTES:]))
    with pytest.raises(CheckpointError):
        MpmSoilCheckpoint.from_bytes(bytes(trailing), 129)
    with pytest.raises(ock.Invalid):
        ock.decode(bytes(trailing), 129)


def _mutations():
    nan, inf = float("nan"), float("inf")
    fb = lambda v: int(np.float32(v).view(np.uint32))  # noqa: E731

    def cfg(**kw):
        return lambda d: d["config"].__setitem__(slice(None), config_words(d["cfg"].copy(**kw)))

    def word(i, v):
        return lambda d: d["words"].__setitem__(i, v)

    return [
        cfg(cell_size=nan), cfg(max_particles=0), cfg(grid_size=0xFFFFFFFF), cfg(poisson=0.5),
        cfg(rheology=Sand(friction_angle=inf)),
        lambda d: d["materials"].__setitem__(1, rheology_words(Snow(1.0, 0.0, 1.0))),
        word(0, fb(nan)), word(4, fb(inf)), word(8, fb(nan)), word(20, fb(inf)), word(32, 0),
        word(33, fb(-1.0)), word(34, fb(nan)), word(35, 3), word(36, fb(nan)),
        lambda d: d.__setitem__("count", 0xFFFFFFFF),
        lambda d: d["words"].pop(),
        lambda d: d.__setitem__("materials", [rheology_words(d["cfg"].rheology)] * 9),
    ]


def test_valid_digest_does_not_bypass_semantic_validation():
    # mpm_checkpoint_tests.rs:103-153, the mutations applied to the raw words
    # and encoded by the oracle encoder (which, like to_bytes, does not validate)
    base = _ck(1)
    MpmSoilCheckpoint.from_bytes(ock.encode(_odict(base)), 129)  # the unmutated base is accepted
    for mutate in _mutations():
        d = _odict(base)
        d["cfg"] = base.cfg
        mutate(d)
        data = ock.encode(d)
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(data, 129)
        with pytest.raises(ock.Invalid):
            ock.decode(data, 129)
    for index in (15, 19):
        data = bytearray(base.to_bytes())
        data[HEADER_BYTES + 4 * index:HEADER_BYTES + 4 * index + 4] = struct.pack("<I", 99)
        data[12:HEADER_BYTES] = blake3(bytes(data[HEADER_BYTES:]))
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(bytes(data), 129)


@pytest.mark.parametrize("count", [0, 1, 129])
def test_lattice_checkpoint_signed_offsets_roundtrip(count):
    old = _ck(count).to_bytes()
    assert old[8:12] == struct.pack("<I", 1)
    for cells in ((1, 0, 0), (-1, 16, -32), (-2 ** 31, 2 ** 31 - 1, 7)):
        ck = _ck(count, window=cells)
        data = ck.to_bytes()
        assert data == ock.encode(_odict(ck))
        assert data[8:12] == struct.pack("<I", 2) and len(data) == len(old) + 12
        back = MpmSoilCheckpoint.from_bytes(data, count)
        assert back.grid_window_cells() == cells
        assert config_words(back.config()) == config_words(ck.cfg)
        assert np.array_equal(back.words, ck.words) and back.to_bytes() == data


def test_lattice_checkpoint_integrity_version_and_bounds():
    data = _ck(2, window=(-1, 2, 3)).to_bytes()
    with pytest.raises(CheckpointError):
        MpmSoilCheckpoint.from_bytes(data, 1)
    for n in range(len(data)):
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(data[:n], 2)
    for i in range(len(data)):
        bad = bytearray(data)
        bad[i] ^= 1
        with pytest.raises(CheckpointError):
            MpmSoilCheckpoint.from_bytes(bytes(bad), 2)
    down = bytearray(data)
    down[8:12] = struct.pack("<I", 1)
    with pytest.raises(CheckpointError):
        MpmSoilCheckpoint.from_bytes(bytes(down), 2)
    zero = bytearray(data)
    zero[44:56] = b"\0" * 12
    zero[12:44] = ock.digest(2, bytes(zero[44:]))
    with pytest.raises(CheckpointError):
        MpmSoilCheckpoint.from_bytes(bytes(zero), 2)
    trailing = bytearray(data + b"\0" * 4)
    trailing[12:44] = ock.digest(2, bytes(trailing[44:]))
    with pytest.raises(CheckpointError):
        MpmSoilCheckpoint.from_bytes(bytes(trailing), 2)


def test_large_checkpoint_roundtrip():
    # 40,000 particles (6.4 MB payload, a threaded digest)
    cfg, mats, _ = soil_fixture(0)
    n = 40_000
    rng = np.random.default_rng(4)
    streams = []
    for s, stride in enumerate((4, 4, 12, 12, 1, 1, 1, 1, 4)):
        if s == 7:
            streams.append(rng.integers(0, 3, n, dtype=np.uint32))
        elif s in (4, 5):
            streams.append(rng.uniform(0.5, 2.0, n).astype(np.float32).view(np.uint32))
        else:
            streams.append(rng.normal(size=n * stride).astype(np.float32).view(np.uint32))
    ck = MpmSoilCheckpoint(cfg=cfg.copy(max_particles=n), grid_window=(0, 0, 0), material_table=mats, count=n,
                           words=np.concatenate(streams))
    data = ck.to_bytes()
    assert data[12:HEADER_BYTES] == blake3(data[HEADER_BYTES:])
    back = MpmSoilCheckpoint.from_bytes(data, n)
    assert np.array_equal(back.words, ck.words)
