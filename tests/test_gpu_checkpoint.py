"""Soil checkpoint capture and restore on the GPU (mpm_checkpoint.rs:388-541),
after the reference's GPU tests (mpm_checkpoint_tests.rs:172-379), which need
a wgpu adapter and run here on the HIP path:

* a restore followed by a capture gives back the same words, bit for bit,
  including the words the kernels never read (w components, column pads,
  tire compaction);
* archive, drop, decode and reload: the history survives, five restores of
  the saved state are exact, stepping continues;
* a restored state steps like the FP64 oracle started from the same words;
* a continued trajectory (restore then step) equals the uninterrupted one bit
  for bit, for the reference's one-particle case in every P2G mode and for a
  4,096-particle block in the deterministic mode;
* rejected restores change nothing; an empty restore empties the handle;
* a latched fault cannot be archived, nor hidden by a status reset.
"""
import numpy as np
import pytest

import oracle as orc
from molla_amd import Implicit, MpmError, MpmSimulation
from molla_amd.checkpoint import MpmSoilCheckpoint, config_words
from tests.fixtures import fill_box, soil_fixture
from tests.gpu_util import log, rel_err

pytestmark = pytest.mark.gpu

MODES = ("fused", "global_atomic", "deterministic")


def _state(count, **cfg_kw):
    cfg, mats, words = soil_fixture(count)
    cfg = cfg.copy(**cfg_kw)
    return MpmSoilCheckpoint(cfg=cfg, grid_window=(0, 0, 0), material_table=mats, count=count, words=words)


def _sim(cfg, mode="fused"):
    sim = MpmSimulation(cfg)
    sim.set_p2g_mode(mode)
    return sim


def test_restore_then_capture_is_exact():
    for count in (0, 1, 129):
        log(f"restore/capture count={count}")
        st = _state(count)
        sim = _sim(st.cfg)
        sim.restore_soil(st)
        got = sim.checkpoint_soil()
        assert got.count == count
        assert np.array_equal(got.words, st.words)
        assert got.to_bytes() == st.to_bytes()
        if count:
            # position w is -0.0 in the fixture: preserved bit for bit
            assert (got.stream(0)[3::4] == 0x80000000).all()
            assert np.array_equal(sim.read_field("material"), st.stream(7))
        sim.close()


@pytest.mark.parametrize("mode", MODES)
def test_archive_drop_reload_preserves_history(mode):
    original = _state(129)
    src = _sim(original.cfg, mode)
    src.restore_soil(original)
    src.step_n(1e-4, 3)
    evolved = src.checkpoint_soil()
    for s in (0, 3):
        assert not np.array_equal(evolved.stream(s), original.stream(s))
    for s in (4, 5, 7, 8):
        assert np.array_equal(evolved.stream(s), original.stream(s)), s
    assert (evolved.stream(0)[3::4] == 0).all()  # a step stores w = 0
    data = evolved.to_bytes()
    src.close()
    loaded = MpmSoilCheckpoint.from_bytes(data, 129)
    dst = _sim(loaded.config(), mode)
    dst.restore_soil(loaded)
    for _ in range(5):
        saved = dst.checkpoint_soil()
        assert saved.to_bytes() == data
        dst.restore_soil(saved)
    dst.step_n(1e-4, 2)
    cont = dst.checkpoint_soil()
    assert not np.array_equal(cont.stream(0), saved.stream(0))
    st = dst.read_status()
    assert st["completed_gathers"] == 2 and not st["faulted"]
    dst.close()


def _oracle_from(ck):
    n = ck.count
    f = lambda s: ck.stream(s).view(np.float32)  # noqa: E731
    col = lambda s: f(s).reshape(n, 3, 4)[:, :, :3].transpose(0, 2, 1)  # noqa: E731  columns -> M[i][j]
    return orc.State(f(0).reshape(n, 4)[:, :3], f(1).reshape(n, 4)[:, :3], C_=col(2), F=col(3), jp=f(6),
                     mass=f(5), vol=f(4), mat=ck.stream(7), dtype=np.float64)


def test_restored_state_steps_like_the_oracle():
    st = _state(129)
    sim = _sim(st.cfg)
    sim.restore_soil(st)
    sim.step_n(1e-4, 3)
    got = sim.checkpoint_soil()
    ref = _oracle_from(st)
    ocfg = st.cfg.copy(materials=st.material_table)
    orc.step(ocfg, ref, 1e-4, 3)
    g = _oracle_from(got)
    errs = {k: rel_err(getattr(g, k), getattr(ref, k)) for k in ("x", "v", "C", "F", "jp")}
    log("restored 129-particle sand/snow/viscous state vs FP64 oracle: " +
        " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert errs["x"] < 1e-6 and errs["v"] < 1e-4 and errs["C"] < 1e-3 and errs["F"] < 1e-5
    assert errs["jp"] < 1e-5
    sim.close()


@pytest.mark.parametrize("implicit", [False, True])
def test_continued_trajectory_and_material_edits_survive_restore(implicit):
    kw = {"time_step": Implicit(2)} if implicit else {}
    state = _state(1, **kw)
    for mode in MODES:
        log(f"continued trajectory implicit={implicit} mode={mode}")
        ref = _sim(state.cfg, mode)
        ref.restore_soil(state)
        ref.step_n(1e-4, 3)
        ref.set_materials(state.material_table)
        ref.set_particle_materials([1])
        saved = ref.checkpoint_soil()
        assert saved.stream(7).tolist() == [1]
        res = _sim(state.cfg, mode)
        res.restore_soil(saved)
        ref.step_n(1e-4, 8)
        res.step_n(1e-4, 8)
        a, b = ref.checkpoint_soil(), res.checkpoint_soil()
        assert np.array_equal(a.words, b.words), mode
        ref.close()
        res.close()


def test_deterministic_continuation_is_bitwise_for_a_block():
    cfg, _, _ = soil_fixture(0)
    x = fill_box([-0.15, 0.2, -0.15], [0.15, 0.5, 0.15], 16, 5)  # 4,096 particles
    cfg = cfg.copy(max_particles=len(x))
    ref = _sim(cfg, "deterministic")
    ref.set_particles(x)
    ref.step_n(1e-4, 5)
    saved = ref.checkpoint_soil()
    res = _sim(cfg, "deterministic")
    res.restore_soil(MpmSoilCheckpoint.from_bytes(saved.to_bytes(), len(x)))
    ref.step_n(1e-4, 10)
    res.step_n(1e-4, 10)
    a, b = ref.checkpoint_soil(), res.checkpoint_soil()
    log(f"4096-particle continuation: {int((a.words != b.words).sum())} differing words")
    assert np.array_equal(a.words, b.words)
    ref.close()
    res.close()


def test_rejected_restore_does_not_mutate_and_empty_restore_empties():
    state = _state(1)
    small = _sim(state.cfg.copy(max_particles=1))
    small.restore_soil(state)
    with pytest.raises(MpmError):
        small.restore_soil(_state(129))
    assert np.array_equal(small.checkpoint_soil().words, state.words)
    small.close()
    sim = _sim(state.cfg)
    sim.restore_soil(state)
    incompatible = _state(1)
    incompatible.cfg = incompatible.cfg.copy(grid_origin=(-0.7, 0.0, -0.8))
    corrupt = _state(1)
    corrupt.words = corrupt.words.copy()
    corrupt.words[0] = np.float32(np.nan).view(np.uint32)
    oversized = _state(129)
    oversized.cfg = oversized.cfg.copy(max_particles=130)
    oversized.count = 130
    oversized.words = np.zeros(130 * 40, np.uint32)
    for bad in (incompatible, corrupt, oversized):
        with pytest.raises(Exception):
            sim.restore_soil(bad)
        assert sim.checkpoint_soil().to_bytes() == state.to_bytes()
    empty = _state(0)
    sim.restore_soil(empty)
    assert sim.n == 0
    sim.step_n(1e-4, 1)
    assert sim.checkpoint_soil().to_bytes() == empty.to_bytes()
    sim.restore_soil(state)
    assert np.array_equal(sim.checkpoint_soil().words, state.words)
    sim.close()


def test_fault_cannot_be_archived_or_hidden_by_status_reset():
    cfg = _state(1).cfg.copy(gravity_y=0.0)
    sim = _sim(cfg)
    sim.set_particles([[-0.69, 0.11, -0.69]])
    good = sim.checkpoint_soil()
    x = np.array([[-0.69, 0.11, -0.69], [0.0, 0.5, 0.0]], np.float32)
    v = np.array([[0, 0, 0], [np.nan, 0, 0]], np.float32)
    sim.set_particles(x, v)
    sim.step_n(1e-4, 1)
    assert sim.read_status()["faulted"]
    with pytest.raises(MpmError):
        sim.checkpoint_soil()
    sim.reset_status()
    with pytest.raises(MpmError):
        sim.checkpoint_soil()
    sim.restore_soil(good)
    assert sim.checkpoint_soil().to_bytes() == good.to_bytes()
    sim.step_n(1e-4, 1)
    assert not sim.read_status()["faulted"]
    sim.checkpoint_soil()
    sim.close()


def test_autotune_keeps_the_archive_rules():
    # The autotuner's trial substeps are not the caller's steps: a reset of a
    # latched fault still blocks an archive afterwards, and a restored handle
    # still captures its own position/velocity w words.
    cfg = _state(1).cfg.copy(gravity_y=0.0)
    sim = _sim(cfg)
    x = np.array([[-0.69, 0.11, -0.69], [0.0, 0.5, 0.0]], np.float32)
    v = np.array([[0, 0, 0], [np.nan, 0, 0]], np.float32)
    sim.set_particles(x, v)
    sim.step_n(1e-4, 1)
    assert sim.read_status()["faulted"]
    sim.reset_status()
    sim.autotune(1e-4, 2)
    assert not sim.read_status()["faulted"]
    with pytest.raises(MpmError):
        sim.checkpoint_soil()
    sim.close()
    st = _state(129)
    sim = _sim(st.cfg)
    sim.restore_soil(st)
    sim.autotune(1e-4, 2)
    assert sim.checkpoint_soil().to_bytes() == st.to_bytes()
    sim.close()


def test_large_state_restore_capture_is_idempotent():
    # 262,144 particles of mixed material: capture(restore(capture)) is exact
    # and a restored copy continues exactly like the source in deterministic mode
    cfg, mats, _ = soil_fixture(0)
    x = fill_box([-0.3, 0.1, -0.3], [0.3, 0.7, 0.3], 64, 8)
    n = len(x)
    cfg = cfg.copy(max_particles=n)
    a = _sim(cfg, "deterministic")
    a.set_particles(x)
    a.set_materials(mats)
    a.set_particle_materials(np.arange(n, dtype=np.uint32) % 3)
    a.step_n(2e-4, 4)
    s1 = a.checkpoint_soil()
    b = _sim(cfg, "deterministic")
    b.restore_soil(s1)
    assert b.checkpoint_soil().to_bytes() == s1.to_bytes()
    a.step_n(2e-4, 4)
    b.step_n(2e-4, 4)
    wa, wb = a.checkpoint_soil().words, b.checkpoint_soil().words
    log(f"{n} particles: continuation differs in {int((wa != wb).sum())} words")
    assert np.array_equal(wa, wb)
    assert config_words(s1.cfg) == config_words(cfg)
    a.close()
    b.close()


def test_batch_and_cpudiff_handles_are_rejected():
    from molla_amd import MpmConfig

    sim = MpmSimulation([MpmConfig.sand(), MpmConfig.sand()])
    with pytest.raises(MpmError):
        sim.checkpoint_soil()
    sim.close()
    sim = MpmSimulation(MpmConfig.cpu_diff_default(), max_particles=8)
    with pytest.raises(MpmError):
        sim.checkpoint_soil()
    sim.close()
