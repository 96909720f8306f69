"""The key query without allocation (mpm_sort_keys_noalloc), through the C
ABI, against the allocating key query (mpm_sort_keys_debug) and the oracle:
bit-exact keys and ids on two calls in a row (the second on a smaller scene,
so stale contents of the shared scratch would show), on a batch scene after
stepping (sorted layout), and a scene above the limit refused."""
import numpy as np
import pytest

import oracle as orc
from molla_amd import MpmConfig, MpmError, MpmSimulation, workloads
from molla_amd.devmem import DeviceArray
from molla_amd.mpm import lib
from tests.gpu_util import log

pytestmark = pytest.mark.gpu


def query(sim, capacity, scene=0):
    # a non-zero fill, so an unwritten word cannot pass as a key
    fill = np.full(capacity, 0x5A5A5A5A, np.uint32)
    with DeviceArray.from_host(fill) as k, DeviceArray.from_host(fill) as i:
        sim.sort_keys_noalloc(k.address, i.address, capacity, scene)
        sim.synchronize()
        return k.download(), i.download()


def test_noalloc_key_query_matches_the_allocating_one():
    cfg = MpmConfig.sand().copy(grid_size=256, cell_size=1 / 256, grid_origin=(0, 0, 0))
    rng = np.random.default_rng(11)
    # first call on 200,000 particles, second on 3,000 (padded to 4,000)
    for n, capacity in ((200_000, 200_000), (3_000, 4_000)):
        x = rng.random((n, 3)).astype(np.float32)
        x[: n // 10] = x[n // 10: 2 * (n // 10)]  # ties
        log(f"noalloc key query: {n} particles, capacity {capacity}")
        sim = MpmSimulation(cfg, max_particles=n)
        sim.set_particles(x)
        got_k, got_i = query(sim, capacity)
        ref_k, ref_i = sim.sort_keys_debug(capacity=capacity)
        ek, ei = orc.sorted_tile_keys(cfg, x, capacity=capacity)
        sim.close()
        assert np.array_equal(got_k, ref_k) and np.array_equal(got_i, ref_i)
        assert np.array_equal(got_k, ek) and np.array_equal(got_i, ei)


def test_noalloc_key_query_on_a_stepped_batch_leaves_stepping_unchanged():
    cfgs, xs, meta = workloads.batched_scenes(6, seed=21, n_lo=1000, n_hi=5000)
    dt = meta["dt"]
    sims = [MpmSimulation(cfgs) for _ in range(2)]
    for s in sims:
        s.set_particles_batch(xs)
        s.set_sort_interval(3)
        s.step_n(dt, 5)  # sorted layout, slots no longer in id order
    for scene in (4, 1):  # a larger scene, then a smaller one
        cnt = len(xs[scene])
        log(f"noalloc key query: batch scene {scene} of {cnt} particles")
        got_k, got_i = query(sims[0], cnt + 17, scene)
        ref_k, ref_i = sims[1].sort_keys_debug(scene, capacity=cnt + 17)
        assert np.array_equal(got_k, ref_k) and np.array_equal(got_i, ref_i)
    # the query leaves the handle's sort state alone: stepping on is unchanged
    # against a handle that never ran it
    ref = MpmSimulation(cfgs)
    ref.set_particles_batch(xs)
    ref.set_sort_interval(3)
    ref.step_n(dt, 5)
    sims[0].step_n(dt, 4)
    ref.step_n(dt, 4)
    assert np.array_equal(sims[0].read_positions(), ref.read_positions())
    for s in (*sims, ref):
        s.close()


def test_noalloc_key_query_refuses_beyond_its_limit():
    cfg = MpmConfig.sand().copy(grid_size=64, cell_size=1 / 64, grid_origin=(0, 0, 0))
    n = 2_097_152 + 1
    x = np.random.default_rng(3).random((n, 3)).astype(np.float32)
    sim = MpmSimulation(cfg, max_particles=n)
    sim.set_particles(x)
    with DeviceArray(n, np.uint32) as k:
        with pytest.raises(MpmError, match="no-allocation limit"):
            sim.sort_keys_noalloc(k.address, k.address, n)
    # invalid arguments are refused before any device work
    assert lib().mpm_sort_keys_noalloc(sim._h, 0, n, None, None) == 1
    sim.close()
