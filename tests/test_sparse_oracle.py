"""The oracle's sparse-pool tile marks (mpm_active_tiles.wgsl:79-124) against
closed forms and the reference's own counts (tests/mpm_sparse_grid_gpu.rs):
one particle marks 8 tiles (cells base-1..base+3 always span two tiles per
axis), the clamp at the grid's faces, particles off the grid mark nothing,
and a sphere's shell reaches radius + 2 cells."""
import numpy as np

import oracle as orc
from molla_amd import MpmConfig


def _cfg(size=64):
    # mpm_sparse_grid_gpu.rs:6-23
    return MpmConfig.sand().copy(grid_size=size, grid_origin=(-1.0, -1.0, -1.0), cell_size=0.0625)


def _lattice(n, offset):
    # mpm_sparse_grid_gpu.rs:25-35
    return np.array([[offset[0] + (i % 7) * 0.021 + 0.00013, offset[1] + (i // 7 % 7) * 0.021 + 0.00031,
                      offset[2] + (i // 49) * 0.021 + 0.00047] for i in range(n)], np.float32)


def test_one_particle_marks_the_two_tiles_around_its_stencil_on_each_axis():
    cfg = _cfg()
    W = 16
    # x = 0.1: xp = 17.6, base 17, cells 16..20 -> tiles 4 and 5 on every axis
    got = orc.marked_tiles(cfg, [[0.1, 0.1, 0.1]])
    want = sorted(tx + W * (ty + W * tz) for tx in (4, 5) for ty in (4, 5) for tz in (4, 5))
    assert got.tolist() == want
    # x = 0: xp = 16, base 15, cells 14..18 -> tiles 3 and 4
    got = orc.marked_tiles(cfg, [[0.0, 0.0, 0.0]])
    assert got.tolist() == sorted(tx + W * (ty + W * tz) for tx in (3, 4) for ty in (3, 4) for tz in (3, 4))


def test_clamp_at_the_faces_and_particles_off_the_grid():
    cfg = _cfg()
    # xp = 0.2: base -1, cells clamp to 0..2: tile 0 only
    assert orc.marked_tiles(cfg, [[-0.9875, -0.9875, -0.9875]]).tolist() == [0]
    # far outside: the clamped range is empty
    assert len(orc.marked_tiles(cfg, [[50.0, 50.0, 50.0]])) == 0
    # the top face: xp = 63.4, base 62, cells 61..63 -> tile 15 only
    x = -1.0 + 63.4 * 0.0625
    assert orc.marked_tiles(cfg, [[x, x, x]]).tolist() == [15 + 16 * (15 + 16 * 15)]


def test_reference_counts():
    cfg = _cfg()
    # overflow test (mpm_sparse_grid_gpu.rs:204-238): two distant particles need 16
    # tiles, one of them 8
    assert len(orc.marked_tiles(cfg, [[0.1] * 3, [1.1] * 3])) == 16
    assert len(orc.marked_tiles(cfg, [[0.1] * 3])) == 8
    # large domains (:240-296): the 129-particle lattice at 0.1 marks 8 tiles
    for size in (32, 96, 512, 1536):
        assert len(orc.marked_tiles(_cfg(size), _lattice(129, [0.1, 0.1, 0.1]))) == 8


def test_sphere_shell_reach():
    cfg = _cfg()
    # radius 0 at the centre of cell 32 (xp = 32): cells 30..34 -> tiles 7 and 8
    got = orc.marked_tiles(cfg, [], spheres=[((1.0, 1.0, 1.0), 0.0)])
    assert got.tolist() == sorted(tx + 16 * (ty + 16 * tz) for tx in (7, 8) for ty in (7, 8) for tz in (7, 8))
    # a sphere covering the domain marks every tile
    assert len(orc.marked_tiles(cfg, [], spheres=[((0.0, 0.0, 0.0), 1000.0)])) == 16 ** 3
