"""The rigid-body coupling of the oracle (rigid_imprint, the rigid mask of the
grid update and of the implicit sweeps, and the legacy shell force) pinned by
the reference's own assertions and closed forms, on the CPU.

* mpm_prescribed_solids_tests.rs:82-245: one cell of mass 0.125 and a known
  momentum under six solids (sphere, two boxes, three cylinders, one shifted
  and one rotated), owned by the first or the second of two bodies: the
  stamped momentum, the unchanged mass, and the wrench on the owner with the
  gravity term, zero on the other body.
* rigid_cell_imprint_wrench (mpm_oracle.rs:54-80) for a legacy sphere: the
  heavy stamp, force -(m v - p) / dt and torque arm x force.
* The shell taper (mpm.wgsl:1196-1212) at half a cell outside a sphere.
* A stamped cell skips gravity and the implicit sweeps.
* Rotated-box containment against an independent rotation-matrix test.
"""
import math

import numpy as np
import pytest

import oracle as orc
from molla_amd import Implicit, MpmConfig
from molla_amd.rigid import MpmPrescribedSolid, SolidBox, SolidCylinder, SolidSphere, body_rows

SQ = 1.0 / math.sqrt(2.0)


def config():
    # mpm_prescribed_solids_tests.rs:14-25
    return MpmConfig.sand().copy(grid_size=16, grid_origin=(-1.0, -1.0, -1.0), cell_size=0.125,
                                 max_particles=129, max_rigid=2, youngs=1000.0, gravity_y=-2.0)


def solid(shape, **kw):
    return MpmPrescribedSolid(shape=shape, linear_velocity=(0.125, -0.25, 0.0625),
                              angular_velocity=(0.5, -0.25, 1.0)).replace(**kw)


CASES = [
    (solid(SolidSphere(0.3)), True),
    (solid(SolidBox((0.2, 0.4, 0.2))), True),
    (solid(SolidBox((0.1, 0.4, 0.2))), False),
    (solid(SolidCylinder(0.3, 0.1)), True),
    (solid(SolidCylinder(0.3, 0.1), center=(0.0, -0.125, 0.0)), False),
    (solid(SolidCylinder(0.3, 0.1), orientation=(0.0, 0.0, SQ, SQ)), False),
]


def cross(a, b):
    return np.cross(np.asarray(a, np.float64), np.asarray(b, np.float64))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_prescribed_solid_single_cell_impulse_and_wrench(case):
    cfg = config()
    body, inside = CASES[case]
    owner = case % 2
    first = body if owner == 0 else body.replace(center=(10.0, 10.0, 10.0))
    G = 16
    world = 9 + 8 * 16 + 9 * 256
    p = np.array([-1 + 9 * 0.125, -1 + 8 * 0.125, -1 + 9 * 0.125])
    dt, mass = 1.0 / 128.0, 0.125
    mom = np.array([-0.03125, 0.0625, 0.015625])
    m = np.zeros(G ** 3)
    pg = np.zeros((G ** 3, 3))
    m[world] = mass
    pg[world] = mom
    m2, p2, v, mask, wrench = orc.rigid_grid(cfg, dt, body_rows(solids=[first, body]), m, pg)
    assert np.array_equal(m2, m), "a prescribed solid keeps the sand mass"
    assert np.all(wrench[1 - owner] == 0.0), "overlap must have one owner"
    arm = p - np.asarray(body.center)
    spin = cross(body.angular_velocity, arm)
    expect = mass * (np.asarray(body.linear_velocity) + spin) if inside else mom
    force = -(expect - mom) / dt + (np.array([0.0, mass * cfg.gravity_y, 0.0]) if inside else 0.0)
    torque = cross(arm, force)
    assert np.abs(p2[world] - expect).max() <= 2e-5
    assert np.abs(wrench[owner][:3] - force).max() <= 2e-5
    assert np.abs(wrench[owner][3:] - torque).max() <= 2e-5
    assert bool(mask[world]) == inside
    if inside:
        # a stamped cell keeps its velocity: no gravity, no walls
        assert np.abs(v[world] - expect / mass).max() <= 1e-12


def test_legacy_sphere_heavy_stamp_and_reference_wrench_formula():
    # rigid_cell_imprint_wrench (mpm_oracle.rs:54-80): impulse to the sand
    # m v - p, force on the sphere -(impulse)/dt, torque arm x force.
    cfg = config()
    G = 16
    world = 9 + 8 * 16 + 9 * 256
    node = np.array([0.125, 0.0, 0.125])
    c, r, vs = (0.1, 0.05, 0.1), 0.2, (0.5, -1.0, 0.25)
    m = np.zeros(G ** 3)
    pg = np.zeros((G ** 3, 3))
    m[world] = 0.25
    pg[world] = (0.01, 0.02, -0.03)
    dt = 1e-3
    m2, p2, v, mask, wrench = orc.rigid_grid(cfg, dt, body_rows(spheres=[(c, r, vs)]), m, pg)
    assert m2[world] == 1.0e6 and mask[world]
    assert np.allclose(p2[world], 1.0e6 * np.asarray(vs, np.float32))
    f = -(0.25 * np.asarray(vs, np.float32) - pg[world]) / dt
    assert np.allclose(wrench[0][:3], f, rtol=1e-6)
    assert np.allclose(wrench[0][3:], cross(node - np.asarray(c, np.float32), f), rtol=1e-6)
    # empty cells inside a legacy sphere are stamped too, with no reaction
    inside_empty = np.where(mask & (m == 0))[0]
    assert inside_empty.size > 0 and np.all(m2[inside_empty] == 1.0e6)


def test_legacy_sphere_shell_taper():
    # mpm_rigid_force_integrate: a cell at d = radius + dx/2 has weight 1/2
    cfg = config()
    G, dx = 16, 0.125
    world = 9 + 8 * 16 + 9 * 256
    node = np.array([0.125, 0.0, 0.125])
    r = 0.1
    center = node - np.array([r + 0.5 * dx, 0.0, 0.0])
    vs = np.array([0.5, 0.0, 0.0])
    m = np.zeros(G ** 3)
    pg = np.zeros((G ** 3, 3))
    m[world], pg[world] = 0.5, (0.1, 0.2, 0.0)
    dt = 1e-3
    cfg = cfg.copy(gravity_y=0.0)
    _, _, v, mask, wrench = orc.rigid_grid(cfg, dt, body_rows(spheres=[(tuple(center), r, tuple(vs))]), m, pg)
    assert not mask[world]
    vcell = v[world]
    assert np.allclose(vcell, [0.2, 0.4, 0.0])
    f = -(0.5 * 0.5 / dt) * (vs - vcell)
    assert np.allclose(wrench[0][:3], f, rtol=1e-5)
    assert np.allclose(wrench[0][3:], cross(node - center, f), rtol=1e-5, atol=1e-9)


def test_stamped_cells_skip_the_implicit_sweeps():
    cfg = config().copy(time_step=Implicit(3))
    G = 16
    m = np.zeros(G ** 3)
    pg = np.zeros((G ** 3, 3))
    idx = lambda x, y, z: x + G * (y + G * z)  # noqa: E731
    for x in range(6, 11):
        m[idx(x, 8, 8)] = 0.1
        pg[idx(x, 8, 8)] = (0.05 * (x - 6), 0.0, 0.0)
    sphere = (tuple(-1 + 0.125 * np.array([8, 8, 8])), 0.05, (2.0, 0.0, 0.0))
    _, _, v, mask, _ = orc.rigid_grid(cfg, 1e-3, body_rows(spheres=[sphere]), m, pg)
    _, _, v0, mask0, _ = orc.rigid_grid(cfg, 1e-3, body_rows(spheres=[]), m, pg)
    c = idx(8, 8, 8)
    assert mask[c] and not mask0.any()
    assert np.allclose(v[c], [2.0, 0.0, 0.0]), "the stamped cell is not smoothed"
    assert not np.isclose(v0[c][0], 2.0)
    # its neighbour is smoothed towards the stamped velocity
    assert v[idx(7, 8, 8)][0] > v0[idx(7, 8, 8)][0] + 1e-6


def test_rotated_box_containment_matches_rotation_matrix():
    rng = np.random.default_rng(0)
    cfg = config().copy(gravity_y=0.0)
    G = 16
    for trial in range(4):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        half = rng.uniform(0.15, 0.45, 3)
        center = rng.uniform(-0.2, 0.2, 3)
        body = MpmPrescribedSolid(shape=SolidBox(tuple(half)), center=tuple(center),
                                  orientation=tuple(q.astype(np.float32)))
        m = np.full(G ** 3, 0.01)
        pg = np.zeros((G ** 3, 3))
        _, _, _, mask, _ = orc.rigid_grid(cfg, 1e-3, body_rows(solids=[body]), m, pg)
        qx, qy, qz, qw = q
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                      [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                      [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
        ii = np.arange(G ** 3)
        nodes = -1 + 0.125 * np.stack([ii % G, ii // G % G, ii // G // G], 1)
        local = (nodes - center) @ R  # R^T (node - c)
        margin = np.abs(np.abs(local) - half).min(axis=1)
        inside = np.all(np.abs(local) < half, axis=1)
        sure = margin > 1e-5
        assert np.array_equal(mask[sure].astype(bool), inside[sure]), trial
        assert inside.any() and (~inside).any()
