"""Rigid bodies on the MPM grid: the host side of ``set_rigid_spheres``
(mpm.rs:1521-1566), ``set_prescribed_solids`` (mpm_prescribed_solids.rs:91-140)
and ``read_rigid_wrenches_with`` (mpm.rs:2908-2955).

Shapes and the prescribed solid mirror ``MpmSolidShape`` / ``MpmPrescribedSolid``
(mpm_prescribed_solids.rs:5-32): a solid's pose is its center and a unit
quaternion (x, y, z, w); a grid node inside it sticks to
``linear_velocity + angular_velocity x (node - center)``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Sequence, Tuple

SHAPE_SPHERE, SHAPE_BOX, SHAPE_CYLINDER = 1, 2, 3


@dataclasses.dataclass(frozen=True)
class SolidSphere:
    radius: float


@dataclasses.dataclass(frozen=True)
class SolidBox:
    half_extents: Tuple[float, float, float]


@dataclasses.dataclass(frozen=True)
class SolidCylinder:
    """Finite cylinder with its axle along local y."""

    radius: float
    half_width: float


@dataclasses.dataclass(frozen=True)
class MpmPrescribedSolid:
    shape: object
    center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    orientation: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 1.0)
    linear_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    angular_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def replace(self, **kw) -> "MpmPrescribedSolid":
        return dataclasses.replace(self, **kw)


class PrescribedSolid(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("dims", C.c_float * 3), ("center", C.c_float * 3),
                ("orientation", C.c_float * 4), ("linear_velocity", C.c_float * 3),
                ("angular_velocity", C.c_float * 3)]


SIGS = {
    "mpm_set_rigid_spheres": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_set_prescribed_solids": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "mpm_read_rigid_wrenches": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
}


def solid_struct(b: MpmPrescribedSolid) -> PrescribedSolid:
    s = PrescribedSolid()
    sh = b.shape
    if isinstance(sh, SolidSphere):
        s.shape, dims = SHAPE_SPHERE, (sh.radius, 0.0, 0.0)
    elif isinstance(sh, SolidBox):
        s.shape, dims = SHAPE_BOX, tuple(sh.half_extents)
    elif isinstance(sh, SolidCylinder):
        s.shape, dims = SHAPE_CYLINDER, (sh.radius, sh.half_width, 0.0)
    else:
        raise TypeError(f"unknown solid shape {sh!r}")
    for k in range(3):
        s.dims[k] = float(dims[k])
        s.center[k] = float(b.center[k])
        s.linear_velocity[k] = float(b.linear_velocity[k])
        s.angular_velocity[k] = float(b.angular_velocity[k])
    for k in range(4):
        s.orientation[k] = float(b.orientation[k])
    return s


def body_rows(spheres=None, solids=None) -> List[List[float]]:
    """The five rows per body that the kernels read (also the oracle's input):
    center + bounding radius, linear velocity, angular velocity, quaternion,
    shape dims + kind (0 legacy sphere, 1 sphere, 2 box, 3 cylinder)."""
    import math

    import numpy as np

    rows = []
    for c, r, v in spheres or []:
        rows += [[*c, r], [*v, 0.0], [0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0, 0.0]]
    for b in solids or []:
        q = np.asarray(b.orientation, np.float32)
        norm = float(np.sqrt(np.float32((q * q).sum())))
        sh = b.shape
        if isinstance(sh, SolidSphere):
            dims, rad = [sh.radius] * 3 + [1.0], sh.radius
        elif isinstance(sh, SolidBox):
            dims, rad = list(sh.half_extents) + [2.0], math.sqrt(sum(x * x for x in sh.half_extents))
        else:
            dims, rad = [sh.radius, sh.half_width, sh.radius, 3.0], math.hypot(sh.radius, sh.half_width)
        rows += [[*b.center, rad], [*b.linear_velocity, 0.0], [*b.angular_velocity, 0.0],
                 [float(np.float32(x) / np.float32(norm)) for x in b.orientation], dims]
    return rows
