"""Host-side configuration types mirroring molla's MPM API.

``MpmConfig``, ``MpmRheology`` and ``MpmTimeStep`` mirror
``crates/molla-solvers/src/mpm.rs:320-433`` (fields, units and presets
``sand`` / ``sand_calibrated`` / ``snow`` / ``viscous`` at :546-655).
``material_row`` packs a rheology exactly like the reference's
``material_row`` (mpm.rs:666-686): ``(rheology_id, dp_alpha, theta_c,
theta_s, hardening, velocity_damping, affine_damping, bulk)``; the eighth
slot ``bulk`` carries this build's weakly-compressible fluid modulus
(rheology id 3), which has no reference counterpart.

Two additions select semantics that the reference keeps in separate code
paths:

* ``model = MODEL_SHADER`` (default) is the GPU step of ``mpm.wgsl``:
  fixed-corotated stress with sigma clamped to [0.1, 10], six padded walls
  with ``boundary_friction``, a position clamp and the plasticity pass.
* ``model = MODEL_CPUDIFF`` is the differentiable CPU step of
  ``molla-autograd/src/mpm_cpu_diff.rs``: corotated Saint-Venant-Kirchhoff
  stress, a floor plane at ``floor_y`` with ``floor_friction`` (1 = sticky),
  no clamps and no plasticity.  The adjoint is defined for this model.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import List, Tuple, Union

MODEL_SHADER = 0
MODEL_CPUDIFF = 1

RHEO_SAND = 0
RHEO_SNOW = 1
RHEO_VISCOUS = 2
RHEO_FLUID = 3

MAX_MPM_MATERIALS = 8  # mpm.rs:661


@dataclass(frozen=True)
class Sand:
    """Drucker-Prager sand; ``friction_angle`` in radians (mpm.rs:324-328)."""

    friction_angle: float = math.radians(30.0)


@dataclass(frozen=True)
class Snow:
    """Stomakhin snow envelope [1-theta_c, 1+theta_s] (mpm.rs:329-338)."""

    theta_c: float = 2.5e-2
    theta_s: float = 7.5e-3
    hardening: float = 10.0


@dataclass(frozen=True)
class Viscous:
    """Exponential velocity / affine damping (mpm.rs:339-348).

    ``Viscous(0, 0)`` is the reference's closest row to a purely elastic
    material: no projection, sigma clamped to [0.05, 20] and F rebuilt.
    """

    velocity_damping: float = 0.0
    affine_damping: float = 0.0


@dataclass(frozen=True)
class Fluid:
    """Weakly-compressible fluid with bulk modulus ``bulk`` (Pa).

    Not in the reference.  Defined as fixed-corotated with mu = 0,
    lambda = bulk (Kirchhoff stress ``bulk * (J - 1) * J * I``) and F reset to
    ``cbrt(det F) * I`` (clamped to [0.05, 20]) after every substep.
    """

    bulk: float = 1.0e4


MpmRheology = Union[Sand, Snow, Viscous, Fluid]


def drucker_prager_alpha(friction_angle: float) -> float:
    """alpha = sqrt(2/3) * 2 sin(phi) / (3 - sin(phi)), f32 like mpm.rs:670-671."""
    import numpy as np

    s = np.float32(math.sin(np.float32(friction_angle)))
    return float(np.float32(np.sqrt(np.float32(2.0) / np.float32(3.0))) * (np.float32(2.0) * s)
                 / (np.float32(3.0) - s))


def material_row(r: MpmRheology) -> Tuple[float, ...]:
    """Pack one rheology row (mpm.rs:666-686, plus the fluid slot)."""
    if isinstance(r, Sand):
        return (0.0, drucker_prager_alpha(r.friction_angle), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    if isinstance(r, Snow):
        return (1.0, 0.0, r.theta_c, r.theta_s, r.hardening, 0.0, 0.0, 0.0)
    if isinstance(r, Viscous):
        return (2.0, 0.0, 0.0, 0.0, 0.0, r.velocity_damping, r.affine_damping, 0.0)
    if isinstance(r, Fluid):
        return (3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, r.bulk)
    raise TypeError(f"unknown rheology {r!r}")


@dataclass(frozen=True)
class Explicit:
    """Explicit symplectic step (mpm.rs:377-381)."""


@dataclass(frozen=True)
class Implicit:
    """Semi-implicit Jacobi step (mpm.rs:382-392): ``jacobi_iters`` grid sweeps per substep."""

    jacobi_iters: int = 3


@dataclass
class MpmConfig:
    """Mirror of ``MpmConfig`` (mpm.rs:394-433) plus model/floor/window fields."""

    grid_size: int = 32
    cell_size: float = 0.05
    grid_origin: Tuple[float, float, float] = (-0.8, 0.0, -0.8)
    gravity_y: float = -9.81
    youngs: float = 5.0e3
    poisson: float = 0.3
    friction_angle: float = math.radians(30.0)
    cohesion: float = 0.0
    boundary_friction: float = 0.5
    particle_mass: float = 1.0e-3
    particle_volume: float = 1.0e-6
    max_particles: int = 16_384
    max_rigid: int = 8
    rheology: MpmRheology = field(default_factory=Sand)
    time_step: object = field(default_factory=Explicit)
    # --- additions ---
    model: int = MODEL_SHADER
    floor_y: float = 0.0
    floor_friction: float = 1.0
    grid_window_cells: Tuple[int, int, int] = (0, 0, 0)
    materials: List[MpmRheology] = field(default_factory=list)

    def copy(self, **kw) -> "MpmConfig":
        return dataclasses.replace(self, **kw)

    def material_rows(self) -> List[Tuple[float, ...]]:
        """The per-step material table (mpm.rs:2406-2420)."""
        if not self.materials:
            return [material_row(self.rheology)]
        if len(self.materials) > MAX_MPM_MATERIALS:
            raise ValueError(f"at most {MAX_MPM_MATERIALS} materials")
        return [material_row(r) for r in self.materials]

    # ---- presets (mpm.rs:546-655) ----
    @staticmethod
    def sand() -> "MpmConfig":
        return MpmConfig()

    @staticmethod
    def sand_calibrated() -> "MpmConfig":
        return MpmConfig(grid_size=48, grid_origin=(-1.20, 0.0, -1.20), youngs=1.0e5,
                         boundary_friction=0.3, particle_mass=1.5e-3, max_particles=32_768,
                         time_step=Implicit(3))

    @staticmethod
    def snow() -> "MpmConfig":
        return MpmConfig(youngs=1.4e5, poisson=0.2, friction_angle=0.0, rheology=Snow())

    @staticmethod
    def viscous() -> "MpmConfig":
        return MpmConfig(youngs=2.0e3, friction_angle=0.0, boundary_friction=0.8,
                         rheology=Viscous(4.0, 8.0))

    @staticmethod
    def cpu_diff_default() -> "MpmConfig":
        """``MpmCpuConfig::default()`` (mpm_cpu_diff.rs:107-122) in this type."""
        return MpmConfig(grid_size=16, cell_size=0.1, grid_origin=(-0.8, -0.05, -0.8),
                         gravity_y=-9.81, youngs=1.0e4, poisson=0.3, particle_volume=1.0e-3,
                         particle_mass=1.0e-2, model=MODEL_CPUDIFF, floor_y=0.0,
                         floor_friction=0.6, rheology=Viscous())
