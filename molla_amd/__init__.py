"""molla_amd -- MI355X-native MPM substep for molla (HIP/gfx950 behind a C ABI).

``MpmConfig`` and the rheology types mirror the reference configuration;
``MpmSimulation`` is the stepping API, backed by ``libmolla_mpm.so``.
"""
from .config import (MODEL_CPUDIFF, MODEL_SHADER, Explicit, Fluid, Implicit, MpmConfig,
                     Sand, Snow, Viscous, material_row)
from .mpm import (P2G_DETERMINISTIC, P2G_FUSED, P2G_GLOBAL_ATOMIC, MpmError, MpmSimulation)
from .checkpoint import CheckpointError, MpmSoilCheckpoint

__all__ = ["MpmConfig", "Sand", "Snow", "Viscous", "Fluid", "Explicit", "Implicit",
           "MODEL_SHADER", "MODEL_CPUDIFF", "material_row", "MpmSimulation", "MpmError",
           "P2G_FUSED", "P2G_GLOBAL_ATOMIC", "P2G_DETERMINISTIC", "MpmSoilCheckpoint", "CheckpointError"]
