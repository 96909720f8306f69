"""molla_amd -- MI355X-native MPM substep for molla (HIP/gfx950 behind a C ABI)."""
from .config import (MODEL_CPUDIFF, MODEL_SHADER, Explicit, Fluid, Implicit, MpmConfig,
                     Sand, Snow, Viscous, material_row)

__all__ = ["MpmConfig", "Sand", "Snow", "Viscous", "Fluid", "Explicit", "Implicit",
           "MODEL_SHADER", "MODEL_CPUDIFF", "material_row"]
