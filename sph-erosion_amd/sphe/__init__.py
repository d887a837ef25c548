"""sphe -- MI355X-native SPH-Erosion simulation step (Python mirror of the C ABI).

The compute lives in libsphe.so (HIP kernels for gfx950 + C++ runtime); this
package binds its C ABI (include/sphe.h) for tests and benchmarks and mirrors
the reference's FluidSystemSPH interface (Erosion/fluid_system.h).
"""
from ._lib import (BUF_DENSITY, BUF_ID, BUF_MASS, BUF_POS, BUF_PRESS, BUF_SED, BUF_VEL, EROSION_ATOMIC,
                   EROSION_OFF, EROSION_TWO_PHASE, FP32, FP32_F64ACC, FP32_NAIVE, FP64, LIB_PATH, SIGNATURES, Params,
                   Particle, SpheError, TerrainDesc, lib)
from .core import Context, default_params, params_dict
from .fluid_system import FluidSystemSPH, Grid

__all__ = ["Context", "FluidSystemSPH", "Grid", "default_params", "params_dict", "lib", "LIB_PATH",
           "SIGNATURES", "Params", "Particle", "TerrainDesc", "SpheError", "FP32", "FP32_NAIVE", "FP64",
           "FP32_F64ACC",
           "EROSION_OFF", "EROSION_ATOMIC", "EROSION_TWO_PHASE", "BUF_POS", "BUF_VEL", "BUF_DENSITY",
           "BUF_MASS", "BUF_PRESS", "BUF_SED", "BUF_ID"]
