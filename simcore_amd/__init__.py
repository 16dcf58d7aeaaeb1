"""simcore_amd -- MI355X-native filament Brownian-dynamics hot path of simcore.

The compute path is libsimcore_amd.so (HIP kernels for gfx950 behind the C
ABI in include/simcore_amd.h).  Importing this package loads it and raises
ImportError if it has not been built; there is no CPU fallback.
"""
from ._lib import (SC_PREC_FP64, SC_PREC_MIXED, EXPORTS, LIB_PATH, ScFilamentParams,
                   ScParams, SimcoreError, UnphysicalResult, lib)
from .system import (FilamentSystem, Filaments, crosslink_params, filament_layout,
                     generate_crosslinks, generate_filaments, params_from_yaml, params_init)
from .ensemble import Ensemble

__all__ = [
    "FilamentSystem", "Filaments", "Ensemble", "generate_crosslinks", "crosslink_params", "generate_filaments", "params_init",
    "params_from_yaml", "filament_layout", "ScParams", "ScFilamentParams",
    "SimcoreError", "UnphysicalResult", "SC_PREC_FP64", "SC_PREC_MIXED",
    "EXPORTS", "LIB_PATH", "lib",
]
