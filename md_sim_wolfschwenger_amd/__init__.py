"""MI355X (gfx950) Brownian/LLG step for magnetic nanoparticle suspensions.

The compute path is libmdsim.so (hand-written HIP kernels + C++ host driver
behind the C ABI of include/mdsim_capi.h); this package is its Python binding.
Nothing here computes physics on the CPU: without the built library every
entry point raises.
"""
from .capi import (BOGACKI_SHAMPINE, DORMAND_PRINCE, HEUN_EULER, PREC_FP64, PREC_MIXED, Config, Context,
                   MdsError, config_from_ini, config_from_string, device_count, init_coords_from_sim,
                   init_orientations, init_positions, init_radii, lib)

__all__ = ["Config", "Context", "MdsError", "config_from_ini", "config_from_string", "device_count", "lib",
           "init_radii", "init_positions", "init_orientations", "init_coords_from_sim",
           "HEUN_EULER", "BOGACKI_SHAMPINE", "DORMAND_PRINCE", "PREC_FP64", "PREC_MIXED"]
