"""elsim_amd -- MI355X-native SBM solver core (HIP kernels for gfx950 behind a
C-ABI; see include/elsim.h).

Python here is host plumbing: the ctypes binding of libelsim_hip.so, synthetic
inputs and the Fortran record format.  Every numerical operation runs in the
HIP library; nothing falls back to the CPU.
"""
from ._lib import (MODE_FP64, MODE_FP64_PCG, MODE_MIXED, MODE_MIXED_MG, OP_CNE, OP_PHE, OP_PHP, STOP_REFERENCE, STOP_RELRES, WS_EXTRAP,
                   WS_LEARNED, WS_PLAIN, DeviceArray,
                   ElsimError, SbmParams, StepStats, check, exported_symbols, lib)
from . import fileio
from .sbm import (Sbm, carve_tunnel, ellipsoid_pack, initial_cnp, potential_problem, reference_params, solve_pot,
                  sphere_pack, synthetic_params, tables, time_control)

__all__ = [
    "MODE_FP64", "MODE_FP64_PCG", "MODE_MIXED", "MODE_MIXED_MG", "OP_CNE", "OP_PHP", "OP_PHE", "STOP_REFERENCE", "STOP_RELRES", "WS_EXTRAP",
    "WS_LEARNED", "WS_PLAIN", "DeviceArray",
    "ElsimError", "SbmParams", "StepStats", "check", "exported_symbols", "lib", "fileio", "Sbm", "carve_tunnel",
    "ellipsoid_pack", "initial_cnp", "reference_params", "sphere_pack", "synthetic_params", "tables",
    "time_control", "potential_problem", "solve_pot",
]
