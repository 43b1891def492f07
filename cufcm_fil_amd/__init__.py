"""MI355X-native wall-corrected RPY mobility + GMRES/Broyden implicit step.

The compute path is the HIP library ``libcufcm_fil.so`` (C ABI in
``include/cufcm_fil.h``); this package is its host-side Python binding.
"""
from ._lib import (CFM_BB, CFM_BS, CFM_FP32C, CFM_FP64, CFM_FULL, CFM_SB, CFM_SS, CfmError,
                   header_symbols)
from .mobility import RpyMobilitySolver

__all__ = ["RpyMobilitySolver", "CfmError", "header_symbols", "CFM_SS", "CFM_SB", "CFM_BB",
           "CFM_BS", "CFM_FULL", "CFM_FP64", "CFM_FP32C"]
