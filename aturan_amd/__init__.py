"""aturan_amd: the MI355X dynamical-core step of ATURAN (RHS + RK4 + BC + red-black Poisson).

The compute path is libaturan_amd.so (hand-written HIP for gfx950) behind the C ABI in
include/atu.h. This package is its Python host binding; it has no CPU fallback.
"""
from ._lib import FIELD, FIELD_NAMES, NFIELDS, NFROZEN, NPROG, LIB_PATH, load  # noqa: F401
from .core import (AtuError, SolveResult, UranusCore, cp_table, default_grid, default_params,  # noqa: F401
                   synthetic_state)

__all__ = ["UranusCore", "AtuError", "SolveResult", "default_grid", "default_params", "synthetic_state",
           "cp_table", "FIELD", "FIELD_NAMES", "NFIELDS", "NPROG", "NFROZEN", "LIB_PATH", "load"]
