"""MI355X-native backend for Anukulan's first-order LP method (PDHG).

The compute lives in libanukulan_hip.so (HIP kernels for gfx950 and the C++
host side behind the C ABI in include/anukulan_hip.h); this package loads it
and exposes it to Python. Importing raises if the library was not built.
"""
from .api import (LP, AdjointResult, AnukError, Context, DeviceVector, Matrix, SolveResult,
                  adjoint_options, default_options, estimate_matrix_norm, pdhg_adjoint, solve_pdhg,
                  solve_pdhg_batch, solve_pdhg_multi)
from ._lib import LIB_PATH

__all__ = ["LP", "AdjointResult", "AnukError", "Context", "DeviceVector", "Matrix", "SolveResult",
           "adjoint_options", "default_options", "estimate_matrix_norm", "pdhg_adjoint",
           "solve_pdhg", "solve_pdhg_batch", "solve_pdhg_multi", "LIB_PATH"]
