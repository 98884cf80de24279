"""MI355X-native backend for Anukulan's first-order LP method (PDHG).

The compute lives in libanukulan_hip.so (HIP kernels for gfx950 and the C++
host side behind the C ABI in include/anukulan_hip.h); this package loads it
and exposes it to Python. Importing raises if the library was not built.
"""
from .api import (LP, AnukError, Context, DeviceVector, Matrix, SolveResult, default_options,
                  estimate_matrix_norm, solve_pdhg, solve_pdhg_batch)
from ._lib import LIB_PATH

__all__ = ["LP", "AnukError", "Context", "DeviceVector", "Matrix", "SolveResult",
           "default_options", "estimate_matrix_norm", "solve_pdhg", "solve_pdhg_batch", "LIB_PATH"]
