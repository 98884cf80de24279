"""Presolve and postsolve through the library (anuk_presolve_*, the
reference's presolve and PostsolveStack: src/anukulan/presolve.hpp:54-489).

    r = presolve_mps("plan.mps")          # or presolve(model) for a model in memory
    if r.status == "reduced":
        lp = r.reduced.to_lp()            # solve this
        x = r.postsolve_primal(x_reduced) # back to the original columns
        y = r.postsolve_dual(y_reduced, reduced_costs, original_objective)

A model presolve proves infeasible or unbounded is a status, not an error.
"""
from __future__ import annotations

from ctypes import byref, c_void_p
from typing import Optional

import numpy as np

from . import _lib
from ._lib import lib
from .api import AnukError, _dp, _f64, _i32, _ip
from .mps import MpsModel, _take, options as mps_options

STATUS = {0: "reduced", 1: "infeasible", 2: "unbounded"}
FLAGS = [f for f, _ in _lib.PresolveOptions._fields_ if f != "max_rounds" and not f.endswith(
    ("tolerance", "relaxation", "bound"))]


def options(**overrides) -> _lib.PresolveOptions:
    """PresolveOptions at the reference's defaults, with overrides by name."""
    o = _lib.PresolveOptions()
    lib.anuk_presolve_default_options(byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"no presolve option {k!r}")
        setattr(o, k, int(v) if k in FLAGS or k == "max_rounds" else float(v))
    return o


class PresolveResult:
    """A presolve's outcome: status, message, counts, the reduced model and
    the postsolve maps back to the original model."""

    def __init__(self, handle):
        self.h = handle
        info = _lib.PresolveInfo()
        lib.anuk_presolve_info_get(self.h, byref(info))
        self.status = STATUS[info.status]
        self.message = lib.anuk_presolve_message(self.h).decode()
        self.dual_exact = bool(info.dual_exact)
        self.original_rows, self.original_cols = info.original_rows, info.original_cols
        self.original_nnz = info.original_nnz
        self.counts = {k: getattr(info, k) for k in _lib.PRESOLVE_COUNTS}
        self.counts["largest_new_finite_bound"] = info.largest_new_finite_bound
        red = lib.anuk_presolve_reduced(self.h)
        self.reduced: Optional[MpsModel] = _take(red, owned=False) if red else None

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.anuk_presolve_free(h)
            self.h = None

    def postsolve_primal(self, reduced_x) -> np.ndarray:
        rx = _f64(reduced_x)
        if self.reduced is None or rx.size != self.reduced.cols:
            raise ValueError("reduced x must have one entry per reduced column")
        x = np.zeros(self.original_cols)
        rc = lib.anuk_postsolve_primal(self.h, _dp(rx), _dp(x))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        return x

    def postsolve_dual(self, reduced_y, reduced_costs=None, original_cost=None) -> np.ndarray:
        ry = _f64(reduced_y)
        if self.reduced is None or ry.size != self.reduced.rows:
            raise ValueError("reduced y must have one entry per reduced row")
        rc = None if reduced_costs is None else _f64(reduced_costs)
        oc = None if original_cost is None else _f64(original_cost)
        if rc is not None and rc.size != self.reduced.cols:
            raise ValueError("reduced costs must have one entry per reduced column")
        if oc is not None and oc.size != self.original_cols:
            raise ValueError("the original cost must have one entry per original column")
        y = np.zeros(self.original_rows)
        code = lib.anuk_postsolve_dual(self.h, _dp(ry), _dp(rc), _dp(oc), _dp(y))
        if code != 0:
            raise AnukError(code, lib.anuk_last_error(None).decode())
        return y


def presolve(model, col_type=None, **opts) -> PresolveResult:
    """Presolves an MpsModel or a generators.NaturalModel held in memory (the
    reduced model carries no names; presolve_mps keeps them)."""
    if isinstance(model, MpsModel):
        ri, ci, v = model.row_index(), model.index, model.value
        col_type = model.col_type if col_type is None else col_type
    else:
        ri, ci, v = model.row_index, model.col_index, model.value
    ri, ci, v = _i32(ri), _i32(ci), _f64(v)
    arrs = [_f64(a) for a in (model.objective, model.row_lower, model.row_upper,
                              model.col_lower, model.col_upper)]
    m = _lib.Model(int(model.rows), int(model.cols), int(v.size), _ip(ri), _ip(ci), _dp(v),
                   _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4]),
                   1 if model.maximize else 0, float(model.objective_offset))
    ct = None if col_type is None else _i32(col_type)
    h = c_void_p()
    rc = lib.anuk_presolve_model(byref(m), _ip(ct) if ct is not None else None,
                                 byref(options(**opts)), byref(h))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    return PresolveResult(h)


def presolve_mps_text(text, source: str = "<text>", reader_options=None, **opts) -> PresolveResult:
    """Reads MPS text with the library's reader and presolves the model,
    names and column types included."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    f = c_void_p()
    rc = lib.anuk_mps_read_text(data, len(data), source.encode(),
                                byref(mps_options(**(reader_options or {}))), byref(f))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    try:
        h = c_void_p()
        rc = lib.anuk_presolve_mps(f, byref(options(**opts)), byref(h))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        return PresolveResult(h)
    finally:
        lib.anuk_mps_free(f)


def presolve_mps(path: str, **opts) -> PresolveResult:
    with open(path, "rb") as f:
        return presolve_mps_text(f.read(), str(path), **opts)
