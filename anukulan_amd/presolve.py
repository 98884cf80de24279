"""Presolve of a case stack (or of one model) in front of the solvers, and
postsolve of their solutions: the C++ behind anuk_presolve_stack, standing
where the reference's presolve and PostsolveStack stand
(src/anukulan/presolve.hpp:485 and 389-396), but taken over a whole stack of
cases that share A, so that the reduced cases share one matrix again and
still run as one batch.

    ps = presolve_stack(base, cases)            # NaturalModel + generators.Case list
    lps = [m.to_lp() for m in ps.models()]      # reduced cases, one shared K
    rs = solve_pdhg_batch(ctx, lps, tolerance=1e-4)
    x, y = ps.postsolve_result(k, lps[k], rs[k])  # case k of the original stack

Reductions: empty, free and singleton rows, redundant rows and redundant
sides of rows, fixed and empty columns, and (with slack_column_singletons=1)
a column alone in an equality row substituted out, its bounds becoming the
row's. A structural reduction is taken only when every case allows it; the
bounds, costs and offsets it moves are moved per case. The substitution is
off by default because it makes its row a ranged row, two opposed rows of
the standard form: on the 24-period case stack it removes 336 columns and
costs the batched PDHG 34% more iterations to 1e-4.

Row duals here are per model row, in the minimisation sense of the LP the
library builds (maximisation negates the objective): positive when the row
holds at its lower side, negative at its upper side; c - A'y are the reduced
costs. model_row_duals / standard_duals convert to and from the duals of the
standard form the solvers return. Every reduction made here inverts exactly
on the dual side, so the postsolved duals are the original LP's, not an
estimate.
"""
from __future__ import annotations

from ctypes import byref, c_void_p
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import lib
from .api import LP, AnukError, SolveResult, _dp, _f64, _ip

STATUS = {0: "reduced", 1: "infeasible", 2: "unbounded"}
RULES = ("empty_rows", "singleton_rows", "redundant_rows", "fixed_columns", "empty_columns",
         "slack_column_singletons", "max_rounds", "tolerance", "pivot_ratio")
COUNTS = ("empty_rows", "free_rows", "singleton_rows", "redundant_rows", "redundant_sides",
          "fixed_columns", "empty_columns", "slack_column_singletons", "rounds")
FILE_INF = 1e30


def _bounds(v) -> np.ndarray:
    """Bounds at or beyond 1e30 in magnitude are infinite, as in the
    standard form (standard_form.cpp)."""
    b = _f64(v).copy()
    b[b >= FILE_INF] = np.inf
    b[b <= -FILE_INF] = -np.inf
    return b


def _csr(rows, cols, ri, ci, v):
    """The model's triplets as CSR with duplicates summed, exact zeros
    dropped and columns sorted in each row (sparse.cpp:24-70)."""
    import scipy.sparse as sp
    a = sp.coo_matrix((_f64(v), (np.asarray(ri, np.int64), np.asarray(ci, np.int64))),
                      shape=(rows, cols)).tocsr()
    a.sum_duplicates()
    a.eliminate_zeros()
    a.sort_indices()
    return (np.ascontiguousarray(a.indptr, np.int32), np.ascontiguousarray(a.indices, np.int32),
            np.ascontiguousarray(a.data, np.float64))


class PresolvedStack:
    """The reduced stack and the record that maps its solutions back."""

    def __init__(self, handle, sign: float, name: str, cases: int):
        self.h = handle
        self.sign = sign
        info = _lib.PresolveInfo()
        lib.anuk_presolve_get_info(self.h, byref(info))
        self.status = STATUS[info.status]
        self.failing_case = info.failing_case
        self.message = lib.anuk_presolve_message(self.h).decode()
        self.rows, self.cols, self.nnz = info.rows, info.cols, info.nnz
        self.original_rows, self.original_cols = info.original_rows, info.original_cols
        self.original_nnz = info.original_nnz
        self.counts = {k: getattr(info, k) for k in COUNTS}
        self.name = name
        self.cases = cases
        if self.status == "reduced":
            self._fetch()

    def _fetch(self):
        self.start = np.zeros(self.rows + 1, np.int32)
        self.index = np.zeros(max(self.nnz, 1), np.int32)
        self.value = np.zeros(max(self.nnz, 1))
        k = self.cases
        self.objective = np.zeros((k, max(self.cols, 1)))
        self.col_lower, self.col_upper = np.zeros_like(self.objective), np.zeros_like(self.objective)
        self.row_lower = np.zeros((k, max(self.rows, 1)))
        self.row_upper = np.zeros_like(self.row_lower)
        self.offset = np.zeros(k)
        self.row_origin = np.zeros(max(self.rows, 1), np.int32)
        self.col_origin = np.zeros(max(self.cols, 1), np.int32)
        rc = lib.anuk_presolve_reduced(self.h, _ip(self.start), _ip(self.index), _dp(self.value),
                                       _dp(self.objective), _dp(self.row_lower), _dp(self.row_upper),
                                       _dp(self.col_lower), _dp(self.col_upper), _dp(self.offset),
                                       _ip(self.row_origin), _ip(self.col_origin))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        self.index, self.value = self.index[:self.nnz], self.value[:self.nnz]
        self.objective, self.col_lower = self.objective[:, :self.cols], self.col_lower[:, :self.cols]
        self.col_upper = self.col_upper[:, :self.cols]
        self.row_lower, self.row_upper = self.row_lower[:, :self.rows], self.row_upper[:, :self.rows]
        self.row_origin, self.col_origin = self.row_origin[:self.rows], self.col_origin[:self.cols]

    def __del__(self):
        if getattr(self, "h", None):
            lib.anuk_presolve_free(self.h)
            self.h = None

    def _need_reduced(self):
        if self.status != "reduced":
            raise ValueError(f"presolve found the stack {self.status}: {self.message}")

    def model(self, k: int = 0):
        """Case k of the reduced stack as a NaturalModel (the original's
        objective sense; triplets in CSR order)."""
        from .generators import NaturalModel
        self._need_reduced()
        ri = np.repeat(np.arange(self.rows, dtype=np.int32), np.diff(self.start))
        return NaturalModel(
            rows=self.rows, cols=self.cols, row_index=ri, col_index=self.index.copy(),
            value=self.value.copy(), objective=self.sign * self.objective[k],
            row_lower=self.row_lower[k].copy(), row_upper=self.row_upper[k].copy(),
            col_lower=self.col_lower[k].copy(), col_upper=self.col_upper[k].copy(),
            maximize=self.sign < 0, objective_offset=float(self.sign * self.offset[k]),
            name=f"{self.name}-presolved-{k}")

    def models(self) -> list:
        return [self.model(k) for k in range(self.cases)]

    def postsolve(self, k: int, x, y):
        """Case k's reduced primal x and model-row duals y mapped to the
        original model's columns and rows."""
        self._need_reduced()
        rx, ry = _f64(x), _f64(y)
        if rx.size != self.cols or ry.size != self.rows:
            raise ValueError(f"postsolve: expected {self.cols} columns and {self.rows} row duals")
        xo, yo = np.zeros(max(self.original_cols, 1)), np.zeros(max(self.original_rows, 1))
        rc = lib.anuk_postsolve(self.h, int(k), _dp(rx), _dp(ry), _dp(xo), _dp(yo))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        return xo[:self.original_cols], yo[:self.original_rows]

    def postsolve_result(self, k: int, reduced_lp: LP, result: SolveResult):
        """postsolve of a solver's result on case k's reduced LP (its duals
        are the standard form's, converted per model row first)."""
        return self.postsolve(k, result.x, model_row_duals(reduced_lp, result.y, self.rows))

    def summary(self) -> str:
        fired = ", ".join(f"{k} {v}" for k, v in self.counts.items() if v and k != "rounds")
        return (f"{self.name}: {self.original_rows}x{self.original_cols} ({self.original_nnz} nnz) -> "
                f"{self.rows}x{self.cols} ({self.nnz} nnz) over {self.cases} case(s) in "
                f"{self.counts['rounds']} rounds; {fired or 'nothing fired'}")


def default_rules(**overrides) -> _lib.PresolveRules:
    r = _lib.PresolveRules()
    lib.anuk_presolve_default_rules(byref(r))
    for k, v in overrides.items():
        if k not in RULES:
            raise TypeError(f"unknown presolve rule {k!r}")
        setattr(r, k, v)
    return r


def presolve_stack(base, cases: Optional[Sequence] = None, **rules) -> PresolvedStack:
    """Presolves the cases (generators.Case: objective, row and column
    bounds on base's matrix; None: base alone). Raises AnukError for
    malformed input; an infeasible or unbounded stack comes back with that
    status and the case that proved it."""
    sign = -1.0 if base.maximize else 1.0
    start, index, value = _csr(base.rows, base.cols, base.row_index, base.col_index, base.value)
    items = list(cases) if cases is not None else [base]
    k = len(items)
    if k == 0:
        raise ValueError("presolve_stack: no cases")

    def stack(field, n, bounds):
        out = np.empty((k, max(n, 1)))
        for i, c in enumerate(items):
            v = _f64(getattr(c, field))
            if v.size != n:
                raise ValueError(f"presolve_stack: case {i} has {v.size} {field} entries, not {n}")
            out[i, :n] = _bounds(v) if bounds else v
        return np.ascontiguousarray(out)

    obj = sign * stack("objective", base.cols, False)
    rl, ru = stack("row_lower", base.rows, True), stack("row_upper", base.rows, True)
    cl, cu = stack("col_lower", base.cols, True), stack("col_upper", base.cols, True)
    off = np.full(k, sign * float(base.objective_offset))
    h = c_void_p()
    r = default_rules(**rules)
    rc = lib.anuk_presolve_stack(base.rows, base.cols, _ip(start), _ip(index), _dp(value), k,
                                 _dp(obj), _dp(rl), _dp(ru), _dp(cl), _dp(cu), _dp(off), byref(r),
                                 byref(h))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    return PresolvedStack(h.value, sign, base.name or "model", k)


def presolve(model, **rules) -> PresolvedStack:
    """A stack of one."""
    return presolve_stack(model, None, **rules)


def model_row_duals(lp: LP, y_standard, model_rows: int) -> np.ndarray:
    """Standard-form duals (one per standard row, >= 0 on inequalities) summed
    per model row with their signs: a ranged row's two standard rows give
    one dual, and a model row with no standard row (a free row) gets 0."""
    e = lp.export()
    y = np.zeros(int(model_rows))
    np.add.at(y, e["row_model"], e["row_sign"] * _f64(y_standard))
    return y


def standard_duals(lp: LP, y_model) -> np.ndarray:
    """Model-row duals split over the standard rows: an equality takes the
    dual as it is; an inequality row takes the part its sign allows."""
    e = lp.export()
    ym = _f64(y_model)
    v = e["row_sign"] * ym[e["row_model"]]
    eq = np.arange(lp.rows) < lp.num_equalities
    return np.where(eq, v, np.maximum(v, 0.0))
