"""Python face of the MI355X PDHG backend.

Thin wrappers over the C ABI: a device context with the LinAlgBackend
operations (src/anukulan/backend.hpp:24-215), host LP handles built the way
to_standard_form builds them (src/anukulan/standard_form.cpp:72-165), and
solve_pdhg with the reference's options and result fields
(src/anukulan/pdhg.hpp:19-541). Errors raise AnukError; nothing here
computes anything itself.
"""
from __future__ import annotations

import ctypes
import itertools
from ctypes import byref, c_double, c_size_t, c_void_p
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import DP, IP, lib

STATUS_NAMES = {0: "optimal", 1: "primal infeasible", 2: "iteration limit", 3: "time limit",
                4: "numerical error"}
TERMINATION = {"full": 0, "primal": 1, "dual": 2}


class AnukError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"anukulan_hip error {code}: {message}")
        self.code = code
        self.message = message


def _dp(a: np.ndarray):
    return a.ctypes.data_as(DP) if a is not None and a.size else None


def _ip(a: np.ndarray):
    return a.ctypes.data_as(IP) if a is not None and a.size else None


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


_ids = itertools.count(1)


class Matrix:
    """A host CSR matrix with an identity for the device cache."""

    def __init__(self, rows: int, cols: int, start, index, value):
        self.rows, self.cols = int(rows), int(cols)
        self.start, self.index, self.value = _i32(start), _i32(index), _f64(value)
        if self.start.shape != (self.rows + 1,):
            raise ValueError("start must have rows+1 entries")
        self.id = next(_ids)
        self._struct = _lib.Csr(self.id, self.rows, self.cols, _ip(self.start),
                                _ip(self.index), _dp(self.value))
        if self._struct.start is None:  # rows == 0: still a valid pointer
            self._struct.start = self.start.ctypes.data_as(IP)

    @property
    def nnz(self) -> int:
        return int(self.start[-1])

    def struct(self):
        return self._struct

    @staticmethod
    def from_scipy(m) -> "Matrix":
        m = m.tocsr()
        m.sort_indices()
        return Matrix(m.shape[0], m.shape[1], m.indptr, m.indices, m.data)

    def transpose(self) -> "Matrix":
        """Counting-sort transpose (rows of the result ascending)."""
        order = np.argsort(self.index, kind="stable")
        rows_of = np.repeat(np.arange(self.rows, dtype=np.int32), np.diff(self.start))
        t_start = np.zeros(self.cols + 1, dtype=np.int64)
        np.add.at(t_start, self.index.astype(np.int64) + 1, 1)
        return Matrix(self.cols, self.rows, np.cumsum(t_start), rows_of[order], self.value[order])


class DeviceVector:
    def __init__(self, ctx: "Context", n: int, ptr: int):
        self.ctx, self.n, self.ptr = ctx, int(n), ptr

    def upload(self, host) -> "DeviceVector":
        h = _f64(host)
        if h.size != self.n:
            raise ValueError("size mismatch")
        self.ctx._call(lib.anuk_upload, _dp(h), self.ptr, self.n)
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        self.ctx._call(lib.anuk_download, self.ptr, _dp(out), self.n)
        return out

    def free(self):
        if self.ptr:
            self.ctx._call(lib.anuk_deallocate, self.ptr)
            self.ptr = None


class Context:
    """One device, one stream: the backend operations of LinAlgBackend."""

    def __init__(self, device: int = 0):
        h = c_void_p()
        rc = lib.anuk_create(int(device), byref(h))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            lib.anuk_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, fn, *args):
        rc = fn(self.h, *args)
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(self.h).decode())

    # memory
    def allocate(self, n: int) -> DeviceVector:
        p = c_void_p()
        self._call(lib.anuk_allocate, int(n), byref(p))
        return DeviceVector(self, n, p.value)

    def vector(self, host) -> DeviceVector:
        h = _f64(host)
        return self.allocate(h.size).upload(h)

    def fill(self, v: DeviceVector, value: float):
        self._call(lib.anuk_fill, v.ptr, v.n, float(value))

    def copy(self, src: DeviceVector, dst: DeviceVector):
        self._call(lib.anuk_copy, src.ptr, dst.ptr, src.n)

    def synchronize(self):
        self._call(lib.anuk_synchronize)

    # matrices
    def prepare(self, a: Matrix):
        self._call(lib.anuk_prepare, byref(a.struct()))

    def release(self):
        self._call(lib.anuk_release)

    def multiply(self, a: Matrix, x: DeviceVector, y: DeviceVector):
        self._call(lib.anuk_multiply, byref(a.struct()), x.ptr, y.ptr)

    def multiply_transpose(self, at: Matrix, y: DeviceVector, x: DeviceVector):
        self._call(lib.anuk_multiply_transpose, byref(at.struct()), y.ptr, x.ptr)

    # reductions
    def dot(self, a: DeviceVector, b: DeviceVector) -> float:
        out = c_double()
        self._call(lib.anuk_dot, a.ptr, b.ptr, a.n, byref(out))
        return out.value

    def two_norm(self, a: DeviceVector) -> float:
        out = c_double()
        self._call(lib.anuk_two_norm, a.ptr, a.n, byref(out))
        return out.value

    def inf_norm(self, a: DeviceVector) -> float:
        out = c_double()
        self._call(lib.anuk_inf_norm, a.ptr, a.n, byref(out))
        return out.value

    def primal_update(self, n, tau, gamma, halpern, x, c, kt_y, lower, upper, anchor, x_out, dx,
                      x_bar):
        self._call(lib.anuk_primal_update, int(n), float(tau), float(gamma), float(halpern), x.ptr,
                   c.ptr, kt_y.ptr, lower.ptr, upper.ptr, anchor.ptr if anchor is not None else None,
                   x_out.ptr, dx.ptr, x_bar.ptr)

    def dual_update(self, m, neq, sigma, gamma, halpern, y, q, k_x_bar, k_x, anchor_y, anchor_kx,
                    y_out, dy, k_out, k_dx):
        self._call(lib.anuk_dual_update, int(m), int(neq), float(sigma), float(gamma),
                   float(halpern), y.ptr, q.ptr, k_x_bar.ptr, k_x.ptr,
                   anchor_y.ptr if anchor_y is not None else None,
                   anchor_kx.ptr if anchor_kx is not None else None, y_out.ptr, dy.ptr, k_out.ptr,
                   k_dx.ptr)

    def accumulate(self, n, weight, v, total):
        self._call(lib.anuk_accumulate, int(n), float(weight), v.ptr, total.ptr)

    def scale_into(self, n, weight, total, out):
        self._call(lib.anuk_scale_into, int(n), float(weight), total.ptr, out.ptr)

    def weighted_norm_squared(self, dx, dy, omega) -> float:
        out = c_double()
        self._call(lib.anuk_weighted_norm_squared, dx.n, dy.n, dx.ptr, dy.ptr, float(omega),
                   byref(out))
        return out.value

    def step_size_terms(self, dx, dy, k_dx, omega):
        i, mv = c_double(), c_double()
        self._call(lib.anuk_step_size_terms, dx.n, dy.n, dx.ptr, dy.ptr, k_dx.ptr, float(omega),
                   byref(i), byref(mv))
        return i.value, mv.value

    def convergence_terms(self, n, m, c, lower, upper, q, row_scale, col_scale, neq, x, k_x, y,
                          kt_y) -> dict:
        dev = [ctypes.cast(v.ptr, DP) for v in (c, lower, upper, q, row_scale, col_scale)]
        p = _lib.ConvProblem(*dev, int(neq))
        t = _lib.ConvTerms()
        self._call(lib.anuk_convergence_terms, int(n), int(m), byref(p), x.ptr, k_x.ptr, y.ptr,
                   kt_y.ptr, byref(t))
        return {f: getattr(t, f) for f, _ in _lib.ConvTerms._fields_}

    def convergence_terms_noalloc(self, n, m, c, lower, upper, q, row_scale, col_scale, neq, x,
                                  k_x, y, kt_y) -> dict:
        """convergence_terms through the allocation-free entry point (the
        library's device scratch array; calls on one device must not
        overlap)."""
        dev = [ctypes.cast(v.ptr, DP) for v in (c, lower, upper, q, row_scale, col_scale)]
        p = _lib.ConvProblem(*dev, int(neq))
        t = _lib.ConvTerms()
        self._call(lib.anuk_convergence_terms_noalloc, int(n), int(m), byref(p), x.ptr, k_x.ptr,
                   y.ptr, kt_y.ptr, byref(t))
        return {f: getattr(t, f) for f, _ in _lib.ConvTerms._fields_}

    def set_profiling(self, on: bool):
        self._call(lib.anuk_set_profiling, 1 if on else 0)

    def profile_report(self) -> str:
        need = c_size_t()
        self._call(lib.anuk_profile_report, None, 0, byref(need))
        buf = ctypes.create_string_buffer(need.value)
        self._call(lib.anuk_profile_report, buf, need.value, byref(need))
        return buf.value.decode()

    def tuning_report(self) -> str:
        """The batched solver's launch-shape tuning: one line per kernel and
        matrix with the timing of every shape and the one chosen."""
        need = c_size_t(0)
        self._call(lib.anuk_tuning_report, None, 0, byref(need))
        buf = ctypes.create_string_buffer(max(1, need.value))
        self._call(lib.anuk_tuning_report, buf, need.value, None)
        return buf.value.decode()



class LP:
    """A standard-form LP held by the library (StandardLp)."""

    def __init__(self, handle):
        self.h = handle
        info = _lib.LpInfo()
        lib.anuk_lp_info_get(self.h, byref(info))
        self.rows, self.cols, self.nnz = info.rows, info.cols, info.nnz
        self.num_equalities = info.num_equalities
        self.objective_scale, self.objective_offset = info.objective_scale, info.objective_offset

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.anuk_lp_free(h)
            self.h = None

    @staticmethod
    def _made(rc, h) -> "LP":
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        return LP(h)

    @staticmethod
    def from_model(rows, cols, row_index, col_index, value, objective, row_lower, row_upper,
                   col_lower, col_upper, maximize=False, objective_offset=0.0) -> "LP":
        ri, ci, v = _i32(row_index), _i32(col_index), _f64(value)
        arrs = [_f64(a) for a in (objective, row_lower, row_upper, col_lower, col_upper)]
        m = _lib.Model(int(rows), int(cols), int(v.size), _ip(ri), _ip(ci), _dp(v),
                       _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4]),
                       1 if maximize else 0, float(objective_offset))
        h = c_void_p()
        return LP._made(lib.anuk_lp_from_model(byref(m), byref(h)), h)

    @staticmethod
    def from_standard(rows, cols, num_equalities, start, index, value, c, q, lower, upper,
                      objective_scale=1.0, objective_offset=0.0) -> "LP":
        s, i, v = _i32(start), _i32(index), _f64(value)
        c, q, lo, hi = _f64(c), _f64(q), _f64(lower), _f64(upper)
        h = c_void_p()
        rc = lib.anuk_lp_from_standard(int(rows), int(cols), int(num_equalities),
                                       s.ctypes.data_as(IP), _ip(i), _dp(v), _dp(c), _dp(q),
                                       _dp(lo), _dp(hi), float(objective_scale),
                                       float(objective_offset), byref(h))
        return LP._made(rc, h)

    def export(self) -> dict:
        o = {"start": np.empty(self.rows + 1, np.int32), "index": np.empty(self.nnz, np.int32),
             "value": np.empty(self.nnz, np.float64), "c": np.empty(self.cols),
             "q": np.empty(self.rows), "lower": np.empty(self.cols), "upper": np.empty(self.cols),
             "row_model": np.empty(self.rows, np.int32), "row_sign": np.empty(self.rows)}
        rc = lib.anuk_lp_export(self.h, o["start"].ctypes.data_as(IP), _ip(o["index"]),
                                _dp(o["value"]), _dp(o["c"]), _dp(o["q"]), _dp(o["lower"]),
                                _dp(o["upper"]), _ip(o["row_model"]), _dp(o["row_sign"]))
        if rc != 0:
            raise AnukError(rc, "export failed")
        o["num_equalities"] = self.num_equalities
        o["objective_scale"] = self.objective_scale
        o["objective_offset"] = self.objective_offset
        return o

    def scaled(self, ruiz_iterations=10, pock_chambolle=True):
        rs, cs = np.empty(self.rows), np.empty(self.cols)
        h = c_void_p()
        rc = lib.anuk_lp_scale(self.h, int(ruiz_iterations), 1 if pock_chambolle else 0, byref(h),
                               _dp(rs), _dp(cs))
        return LP._made(rc, h), rs, cs

    def scaled_on_device(self, ctx: "Context", ruiz_iterations=10, pock_chambolle=True):
        """scaled() computed on the device (the same bits)."""
        rs, cs = np.empty(self.rows), np.empty(self.cols)
        h = c_void_p()
        ctx._call(lib.anuk_lp_scale_device, self.h, int(ruiz_iterations),
                  1 if pock_chambolle else 0, byref(h), _dp(rs), _dp(cs))
        return LP(h), rs, cs

    def evaluate_residual(self, x, y) -> dict:
        x, y = _f64(x), _f64(y)
        r = _lib.PdhgResidual()
        rc = lib.anuk_evaluate_residual(self.h, _dp(x), _dp(y), byref(r))
        if rc != 0:
            raise AnukError(rc, lib.anuk_last_error(None).decode())
        return {f: getattr(r, f) for f, _ in _lib.PdhgResidual._fields_}


def default_options(**overrides) -> _lib.PdhgOptions:
    o = _lib.PdhgOptions()
    lib.anuk_pdhg_default_options(byref(o))
    for k, v in overrides.items():
        if k == "termination" and isinstance(v, str):
            v = TERMINATION[v]
        if not hasattr(o, k) or k in ("warm_x", "warm_y", "stop"):
            raise KeyError(f"unknown or non-scalar option {k}")
        setattr(o, k, v)
    return o


@dataclass
class SolveResult:
    status: str
    status_code: int
    objective: float
    iterations: int
    restarts: int
    polish_attempts: int
    polish_iterations: int
    polished: bool
    solve_seconds: float
    matrix_norm_estimate: float
    final_step_size: float
    final_primal_weight: float
    phases: dict
    residual: dict
    message: str
    x: np.ndarray = field(repr=False)
    y: np.ndarray = field(repr=False)


def result_from(r: _lib.PdhgResult, x, y) -> SolveResult:
    return SolveResult(
        status=STATUS_NAMES.get(r.status, "unknown"), status_code=r.status, objective=r.objective,
        iterations=r.iterations, restarts=r.restarts, polish_attempts=r.polish_attempts,
        polish_iterations=r.polish_iterations, polished=bool(r.polished),
        solve_seconds=r.solve_seconds, matrix_norm_estimate=r.matrix_norm_estimate,
        final_step_size=r.final_step_size, final_primal_weight=r.final_primal_weight,
        phases={k: getattr(r, "phase_" + k) for k in
                ("scaling", "matrix_norm", "setup", "iterating", "checks", "polishing")},
        residual={f: getattr(r.residual, f) for f, _ in _lib.PdhgResidual._fields_},
        message=r.message.decode(errors="replace"), x=x, y=y)


def solve_pdhg(ctx: Context, lp: LP, options: Optional[_lib.PdhgOptions] = None,
               warm_x=None, warm_y=None, **overrides) -> SolveResult:
    o = options if options is not None else default_options(**overrides)
    if options is not None and overrides:
        for k, v in overrides.items():
            setattr(o, k, TERMINATION[v] if k == "termination" and isinstance(v, str) else v)
    keep = []
    if warm_x is not None and warm_y is not None:
        wx, wy = _f64(warm_x), _f64(warm_y)
        keep += [wx, wy]
        o.warm_x, o.warm_y = _dp(wx), _dp(wy)
    x, y = np.empty(lp.cols), np.empty(lp.rows)
    r = _lib.PdhgResult()
    ctx._call(lib.anuk_solve_pdhg, lp.h, byref(o), byref(r), _dp(x), _dp(y))
    return result_from(r, x, y)


def solve_pdhg_batch(ctx: Context, lps, options: Optional[_lib.PdhgOptions] = None,
                     **overrides) -> list:
    """Solves LPs that share one constraint matrix together. Each lane's row
    sums and steps are the plain solve's operations in the same order; the
    check reductions fold in a different fixed tree, so result i reaches what
    solve_pdhg(ctx, lps[i], ...) reaches (same status and decisions in the
    tested stacks, objective within rounding), not necessarily its bits."""
    lps = list(lps)
    if not lps:
        return []
    o = options if options is not None else default_options(**overrides)
    if options is not None and overrides:
        for k, v in overrides.items():
            setattr(o, k, TERMINATION[v] if k == "termination" and isinstance(v, str) else v)
    count = len(lps)
    handles = (c_void_p * count)(*[lp.h for lp in lps])
    results = (_lib.PdhgResult * count)()
    n, m = lps[0].cols, lps[0].rows
    xs, ys = np.empty((count, n)), np.empty((count, m))
    ctx._call(lib.anuk_solve_pdhg_batch, handles, count, byref(o), results, _dp(xs), _dp(ys))
    return [result_from(results[i], xs[i].copy(), ys[i].copy()) for i in range(count)]


def solve_pdhg_multi(ctx: Context, lps, options: Optional[_lib.PdhgOptions] = None,
                     **overrides) -> list:
    """Solves independent LPs of any size (each with its own K) in one set of
    launches over their block-diagonal stack; result i is what
    solve_pdhg(ctx, lps[i], ...) reaches."""
    lps = list(lps)
    if not lps:
        return []
    o = options if options is not None else default_options(**overrides)
    if options is not None and overrides:
        for k, v in overrides.items():
            setattr(o, k, TERMINATION[v] if k == "termination" and isinstance(v, str) else v)
    count = len(lps)
    handles = (c_void_p * count)(*[lp.h for lp in lps])
    results = (_lib.PdhgResult * count)()
    xs = np.empty(sum(lp.cols for lp in lps))
    ys = np.empty(sum(lp.rows for lp in lps))
    ctx._call(lib.anuk_solve_pdhg_multi, handles, count, byref(o), results, _dp(xs), _dp(ys))
    out, xo, yo = [], 0, 0
    for i, lp in enumerate(lps):
        out.append(result_from(results[i], xs[xo:xo + lp.cols].copy(), ys[yo:yo + lp.rows].copy()))
        xo += lp.cols
        yo += lp.rows
    return out


def estimate_matrix_norm(ctx: Context, lp: LP, iterations=200, tolerance=1e-8) -> float:
    out = c_double()
    ctx._call(lib.anuk_estimate_matrix_norm, lp.h, int(iterations), float(tolerance), byref(out))
    return out.value


# ---- sensitivities -----------------------------------------------------------

ADJOINT_METHODS = {"cg": 0, "fixed_point": 1}
ACTIVE_SETS = {"auto": 0, "point": 1, "basis": 2}


@dataclass
class AdjointResult:
    """Gradients of L = w_x'x + w_y'y with respect to one LP's standard-form
    data (c, q, lower, upper), and how the adjoint solve went."""
    grad_c: np.ndarray
    grad_q: np.ndarray
    grad_lower: np.ndarray
    grad_upper: np.ndarray
    iterations: int
    converged: bool
    residual_primal: float
    residual_dual: float
    free_columns: int
    binding_rows: int
    lower_columns: int
    upper_columns: int
    basis_completed: bool = False

    def model_gradients(self, lp: LP, model_rows: int) -> dict:
        """The same gradients with respect to the natural model's data:
        objective, row_lower, row_upper, col_lower, col_upper. A standard row
        is sign * (model row) with q = sign * (the bound it came from); for an
        equality row both row_lower and row_upper report dL/db for its value b,
        which the two bounds share."""
        e = lp.export()
        lo, up = np.zeros(model_rows), np.zeros(model_rows)
        for r in range(lp.rows):
            i, sign, g = int(e["row_model"][r]), float(e["row_sign"][r]), float(self.grad_q[r])
            if r < lp.num_equalities:
                lo[i] += sign * g
                up[i] += sign * g
            elif sign > 0:
                lo[i] += g
            else:
                up[i] -= g
        return {"objective": lp.objective_scale * self.grad_c, "row_lower": lo, "row_upper": up,
                "col_lower": self.grad_lower.copy(), "col_upper": self.grad_upper.copy()}


def adjoint_options(**overrides) -> _lib.AdjointOptions:
    o = _lib.AdjointOptions()
    lib.anuk_adjoint_default_options(byref(o))
    for k, v in overrides.items():
        if k == "method" and isinstance(v, str):
            v = ADJOINT_METHODS[v]
        if k == "active_set" and isinstance(v, str):
            v = ACTIVE_SETS[v]
        if not hasattr(o, k):
            raise KeyError(f"unknown adjoint option {k}")
        setattr(o, k, v)
    return o


def pdhg_adjoint(ctx: Context, lps, xs, ys, w_x=None, w_y=None,
                 scaling: Optional[_lib.PdhgOptions] = None, **overrides) -> list:
    """Gradients of L = w_x'x + w_y'y at PDHG solutions (xs[i], ys[i]) of LPs
    that share K, one wave64 lane per LP. A single LP may be passed bare."""
    single = isinstance(lps, LP)
    if single:
        lps, xs, ys = [lps], [xs], [ys]
        w_x = None if w_x is None else [w_x]
        w_y = None if w_y is None else [w_y]
    lps = list(lps)
    count = len(lps)
    if count == 0:
        return []
    n, m = lps[0].cols, lps[0].rows
    X = np.ascontiguousarray(np.asarray(xs, dtype=np.float64).reshape(count, n))
    Y = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(count, m))
    WX = None if w_x is None else np.ascontiguousarray(np.asarray(w_x, dtype=np.float64).reshape(count, n))
    WY = None if w_y is None else np.ascontiguousarray(np.asarray(w_y, dtype=np.float64).reshape(count, m))
    gc, gl, gu = np.zeros((count, n)), np.zeros((count, n)), np.zeros((count, n))
    gq = np.zeros((count, m))
    info = (_lib.AdjointInfo * count)()
    sc = scaling if scaling is not None else default_options()
    o = adjoint_options(**overrides)
    handles = (c_void_p * count)(*[lp.h for lp in lps])
    ctx._call(lib.anuk_pdhg_adjoint, handles, count, byref(sc), byref(o), _dp(X), _dp(Y),
              _dp(WX), _dp(WY), _dp(gc), _dp(gq), _dp(gl), _dp(gu), info)
    out = [AdjointResult(gc[i].copy(), gq[i].copy(), gl[i].copy(), gu[i].copy(),
                         info[i].iterations, bool(info[i].converged), info[i].residual_primal,
                         info[i].residual_dual, info[i].free_columns, info[i].binding_rows,
                         info[i].lower_columns, info[i].upper_columns,
                         bool(info[i].basis_completed)) for i in range(count)]
    return out[0] if single else out
