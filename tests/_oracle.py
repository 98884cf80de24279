"""Test-side bindings of the checkers under oracle/ (test infrastructure).

liboracle.so       the plain-C restatement of the reference arithmetic
libanukulan_ref.so the reference's own CPU path built from its sources

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
these, and only to check or time against; the product never loads them.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libanukulan_ref.so")
# The reference's CPU path plus the LinAlgBackend adapter over the HIP
# library: RefLp handles made by it solve with threads=-1 on the device.
REF_HIP_PATH = os.path.join(ROOT, "oracle", "_ref", "libanukulan_ref_hip.so")
CONTRACT_PATH = os.path.join(ROOT, "oracle", "_ref", "test_backend_hip")

DP = POINTER(c_double)
IP = POINTER(c_int32)


def _dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(IP) if a is not None else None


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class ConvProblemC(ctypes.Structure):
    _fields_ = [("c", DP), ("lower", DP), ("upper", DP), ("q", DP), ("row_scale", DP),
                ("col_scale", DP), ("num_equalities", c_int32)]


class ConvTermsC(ctypes.Structure):
    _fields_ = [(k, c_double) for k in ("primal_sq", "primal_inf", "q_dot_y", "dual_sq",
                                        "dual_inf", "bound_term", "c_dot_x")]


TERM_NAMES = [k for k, _ in ConvTermsC._fields_]


def _load_oracle():
    if not os.path.exists(ORACLE_PATH):
        return None
    L = ctypes.CDLL(ORACLE_PATH)
    L.orc_from_triplets.restype = c_int64
    L.orc_from_triplets.argtypes = [c_int32, c_int32, c_int64, IP, IP, DP, IP, IP, DP]
    L.orc_transpose.argtypes = [c_int32, c_int32, IP, IP, DP, IP, IP, DP]
    L.orc_spmv.argtypes = [c_int32, IP, IP, DP, DP, DP]
    for f in ("orc_dot",):
        getattr(L, f).restype = c_double
        getattr(L, f).argtypes = [DP, DP, c_int32]
    for f in ("orc_two_norm", "orc_inf_norm"):
        getattr(L, f).restype = c_double
        getattr(L, f).argtypes = [DP, c_int32]
    L.orc_primal_update.argtypes = [c_int32, c_double, c_double, c_double] + [DP] * 9
    L.orc_dual_update.argtypes = [c_int32, c_int32, c_double, c_double, c_double] + [DP] * 10
    L.orc_accumulate.argtypes = [c_int32, c_double, DP, DP]
    L.orc_scale_into.argtypes = [c_int32, c_double, DP, DP]
    L.orc_weighted_norm_squared.restype = c_double
    L.orc_weighted_norm_squared.argtypes = [c_int32, c_int32, DP, DP, c_double]
    L.orc_step_size_terms.argtypes = [c_int32, c_int32, DP, DP, DP, c_double, DP, DP]
    L.orc_convergence_terms.argtypes = [c_int32, c_int32, POINTER(ConvProblemC), DP, DP, DP, DP,
                                        POINTER(ConvTermsC)]
    L.orc_residual_from_terms.argtypes = [POINTER(ConvTermsC), c_double, c_double, c_double,
                                          c_double, DP]
    L.orc_estimate_matrix_norm.restype = c_double
    L.orc_estimate_matrix_norm.argtypes = [c_int32, c_int32, IP, IP, DP, IP, IP, DP, c_int,
                                           c_double]
    L.orc_is_infeasibility_certificate.restype = c_int
    L.orc_is_infeasibility_certificate.argtypes = [c_int32, c_int32, c_int32, IP, IP, DP, DP, DP,
                                                   DP, DP, c_double]
    L.orc_scale_lp.argtypes = [c_int32, c_int32, IP, IP, DP, DP, DP, DP, DP, DP, DP, c_int, c_int]
    L.orc_to_standard_rows.argtypes = [c_int32, IP, IP, DP, DP, DP, IP, IP, POINTER(c_int64), DP,
                                       IP, DP, IP, IP, DP]
    return L


oracle = _load_oracle()


# ---- plain-C restatement, numpy in / numpy out ---------------------------------

def csr_from_triplets(rows, cols, ri, ci, v):
    ri, ci, v = i32(ri), i32(ci), f64(v)
    start = np.zeros(rows + 1, np.int32)
    index = np.zeros(max(v.size, 1), np.int32)
    value = np.zeros(max(v.size, 1))
    nnz = oracle.orc_from_triplets(rows, cols, v.size, _ip(ri), _ip(ci), _dp(v), _ip(start),
                                   _ip(index), _dp(value))
    return start, index[:nnz].copy(), value[:nnz].copy()


def transpose(rows, cols, start, index, value):
    nnz = int(start[-1])
    ts = np.zeros(cols + 1, np.int32)
    ti = np.zeros(max(nnz, 1), np.int32)
    tv = np.zeros(max(nnz, 1))
    oracle.orc_transpose(rows, cols, _ip(i32(start)), _ip(i32(index)), _dp(f64(value)), _ip(ts),
                         _ip(ti), _dp(tv))
    return ts, ti[:nnz].copy(), tv[:nnz].copy()


def spmv(start, index, value, x):
    rows = len(start) - 1
    y = np.zeros(rows)
    oracle.orc_spmv(rows, _ip(i32(start)), _ip(i32(index)), _dp(f64(value)), _dp(f64(x)), _dp(y))
    return y


def dot(a, b):
    a, b = f64(a), f64(b)
    return oracle.orc_dot(_dp(a), _dp(b), a.size)


def two_norm(a):
    a = f64(a)
    return oracle.orc_two_norm(_dp(a), a.size)


def inf_norm(a):
    a = f64(a)
    return oracle.orc_inf_norm(_dp(a), a.size)


def primal_update(tau, gamma, h, x, c, kty, lo, hi, anchor):
    x, c, kty, lo, hi = (f64(a) for a in (x, c, kty, lo, hi))
    n = x.size
    xo, dx, xb = np.zeros(n), np.zeros(n), np.zeros(n)
    oracle.orc_primal_update(n, tau, gamma, h, _dp(x), _dp(c), _dp(kty), _dp(lo), _dp(hi),
                             _dp(f64(anchor)) if anchor is not None else None, _dp(xo), _dp(dx),
                             _dp(xb))
    return xo, dx, xb


def dual_update(neq, sigma, gamma, h, y, q, kxb, kx, ay, ak):
    y, q, kxb, kx = (f64(a) for a in (y, q, kxb, kx))
    m = y.size
    yo, dy, ko, kd = np.zeros(m), np.zeros(m), np.zeros(m), np.zeros(m)
    oracle.orc_dual_update(m, neq, sigma, gamma, h, _dp(y), _dp(q), _dp(kxb), _dp(kx),
                           _dp(f64(ay)) if ay is not None else None,
                           _dp(f64(ak)) if ak is not None else None, _dp(yo), _dp(dy), _dp(ko),
                           _dp(kd))
    return yo, dy, ko, kd


def weighted_norm_squared(dx, dy, omega):
    dx, dy = f64(dx), f64(dy)
    return oracle.orc_weighted_norm_squared(dx.size, dy.size, _dp(dx), _dp(dy), omega)


def step_size_terms(dx, dy, kdx, omega):
    dx, dy, kdx = f64(dx), f64(dy), f64(kdx)
    i, mv = c_double(), c_double()
    oracle.orc_step_size_terms(dx.size, dy.size, _dp(dx), _dp(dy), _dp(kdx), omega, byref(i),
                               byref(mv))
    return i.value, mv.value


def convergence_terms(c, lo, hi, q, rs, cs, neq, x, kx, y, kty):
    arrs = [f64(a) for a in (c, lo, hi, q, rs, cs, x, kx, y, kty)]
    p = ConvProblemC(_dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4]),
                     _dp(arrs[5]), neq)
    t = ConvTermsC()
    oracle.orc_convergence_terms(arrs[0].size, arrs[3].size, byref(p), _dp(arrs[6]),
                                 _dp(arrs[7]), _dp(arrs[8]), _dp(arrs[9]), byref(t))
    return {k: getattr(t, k) for k in TERM_NAMES}


def residual_from_terms(terms: dict, q_two, c_two, q_inf, c_inf):
    t = ConvTermsC(*[terms[k] for k in TERM_NAMES])
    out = np.zeros(12)
    oracle.orc_residual_from_terms(byref(t), q_two, c_two, q_inf, c_inf, _dp(out))
    return out


def estimate_matrix_norm(m, n, ks, ki, kv, ts, ti, tv, iterations=200, tol=1e-8):
    return oracle.orc_estimate_matrix_norm(m, n, _ip(i32(ks)), _ip(i32(ki)), _dp(f64(kv)),
                                           _ip(i32(ts)), _ip(i32(ti)), _dp(f64(tv)), iterations,
                                           tol)


def is_infeasibility_certificate(neq, ts, ti, tv, q, lo, hi, dy, tol=1e-8):
    q, lo, hi, dy = f64(q), f64(lo), f64(hi), f64(dy)
    return bool(oracle.orc_is_infeasibility_certificate(
        q.size, lo.size, neq, _ip(i32(ts)), _ip(i32(ti)), _dp(f64(tv)), _dp(q), _dp(lo), _dp(hi),
        _dp(dy), tol))


def scale_lp(m, n, start, index, value, q, c, lo, hi, ruiz=10, pc=True):
    value, q, c, lo, hi = (f64(a).copy() for a in (value, q, c, lo, hi))
    rs, cs = np.zeros(m), np.zeros(n)
    oracle.orc_scale_lp(m, n, _ip(i32(start)), _ip(i32(index)), _dp(value), _dp(q), _dp(c),
                        _dp(lo), _dp(hi), _dp(rs), _dp(cs), ruiz, 1 if pc else 0)
    return value, q, c, lo, hi, rs, cs


def to_standard_form(rows, cols, ri, ci, v, objective, row_lower, row_upper, col_lower,
                     col_upper, maximize=False):
    """to_standard_form of a natural model, from the restated pieces."""
    a_start, a_index, a_value = csr_from_triplets(rows, cols, ri, ci, v)
    rl, ru = f64(row_lower), f64(row_upper)
    rl = np.where(rl <= -1e30, -np.inf, rl)
    ru = np.where(ru >= 1e30, np.inf, ru)
    srows, neq, count = c_int32(), c_int32(), c_int64()
    oracle.orc_to_standard_rows(rows, _ip(a_start), _ip(a_index), _dp(a_value), _dp(rl), _dp(ru),
                                byref(srows), byref(neq), byref(count), None, None, None, None,
                                None, None)
    S = srows.value
    q, rm, rsg = np.zeros(S), np.zeros(S, np.int32), np.zeros(S)
    tri, tci, tv = (np.zeros(max(count.value, 1), np.int32), np.zeros(max(count.value, 1), np.int32),
                    np.zeros(max(count.value, 1)))
    oracle.orc_to_standard_rows(rows, _ip(a_start), _ip(a_index), _dp(a_value), _dp(rl), _dp(ru),
                                byref(srows), byref(neq), byref(count), _dp(q), _ip(rm), _dp(rsg),
                                _ip(tri), _ip(tci), _dp(tv))
    k = count.value
    start, index, value = csr_from_triplets(S, cols, tri[:k], tci[:k], tv[:k])
    sc = -1.0 if maximize else 1.0
    cl, cu = f64(col_lower), f64(col_upper)
    return {"start": start, "index": index, "value": value, "q": q, "row_model": rm,
            "row_sign": rsg, "num_equalities": neq.value, "c": sc * f64(objective),
            "lower": np.where(cl <= -1e30, -np.inf, cl), "upper": np.where(cu >= 1e30, np.inf, cu),
            "objective_scale": sc}


# ---- the reference build ---------------------------------------------------------

def _load_ref(path=REF_PATH):
    if not os.path.exists(path):
        return None
    from anukulan_amd import _lib as P  # struct layouts only: the shared header's types
    L = ctypes.CDLL(path)
    L.ref_last_error.restype = c_char_p
    L.ref_default_thread_count.restype = c_int
    L.ref_pdhg_default_options.argtypes = [POINTER(P.PdhgOptions)]
    L.ref_lp_from_model.restype = c_void_p
    L.ref_lp_from_model.argtypes = [POINTER(P.Model)]
    L.ref_lp_from_mps.restype = c_void_p
    L.ref_lp_from_mps.argtypes = [c_char_p]
    L.ref_lp_from_standard.restype = c_void_p
    L.ref_lp_from_standard.argtypes = [c_int32, c_int32, c_int32, IP, IP, DP, DP, DP, DP, DP,
                                       c_double, c_double]
    L.ref_lp_free.argtypes = [c_void_p]
    L.ref_lp_info.argtypes = [c_void_p, POINTER(P.LpInfo)]
    L.ref_lp_export.argtypes = [c_void_p, IP, IP, DP, DP, DP, DP, DP, IP, DP]
    L.ref_lp_scale.restype = c_void_p
    L.ref_lp_scale.argtypes = [c_void_p, c_int32, c_int32, DP, DP]
    L.ref_solve_pdhg.argtypes = [c_void_p, POINTER(P.PdhgOptions), c_int32, POINTER(P.PdhgResult),
                                 DP, DP]
    L.ref_estimate_matrix_norm.restype = c_double
    L.ref_estimate_matrix_norm.argtypes = [c_void_p, c_int32, c_double, c_int32]
    L.ref_evaluate_residual.argtypes = [c_void_p, DP, DP, POINTER(P.PdhgResidual)]
    L.ref_simplex.argtypes = [c_void_p, DP, DP, DP, IP, IP, IP]
    L.ref_crossover.argtypes = [c_void_p, DP, DP, IP, IP]
    if hasattr(L, "ref_simplex_warm"):
        L.ref_simplex_warm.argtypes = [c_void_p, c_int32, IP, c_int32, IP, c_int32, DP, DP, DP,
                                       IP, IP, IP, IP, IP, POINTER(c_int64), DP]
    L.ref_op_multiply.argtypes = [c_int32, c_int32, c_int32, IP, IP, DP, DP, DP]
    L.ref_op_transpose.argtypes = [c_int32, c_int32, IP, IP, DP, IP, IP, DP]
    if hasattr(L, "ref_time_multiply"):
        L.ref_time_multiply.restype = c_double
        L.ref_time_multiply.argtypes = [c_int32, c_int32, c_int32, IP, IP, DP, DP, DP, c_int32]
    for f in ("ref_op_dot",):
        getattr(L, f).restype = c_double
        getattr(L, f).argtypes = [DP, DP, c_int32]
    for f in ("ref_op_two_norm", "ref_op_inf_norm"):
        getattr(L, f).restype = c_double
        getattr(L, f).argtypes = [DP, c_int32]
    L.ref_op_primal_update.argtypes = [c_int32, c_double, c_double, c_double] + [DP] * 9
    L.ref_op_dual_update.argtypes = [c_int32, c_int32, c_double, c_double, c_double] + [DP] * 10
    L.ref_op_accumulate.argtypes = [c_int32, c_double, DP, DP]
    L.ref_op_scale_into.argtypes = [c_int32, c_double, DP, DP]
    L.ref_op_weighted_norm_squared.restype = c_double
    L.ref_op_weighted_norm_squared.argtypes = [c_int32, c_int32, DP, DP, c_double]
    L.ref_op_step_size_terms.argtypes = [c_int32, c_int32, DP, DP, DP, c_double, DP, DP]
    L.ref_op_convergence_terms.argtypes = [c_int32, c_int32, POINTER(ConvProblemC), DP, DP, DP,
                                           DP, POINTER(ConvTermsC)]
    L.ref_mps_read_text.restype = c_void_p
    L.ref_mps_read_text.argtypes = [c_char_p, ctypes.c_int64, c_int32, c_int32, c_double]
    L.ref_mps_free.argtypes = [c_void_p]
    L.ref_mps_ok.argtypes = [c_void_p]
    L.ref_mps_error.restype = c_char_p
    L.ref_mps_error.argtypes = [c_void_p]
    L.ref_mps_sizes.argtypes = [c_void_p, POINTER(ctypes.c_int64)]
    L.ref_mps_offset.restype = c_double
    L.ref_mps_offset.argtypes = [c_void_p]
    L.ref_mps_name.restype = c_char_p
    L.ref_mps_name.argtypes = [c_void_p]
    L.ref_mps_arrays.argtypes = [c_void_p, IP, IP, DP, IP, IP, DP, DP, DP, DP, DP, DP, IP]
    for f in ("ref_mps_row_name", "ref_mps_col_name", "ref_mps_warning"):
        getattr(L, f).restype = c_char_p
        getattr(L, f).argtypes = [c_void_p, c_int32]
    L.ref_presolve_mps.restype = c_void_p
    L.ref_presolve_mps.argtypes = [c_char_p, ctypes.c_int64, IP, DP]
    L.ref_presolve_free.argtypes = [c_void_p]
    L.ref_presolve_status.argtypes = [c_void_p]
    L.ref_presolve_message.restype = c_char_p
    L.ref_presolve_message.argtypes = [c_void_p]
    L.ref_presolve_dual_exact.argtypes = [c_void_p]
    L.ref_presolve_reduced.restype = c_void_p
    L.ref_presolve_reduced.argtypes = [c_void_p]
    L.ref_presolve_counts.argtypes = [c_void_p, POINTER(ctypes.c_int64), DP]
    L.ref_postsolve_primal.argtypes = [c_void_p, DP, ctypes.c_int64, DP]
    L.ref_postsolve_dual.argtypes = [c_void_p, DP, ctypes.c_int64, DP, ctypes.c_int64, DP,
                                     ctypes.c_int64, DP]
    L.ref_presolve_lp.restype = c_void_p
    L.ref_presolve_lp.argtypes = [c_void_p]
    return L


ref = _load_ref()


def load_ref_hip():
    """The reference build with the HIP adapter (loads libanukulan_hip.so;
    solving with it needs a GPU), or None when it was not built."""
    return _load_ref(REF_HIP_PATH)


def ref_time_multiply(start, index, value, x, threads=1, reps=10):
    """Seconds per y = A x of the reference backend (threads > 1: its
    threaded backend), the matrix built once before the clock starts."""
    rows = len(start) - 1
    y = np.zeros(rows)
    return ref.ref_time_multiply(threads, rows, len(x), _ip(i32(start)), _ip(i32(index)),
                                 _dp(f64(value)), _dp(f64(x)), _dp(y), reps)


def ref_default_thread_count() -> int:
    """The reference's default_thread_count() (threading.cpp:162)."""
    return int(ref.ref_default_thread_count())


def ref_read_mps(text, format=0, negative_up=0, infinity_threshold=1e30) -> dict:
    """The reference reader's result on `text` (read_mps_stream, source
    "<text>"): {"ok", "error"} and, when ok, every field of the model."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ref.ref_mps_read_text(data, len(data), format, negative_up, infinity_threshold)
    if not h:
        raise RuntimeError(ref.ref_last_error().decode())
    try:
        return _ref_mps_export(h)
    finally:
        ref.ref_mps_free(h)


PRESOLVE_COUNTS = ("rows_removed", "cols_removed", "nonzeros_removed", "empty_rows",
                   "singleton_rows", "redundant_rows", "forcing_rows", "duplicate_rows",
                   "fixed_columns", "empty_columns", "free_column_singletons",
                   "slack_column_singletons", "inequality_column_singletons",
                   "slack_singletons_declined", "doubleton_equations", "dual_fixed_columns",
                   "coefficients_tightened", "bounds_tightened", "bounds_made_finite", "rounds")
PRESOLVE_FLAGS = ("empty_rows", "singleton_rows", "redundant_rows", "forcing_rows",
                  "duplicate_rows", "fixed_columns", "empty_columns", "free_column_singletons",
                  "bound_tightening", "slack_column_singletons", "inequality_column_singletons",
                  "dual_fixing", "doubleton_equations", "coefficient_tightening")


class RefPresolve:
    """The reference's presolve (presolve.cpp:1215-1317) of an MPS text, read
    by the reference's reader; keyword options as PresolveOptions."""

    def __init__(self, text, **opts):
        data = text.encode() if isinstance(text, str) else bytes(text)
        defaults = dict.fromkeys(PRESOLVE_FLAGS, 1)
        defaults["coefficient_tightening"] = 0
        flags = (c_int32 * 15)(*[int(opts.get(f, defaults[f])) for f in PRESOLVE_FLAGS],
                               int(opts.get("max_rounds", 30)))
        tols = (c_double * 4)(opts.get("feasibility_tolerance", 1e-9),
                              opts.get("infeasibility_tolerance", 1e-7),
                              opts.get("bound_relaxation", 0.0),
                              opts.get("max_new_finite_bound", 1e7))
        self.h = ref.ref_presolve_mps(data, len(data), flags, tols)
        if not self.h:
            raise RuntimeError(ref.ref_last_error().decode())
        self.status = {0: "reduced", 1: "infeasible", 2: "unbounded"}[ref.ref_presolve_status(self.h)]
        self.message = ref.ref_presolve_message(self.h).decode()
        self.dual_exact = bool(ref.ref_presolve_dual_exact(self.h))
        counts = (ctypes.c_int64 * 20)()
        largest = c_double()
        ref.ref_presolve_counts(self.h, counts, byref(largest))
        self.counts = dict(zip(PRESOLVE_COUNTS, counts))
        self.counts["largest_new_finite_bound"] = largest.value
        self.reduced = (_ref_mps_export(ref.ref_presolve_reduced(self.h))
                        if self.status == "reduced" else None)

    def __del__(self):
        if getattr(self, "h", None):
            ref.ref_presolve_free(self.h)
            self.h = None

    def postsolve_primal(self, rx):
        rx = f64(rx)
        x = np.zeros(max(1, len(self.reduced["objective"]) + self.counts["cols_removed"]))
        ref.ref_postsolve_primal(self.h, _dp(rx), rx.size, _dp(x))
        return x[:len(self.reduced["objective"]) + self.counts["cols_removed"]]

    def postsolve_dual(self, ry, rc, oc, rows):
        ry, rc, oc = f64(ry), f64(rc), f64(oc)
        y = np.zeros(max(1, rows))
        ref.ref_postsolve_dual(self.h, _dp(ry), ry.size, _dp(rc), rc.size, _dp(oc), oc.size, _dp(y))
        return y[:rows]

    def lp(self):
        return RefLp(ref.ref_presolve_lp(self.h))


def _ref_mps_export(h) -> dict:
    """Every field of the model behind a reference MpsReadResult handle."""
    out = {"ok": bool(ref.ref_mps_ok(h)), "error": ref.ref_mps_error(h).decode()}
    sz = (ctypes.c_int64 * 9)()
    ref.ref_mps_sizes(h, sz)
    rows, cols, nnz, qnnz = sz[0], sz[1], sz[2], sz[3]
    out["warnings"] = [ref.ref_mps_warning(h, k).decode() for k in range(sz[8])]
    if not out["ok"]:
        return out
    a = {"start": np.zeros(rows + 1, np.int32), "index": np.zeros(nnz, np.int32),
         "value": np.zeros(nnz), "objective": np.zeros(cols), "row_lower": np.zeros(rows),
         "row_upper": np.zeros(rows), "col_lower": np.zeros(cols), "col_upper": np.zeros(cols),
         "col_type": np.zeros(cols, np.int32)}
    q = (np.zeros(cols + 1, np.int32), np.zeros(qnnz, np.int32), np.zeros(qnnz)) if qnnz else None

    def p(x, t):
        return x.ctypes.data_as(t) if x is not None and x.size else None
    ref.ref_mps_arrays(h, p(a["start"], IP), p(a["index"], IP), p(a["value"], DP),
                       p(q[0], IP) if q else None, p(q[1], IP) if q else None,
                       p(q[2], DP) if q else None, p(a["objective"], DP),
                       p(a["row_lower"], DP), p(a["row_upper"], DP), p(a["col_lower"], DP),
                       p(a["col_upper"], DP), p(a["col_type"], IP))
    out.update(a)
    out.update(rows=rows, cols=cols, maximize=bool(sz[4]), free_rows_dropped=sz[5],
               objective_nonzeros=sz[6], fixed_format=bool(sz[7]), hessian=q,
               objective_offset=ref.ref_mps_offset(h), name=ref.ref_mps_name(h).decode(),
               row_names=[ref.ref_mps_row_name(h, i).decode() for i in range(rows)],
               col_names=[ref.ref_mps_col_name(h, j).decode() for j in range(cols)])
    return out


class RefLp:
    """A StandardLp held by the reference build (`lib`: the library that
    made the handle, libanukulan_ref.so unless given)."""

    def __init__(self, handle, lib=None):
        self.L = lib if lib is not None else ref
        if not handle:
            raise RuntimeError("reference build: " + self.L.ref_last_error().decode())
        from anukulan_amd import _lib as P
        self.h = handle
        info = P.LpInfo()
        self.L.ref_lp_info(self.h, byref(info))
        self.rows, self.cols, self.nnz = info.rows, info.cols, info.nnz
        self.num_equalities = info.num_equalities
        self.objective_scale, self.objective_offset = info.objective_scale, info.objective_offset

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_lp_free(self.h)
            self.h = None

    @staticmethod
    def from_natural(nm, lib=None):
        from anukulan_amd import _lib as P
        L = lib if lib is not None else ref
        ri, ci, v = i32(nm.row_index), i32(nm.col_index), f64(nm.value)
        arrs = [f64(a) for a in (nm.objective, nm.row_lower, nm.row_upper, nm.col_lower,
                                 nm.col_upper)]
        m = P.Model(nm.rows, nm.cols, v.size, _ip(ri), _ip(ci), _dp(v), *[_dp(a) for a in arrs],
                    1 if nm.maximize else 0, nm.objective_offset)
        return RefLp(L.ref_lp_from_model(byref(m)), L)

    @staticmethod
    def from_mps(text: str):
        return RefLp(ref.ref_lp_from_mps(text.encode()))

    @staticmethod
    def from_export(e: dict):
        rows, cols = len(e["q"]), len(e["c"])
        return RefLp(ref.ref_lp_from_standard(
            rows, cols, e["num_equalities"], _ip(i32(e["start"])), _ip(i32(e["index"])),
            _dp(f64(e["value"])), _dp(f64(e["c"])), _dp(f64(e["q"])), _dp(f64(e["lower"])),
            _dp(f64(e["upper"])), e.get("objective_scale", 1.0), e.get("objective_offset", 0.0)))

    def export(self):
        o = {"start": np.zeros(self.rows + 1, np.int32), "index": np.zeros(max(self.nnz, 1), np.int32),
             "value": np.zeros(max(self.nnz, 1)), "c": np.zeros(self.cols), "q": np.zeros(self.rows),
             "lower": np.zeros(self.cols), "upper": np.zeros(self.cols),
             "row_model": np.zeros(self.rows, np.int32), "row_sign": np.zeros(self.rows)}
        self.L.ref_lp_export(self.h, _ip(o["start"]), _ip(o["index"]), _dp(o["value"]), _dp(o["c"]),
                          _dp(o["q"]), _dp(o["lower"]), _dp(o["upper"]), _ip(o["row_model"]),
                          _dp(o["row_sign"]))
        o["index"], o["value"] = o["index"][:self.nnz], o["value"][:self.nnz]
        o["num_equalities"] = self.num_equalities
        o["objective_scale"] = self.objective_scale
        o["objective_offset"] = self.objective_offset
        return o

    def scaled(self, ruiz=10, pc=True):
        rs, cs = np.zeros(self.rows), np.zeros(self.cols)
        h = self.L.ref_lp_scale(self.h, ruiz, 1 if pc else 0, _dp(rs), _dp(cs))
        return RefLp(h, self.L), rs, cs

    def solve(self, threads=1, warm_x=None, warm_y=None, **overrides):
        from anukulan_amd import _lib as P
        from anukulan_amd.api import TERMINATION, result_from
        o = P.PdhgOptions()
        ref.ref_pdhg_default_options(byref(o))
        for k, v in overrides.items():
            setattr(o, k, TERMINATION[v] if k == "termination" and isinstance(v, str) else v)
        keep = []
        if warm_x is not None:
            wx, wy = f64(warm_x), f64(warm_y)
            keep += [wx, wy]
            o.warm_x, o.warm_y = _dp(wx), _dp(wy)
        x, y = np.zeros(self.cols), np.zeros(self.rows)
        r = P.PdhgResult()
        rc = self.L.ref_solve_pdhg(self.h, byref(o), threads, byref(r), _dp(x), _dp(y))
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        return result_from(r, x, y)

    def matrix_norm(self, iterations=200, tol=1e-8, use_backend=True):
        return self.L.ref_estimate_matrix_norm(self.h, iterations, tol, 1 if use_backend else 0)

    def evaluate_residual(self, x, y):
        from anukulan_amd import _lib as P
        r = P.PdhgResidual()
        self.L.ref_evaluate_residual(self.h, _dp(f64(x)), _dp(f64(y)), byref(r))
        return {k: getattr(r, k) for k, _ in P.PdhgResidual._fields_}

    def crossover(self, x, y=None):
        """The reference's crossover_basis of the point (x, y): its counts and
        the basis it hands over."""
        info = np.zeros(6, np.int32)
        basic = np.zeros(self.rows, np.int32)
        xv = f64(x)
        yv = f64(y) if y is not None else None
        rc = self.L.ref_crossover(self.h, _dp(xv), _dp(yv), _ip(info), _ip(basic))
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        return {"ok": bool(info[0]), "candidates": int(info[1]), "pushed": int(info[2]),
                "rejected_small_pivot": int(info[3]), "logicals_remaining": int(info[4]),
                "stopped_on_refactorization": bool(info[5]), "basic": basic}

    def simplex_warm(self, dual=True, start=None):
        """The reference simplex (dual by default, falling back to the
        primal as solve_lp does), warm-started from `start` = (basic,
        statuses) of an earlier result when given: what the reference's
        case-stack workflow does for each case (app/anukulan_cli.cpp:1065-1230,
        `family --from-first`)."""
        x, y = np.zeros(self.cols), np.zeros(self.rows)
        obj, st, bc, vc = c_double(), c_int32(), c_int32(), c_int32()
        piv, sec = c_int64(), c_double()
        basic = np.zeros(self.rows + self.cols, np.int32)
        vstat = np.zeros(self.rows + self.cols, np.int32)
        sb = ss = None
        if start is not None:
            sb, ss = i32(start[0]), i32(start[1])
        rc = self.L.ref_simplex_warm(self.h, 1 if dual else 0, _ip(sb), 0 if sb is None else sb.size,
                                  _ip(ss), 0 if ss is None else ss.size, _dp(x), _dp(y),
                                  byref(obj), byref(st), _ip(basic), byref(bc), _ip(vstat),
                                  byref(vc), byref(piv), byref(sec))
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        return {"status": st.value, "objective": obj.value, "x": x, "y": y,
                "basic": basic[:bc.value].copy(), "var_status": vstat[:vc.value].copy(),
                "pivots": piv.value, "seconds": sec.value}

    def simplex(self):
        x, y = np.zeros(self.cols), np.zeros(self.rows)
        obj, st = c_double(), c_int32()
        basic = np.zeros(self.rows + self.cols, np.int32)
        cnt = c_int32()
        self.L.ref_simplex(self.h, _dp(x), _dp(y), byref(obj), byref(st), _ip(basic), byref(cnt))
        return {"status": st.value, "objective": obj.value, "x": x, "y": y,
                "basic": basic[:cnt.value].copy()}


def node_verdict(result) -> dict:
    """What the reference tree reads off a first-order node solve
    (branch_and_bound.cpp:533-564): proved infeasible on a Farkas
    certificate; `optimal`, with the dual objective as the subtree's bound,
    only when the solve converged and its dual is feasible to
    1e-7 * max(1, |dual objective|); otherwise the bound stays -inf."""
    dual_obj = result.residual["dual_objective"]
    enough = 1e-7 * max(1.0, abs(dual_obj))
    optimal = result.status == "optimal" and result.residual["dual_residual_inf"] <= enough
    return {"optimal": bool(optimal), "proved_infeasible": result.status == "primal infeasible",
            "bound": dual_obj if optimal else -float("inf")}
