"""Exact finish of first-order points (anuk_exact_finish): the crossover of
a PDHG point to a starting basis and a bounded primal simplex from it to an
optimal vertex, on the host cores, one LP per thread.

The reference finishes a first-order seed the same way: crossover_basis
(src/anukulan/crossover.cpp:35-229) and its simplex warm-started from that
basis (app/anukulan_cli.cpp:592-631). The device solves every case of a
stack to moderate accuracy (solve_pdhg_batch); the cases that decide
something are finished here:

    res = solve_pdhg_batch(ctx, lps, tolerance=1e-6)
    fin = exact_finish(lps, [r.x for r in res], [r.y for r in res])
    fin[i].objective, fin[i].x, fin[i].basic

The reference's case-stack workflow (app/anukulan_cli.cpp:1065-1230) starts
every what-if case from the base plan's final basis instead
(anuk_exact_finish_basis); a FinishResult is such a start:

    base = exact_finish(lps[:1], [res[0].x], [res[0].y])[0]
    rest = exact_finish(lps[1:], start=base)
"""
from __future__ import annotations

from ctypes import byref, c_void_p
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import lib
from .api import AnukError, _dp, _ip

STATUS = {0: "optimal", 1: "infeasible", 2: "unbounded", 3: "iteration limit", 4: "numerical error"}
INFO = ("iterations", "phase_one_iterations", "refactorizations", "bland_iterations",
        "candidates", "pushed", "rejected_small_pivot", "logicals_remaining")


def options(**overrides) -> _lib.FinishOptions:
    o = _lib.FinishOptions()
    lib.anuk_finish_default_options(byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"no finish option {k}")
        setattr(o, k, v)
    return o


@dataclass
class FinishResult:
    status: str
    objective: float
    started_warm: bool
    stopped_on_refactorization: bool
    primal_violation: float
    dual_violation: float
    seconds: float
    message: str
    info: dict
    x: np.ndarray = field(repr=False)
    y: np.ndarray = field(repr=False)
    basic: np.ndarray = field(repr=False)
    # each column of [K | -I]: 0 basic, 1 at lower, 2 at upper, 3 free
    state: np.ndarray = field(repr=False, default=None)

    @property
    def iterations(self) -> int:
        return self.info["iterations"]


def exact_finish(lps, xs=None, ys=None, start=None, **overrides) -> list:
    """Finishes each LP of `lps` exactly. xs[i] (cols) and ys[i] (rows) are
    its first-order point; xs=None starts every LP cold from all logicals,
    ys=None places nonbasic columns by proximity alone. start (a
    FinishResult of an LP of the same shape, or a (basic, state) pair)
    starts every LP from that basis instead of its point."""
    lps = list(lps)
    if not lps:
        return []
    count = len(lps)
    cols = [lp.cols for lp in lps]
    rows = [lp.rows for lp in lps]

    def stacked(vs, sizes, what):
        if vs is None:
            return None
        vs = list(vs)
        if len(vs) != count or any(np.asarray(v).shape != (s,) for v, s in zip(vs, sizes)):
            raise ValueError(f"one {what} vector of the LP's size per LP")
        return np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64) for v in vs])
                                    if count else np.zeros(0))
    x = stacked(xs, cols, "primal")
    y = stacked(ys, rows, "dual")
    sb = ss = None
    if start is not None:
        basic, state = (start.basic, start.state) if isinstance(start, FinishResult) else start
        sb = np.ascontiguousarray(basic, np.int32)
        ss = np.ascontiguousarray(state, np.int32)
        if sb.shape != (rows[0],) or ss.shape != (rows[0] + cols[0],):
            raise ValueError("a start basis holds one column per row and one state per column")
    o = options(**overrides)
    handles = (c_void_p * count)(*[lp.h for lp in lps])
    results = (_lib.FinishResult * count)()
    xo, yo = np.empty(sum(cols)), np.empty(sum(rows))
    bo = np.empty(sum(rows), np.int32)
    so = np.empty(sum(rows) + sum(cols), np.int32)
    rc = lib.anuk_exact_finish_basis(handles, count, _dp(x) if x is not None else None,
                                     _dp(y) if y is not None else None,
                                     _ip(sb) if sb is not None else None,
                                     _ip(ss) if ss is not None else None, byref(o), results,
                                     _dp(xo), _dp(yo), _ip(bo), _ip(so))
    if rc != 0:
        raise AnukError(rc, lib.anuk_last_error(None).decode())
    out, co, ro = [], 0, 0
    for i, r in enumerate(results):
        out.append(FinishResult(
            status=STATUS.get(r.status, "unknown"), objective=r.objective,
            started_warm=bool(r.started_warm),
            stopped_on_refactorization=bool(r.stopped_on_refactorization),
            primal_violation=r.primal_violation, dual_violation=r.dual_violation,
            seconds=r.seconds, message=r.message.decode(errors="replace"),
            info={k: int(getattr(r, k)) for k in INFO},
            x=xo[co:co + cols[i]].copy(), y=yo[ro:ro + rows[i]].copy(),
            basic=bo[ro:ro + rows[i]].copy(),
            state=so[co + ro:co + ro + cols[i] + rows[i]].copy()))
        co += cols[i]
        ro += rows[i]
    return out
