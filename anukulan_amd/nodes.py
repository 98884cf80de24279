"""Branch-and-bound node relaxations, solved together on the device
(anuk_solve_node_relaxations).

The reference's tree solves a node's relaxation with solve_pdhg on the
working LP with the node's column bounds, one node at a time, and reads a
bound off it only from the dual objective, and only when the dual is
feasible (src/anukulan/branch_and_bound.cpp:528-566). The nodes of a tree
share the working LP's constraint matrix, so a frontier of nodes is a case
stack and is solved as one batch here, node i getting what its own solve
gives.

    lo, hi = np.tile(std["lower"], (k, 1)), np.tile(std["upper"], (k, 1))
    hi[0, j] = 2.0; lo[1, j] = 3.0        # the two children of x_j = 2.4
    res = solve_node_relaxations(ctx, lp, lo, hi)
    res[0].optimal, res[0].bound, res[0].x
"""
from __future__ import annotations

from ctypes import byref
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import lib
from .api import LP, Context, SolveResult, TERMINATION, _dp, default_options, result_from


@dataclass
class NodeResult:
    """NodeSolution of a first-order node solve (branch_and_bound.cpp:44-57):
    `bound` is a lower bound on the node's subtree in standard form (the dual
    objective) when `optimal`, else -inf."""
    optimal: bool
    proved_infeasible: bool
    bound: float
    relaxation: SolveResult = field(repr=False)

    @property
    def x(self) -> np.ndarray:
        return self.relaxation.x

    @property
    def y(self) -> np.ndarray:
        return self.relaxation.y


def branch_frontier(lower, upper, columns, values):
    """Bounds of the 2^d nodes below a node with bounds (lower, upper) after
    branching on `columns` at the fractional `values` (the down child
    upper min(u_j, floor(v)), the up child lower max(l_j, ceil(v)), as the
    reference's children are made, branch_and_bound.cpp:1820-1826). Node i
    takes the up branch on columns[k] when bit k of i is set. Returns
    (lower, upper), each (2^d, cols). A value whose floor is below the
    column's lower bound or whose ceiling is above its upper bound is an
    error: one child's box would be empty, and the reference makes no such
    child (branch_and_bound.cpp:1827)."""
    lower, upper = np.asarray(lower, np.float64), np.asarray(upper, np.float64)
    d = len(columns)
    if len(values) != d:
        raise ValueError("one value per branched column")
    lo = np.tile(lower, (1 << d, 1))
    hi = np.tile(upper, (1 << d, 1))
    for k, (j, v) in enumerate(zip(columns, values)):
        up = (np.arange(1 << d) >> k) & 1 == 1
        hi[~up, j] = np.minimum(hi[~up, j], np.floor(v))
        lo[up, j] = np.maximum(lo[up, j], np.ceil(v))
        if np.floor(v) < lower[j] or np.ceil(v) > upper[j]:
            raise ValueError(f"branching column {j} at {v} leaves a child with an empty box "
                             "(the reference makes no such child)")
    return lo, hi


def solve_node_relaxations(ctx: Context, lp: LP, lower, upper,
                           options: Optional[_lib.PdhgOptions] = None, **overrides) -> list:
    """Solves the relaxations of len(lower) nodes of `lp`: node i has column
    bounds lower[i], upper[i] (standard-form columns) and everything else of
    `lp`. Every node starts cold, as in the reference."""
    lower = np.ascontiguousarray(np.atleast_2d(np.asarray(lower, dtype=np.float64)))
    upper = np.ascontiguousarray(np.atleast_2d(np.asarray(upper, dtype=np.float64)))
    n, m = lp.cols, lp.rows
    if lower.shape != upper.shape or (lower.size and lower.shape[1] != n):
        raise ValueError(f"node bounds must both be (nodes, {n})")
    count = lower.shape[0] if lower.size or n == 0 else 0
    if count == 0:
        return []
    o = options if options is not None else default_options(**overrides)
    if options is not None and overrides:
        for k, v in overrides.items():
            setattr(o, k, TERMINATION[v] if k == "termination" and isinstance(v, str) else v)
    results = (_lib.NodeResult * count)()
    xs, ys = np.empty((count, n)), np.empty((count, m))
    ctx._call(lib.anuk_solve_node_relaxations, lp.h, count, _dp(lower), _dp(upper), byref(o),
              results, _dp(xs), _dp(ys))
    return [NodeResult(optimal=bool(r.optimal), proved_infeasible=bool(r.proved_infeasible),
                       bound=r.bound, relaxation=result_from(r.relaxation, xs[i].copy(), ys[i].copy()))
            for i, r in enumerate(results)]
