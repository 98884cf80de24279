"""Writes tests/golden/nodes_reference.json from the reference build.

Branch-and-bound frontiers of the refinery plan read as a MILP whose crude
purchases and product sales come in whole units (the `buy` and `sell`
columns integer): the root
relaxation is solved exactly by the reference's simplex, the `depth` most
fractional of them are branched on (floor / ceil children, as
branch_and_bound.cpp:1820-1826 makes them), and every node of the full
frontier of 2^depth nodes is solved the way the reference tree's first-order
node solver solves it: solve_pdhg on the working LP with the node's bounds,
default options (branch_and_bound.cpp:528-566), through oracle/_ref. Each
node's status, iterations, objective, dual objective, dual residual and the
tree's verdict are recorded, with the node LP's exact optimum so a bound can
be checked to be one: the reference's simplex where it finishes (its primal
simplex stops with a numerical error on most of these nodes, which is why
the reference tree retries nodes cold and with the other algorithm), else
HiGHS's dual simplex (scipy's linprog).

    make -C oracle && python tests/golden/make_node_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from anukulan_amd import generators as G  # noqa: E402
from anukulan_amd.nodes import branch_frontier  # noqa: E402

FRONTIERS = [{"periods": 4, "depth": 4}, {"periods": 12, "depth": 6}]


def branch_points(nm, e, x, depth):
    """The `depth` most fractional purchases and sales at x (fractional part
    nearest one half first, then by column), with their values. A column at
    a fractional bound is not branched on: one of its children would
    have lower > upper, and the reference makes no such child
    (branch_and_bound.cpp:1827)."""
    integer = [j for t in nm.tags["buy"] + nm.tags["sell"] for j in t]
    frac = [(abs((x[j] - np.floor(x[j])) - 0.5), j) for j in integer
            if abs(x[j] - np.round(x[j])) > 1e-6
            and e["lower"][j] <= np.floor(x[j]) and np.ceil(x[j]) <= e["upper"][j]]
    frac.sort()
    cols = [j for _, j in frac[:depth]]
    return cols, [float(x[j]) for j in cols]


def highs_optimum(e):
    """min c'x over the standard-form LP (the first num_equalities rows of
    K x = q, the rest K x >= q, lower <= x <= upper), by HiGHS; +inf when
    HiGHS proves the node infeasible (to its 1e-9 feasibility tolerance: a
    branch that overdraws a capacity by less than PDHG's relative tolerance
    of 1e-6 still reads as converged to the first-order method, and any
    bound is a bound on an empty node)."""
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    m, n = len(e["q"]), len(e["c"])
    K = csr_matrix((e["value"], e["index"], e["start"]), shape=(m, n))
    neq = e["num_equalities"]
    r = linprog(e["c"], A_ub=-K[neq:], b_ub=-e["q"][neq:], A_eq=K[:neq], b_eq=e["q"][:neq],
                bounds=list(zip(np.where(np.isinf(e["lower"]), None, e["lower"]),
                                np.where(np.isinf(e["upper"]), None, e["upper"]))),
                method="highs-ds", options={"primal_feasibility_tolerance": 1e-9,
                                            "dual_feasibility_tolerance": 1e-9})
    if r.status == 2:
        return float("inf")
    assert r.status == 0, r.message
    return float(r.fun)


def frontier(spec):
    nm = G.refinery_model(spec["periods"])
    e = nm.to_lp().export()
    root = _oracle.RefLp.from_export(e).simplex()
    assert root["status"] == 0
    cols, vals = branch_points(nm, e, root["x"], spec["depth"])
    lo, hi = branch_frontier(e["lower"], e["upper"], cols, vals)
    nodes = []
    for i in range(lo.shape[0]):
        ne = dict(e, lower=lo[i], upper=hi[i])
        R = _oracle.RefLp.from_export(ne)
        r = R.solve(threads=8)
        v = _oracle.node_verdict(r)
        sx = R.simplex()
        exact, how = None, "reference simplex"
        if sx["status"] == 0:
            exact = float(np.dot(e["c"], sx["x"]))
        else:
            exact, how = highs_optimum(ne), "highs"
            if exact == float("inf"):
                how = "highs: infeasible"
        nodes.append({
            "status": r.status, "iterations": r.iterations, "restarts": r.restarts,
            "polish_attempts": r.polish_attempts, "objective": r.objective.hex(),
            "dual_objective": r.residual["dual_objective"].hex(),
            "dual_residual_inf": r.residual["dual_residual_inf"].hex(),
            "optimal": v["optimal"], "proved_infeasible": v["proved_infeasible"],
            "bound": float(v["bound"]).hex(),
            "simplex_status": int(sx["status"]),
            # standard form, like the bound: c'x at an optimal vertex
            "exact_objective": float(exact).hex(), "exact_from": how})
        print(spec, i, r.status, r.iterations, v["optimal"], sx["status"], flush=True)
    return {"periods": spec["periods"], "depth": spec["depth"], "columns": cols,
            "values": [v.hex() for v in vals], "root_objective": float(root["objective"]).hex(),
            "nodes": nodes}


def main():
    out = {"frontiers": {f"refinery{s['periods']}_depth{s['depth']}": frontier(s) for s in FRONTIERS}}
    path = os.path.join(HERE, "nodes_reference.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
