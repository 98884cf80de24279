"""Exact finish from device first-order points (solve_pdhg_batch /
solve_pdhg on the GPU, then anuk_exact_finish on the host).

The reference's own claims for this join, restated on our points
(test_crossover.cpp:109-170): starting the simplex from the crossover basis
of a first-order point never changes the optimum, and it saves pivots. On
the 24-period case stack the finished answers are the reference simplex's
FP64 optima (tests/golden/pdhg_reference.json stack24) to 1e-9.
"""
import json
import os

import numpy as np
import pytest

from anukulan_amd import generators as G, solve_pdhg, solve_pdhg_batch
from anukulan_amd.finish import exact_finish
from test_finish import random_lp, rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
STACK = json.load(open(os.path.join(HERE, "golden", "pdhg_reference.json")))["stack24"]


def test_stack_finished_from_batched_points(ctx):
    base, cases = G.case_stack(STACK["periods"])
    lps = [G.case_model(base, c).to_lp() for c in cases[:8]]
    pts = solve_pdhg_batch(ctx, lps, tolerance=1e-8)
    warm = exact_finish(lps, [p.x for p in pts], [p.y for p in pts])
    cold = exact_finish(lps)
    for i, (w, c) in enumerate(zip(warm, cold)):
        assert w.status == c.status == "optimal", i
        assert w.started_warm, i
        assert rel(w.objective, STACK["simplex_objective"][i]) <= 1e-9, (i, w.objective)
        assert rel(w.objective, c.objective) <= 1e-9
    assert sum(w.iterations for w in warm) < 0.7 * sum(c.iterations for c in cold)


def test_starting_from_the_crossover_never_changes_the_optimum(ctx):
    compared = fewer = 0
    for seed in range(1, 31):
        lp = random_lp(seed, 14 + seed % 7, 18 + seed % 9).to_lp()
        cold = exact_finish([lp])[0]
        if cold.status != "optimal":
            continue
        p = solve_pdhg(ctx, lp, tolerance=1e-4, gap_tolerance=1e-4, max_iterations=20000)
        warm = exact_finish([lp], [p.x], [p.y])[0]
        assert warm.status == "optimal", seed
        assert abs(warm.objective - cold.objective) <= 1e-7 * max(1.0, abs(cold.objective)), seed
        compared += 1
        fewer += warm.iterations < cold.iterations
    assert compared >= 20
    print(f"\n{compared} LPs: the same optimum every time, {fewer} with fewer pivots warm")
