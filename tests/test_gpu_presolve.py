"""refinery.mps presolved by the library, solved by the plain GPU PDHG and
postsolved, against the reference's presolve + solve_pdhg on the same file
(tests/golden/presolve_reference.json, make_presolve_fixtures.py): same
status, objective within 1e-8 relative, iterations within 2%; and the
postsolved point against the original model's optimum
(bench/results/ipm_refinery.txt:19).
"""
import gzip
import json
import os

import numpy as np
import pytest

from anukulan_amd import solve_pdhg
from anukulan_amd.mps import read_mps_text
from anukulan_amd.presolve import presolve_mps_text

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIX = json.load(open(os.path.join(GOLD, "presolve_reference.json")))["pdhg"]["refinery_presolved"]
OPTIMUM = 2.426062266887e10


@pytest.fixture(scope="module")
def refinery():
    raw = gzip.decompress(open(os.path.join(GOLD, "refinery.mps.gz"), "rb").read())
    return read_mps_text(raw), presolve_mps_text(raw)


@pytest.mark.parametrize("tol", ["1e-4", "1e-6"])
def test_presolved_refinery_matches_reference(ctx, refinery, tol):
    model, pre = refinery
    lp = pre.reduced.to_lp()
    assert (lp.rows, lp.cols, lp.nnz) == (FIX["rows"], FIX["cols"], FIX["nnz"])
    want = FIX[tol]
    r = solve_pdhg(ctx, lp, tolerance=float(tol), max_iterations=2_000_000, time_limit_seconds=1e9)
    print(f"\npresolved refinery.mps at {tol}: {r.iterations} iterations "
          f"(reference {want['iterations']}), objective {r.objective!r}")
    assert r.status == want["status"] == "optimal"
    assert abs(r.objective - want["objective"]) <= 1e-8 * abs(want["objective"])
    assert abs(r.iterations - want["iterations"]) <= 0.02 * want["iterations"]
    # back in the original model's columns, with its own objective
    x = pre.postsolve_primal(r.x)
    assert x.shape == (model.cols,)
    obj = float(model.objective @ x) + model.objective_offset
    assert abs(obj - r.objective) <= 1e-9 * abs(r.objective)
    if tol == "1e-6":
        assert abs(obj - OPTIMUM) <= 1e-6 * OPTIMUM
