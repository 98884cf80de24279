"""refinery.mps (the reference's data/refinery/refinery.mps, 1380 x 4368)
read by the library's MPS reader and solved by the plain GPU PDHG, against the
reference's solve_pdhg on the same file (tests/golden/mps_reference.json,
written by make_mps_fixtures.py): same status, objective within 1e-8
relative, iterations within 2%; and against the optimum of the reference's
interior point method (bench/results/ipm_refinery.txt:19).
"""
import gzip
import json
import os

import pytest

from anukulan_amd import solve_pdhg, solve_pdhg_batch
from anukulan_amd.mps import read_mps_text

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIX = json.load(open(os.path.join(GOLD, "mps_reference.json")))["pdhg"]["refinery"]


@pytest.fixture(scope="module")
def refinery_lp():
    raw = gzip.decompress(open(os.path.join(GOLD, "refinery.mps.gz"), "rb").read())
    return read_mps_text(raw, "refinery.mps").to_lp()


@pytest.mark.parametrize("tol", ["1e-4", "1e-6"])
def test_refinery_mps_matches_reference_pdhg(ctx, refinery_lp, tol):
    want = FIX[tol]
    r = solve_pdhg(ctx, refinery_lp, tolerance=float(tol), max_iterations=2_000_000,
                   time_limit_seconds=1e9)
    print(f"\nrefinery.mps at {tol}: {r.iterations} iterations (reference {want['iterations']}), "
          f"objective {r.objective!r} (reference {want['objective']!r})")
    assert r.status == want["status"] == "optimal"
    assert abs(r.objective - want["objective"]) <= 1e-8 * abs(want["objective"])
    assert abs(r.iterations - want["iterations"]) <= 0.02 * want["iterations"]
    # rel KKT 1e-6 puts the objective within ~1e-7 of the true optimum here
    if tol == "1e-6":
        assert abs(r.objective - FIX["optimum"]) <= 1e-6 * abs(FIX["optimum"])


def test_refinery_mps_batched_copies_agree(ctx, refinery_lp):
    rs = solve_pdhg_batch(ctx, [refinery_lp] * 3, tolerance=1e-4, max_iterations=2_000_000,
                          time_limit_seconds=1e9)
    assert all(r.status == "optimal" for r in rs)
    assert rs[0].iterations == rs[1].iterations == rs[2].iterations
    assert abs(rs[0].iterations - FIX["1e-4"]["iterations"]) <= 0.02 * FIX["1e-4"]["iterations"]
