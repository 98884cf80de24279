"""End-to-end PDHG parity on the GPU.

The bar: on the same seeded LP and tolerance, the final
objective within 1e-8 relative of the reference CPU PDHG's and the iteration
count within 2%. The reference numbers come from the committed fixture
tests/golden/pdhg_reference.json (written by the reference's own solve_pdhg
through oracle/_ref); where the reference build is present the same solve is
also run live. The known-answer LPs are those of tests/test_pdhg.cpp.
"""
import json
import os

import numpy as np
import pytest

from anukulan_amd import LP, generators as G, solve_pdhg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "pdhg_reference.json")))


def relerr(a, b):
    return abs(a - b) / max(1.0, abs(b))


def known_lp(name):
    from golden.known_lps import natural
    return natural(name).to_lp()


# small_lp_engine 0 solves these small LPs as batches of one; 1 keeps the
# plain fused kernels. Both must meet the same bar.
ENGINES = pytest.mark.parametrize("engine", [0, 1], ids=["small-lp", "plain-kernels"])


@ENGINES
@pytest.mark.parametrize("name", ["wyndor", "equalities", "boxed_free", "ranged", "degenerate",
                                  "offset", "empty_matrix", "norm_diag"])
def test_known_optimum(ctx, name, engine):
    from golden.known_lps import KNOWN
    d = KNOWN[name]
    r = solve_pdhg(ctx, known_lp(name), tolerance=1e-8, small_lp_engine=engine)
    assert r.status == "optimal", r
    assert abs(r.objective - d["optimum"]) <= 1e-6 * max(1.0, abs(d["optimum"]))
    if d["x"] is not None:
        assert np.allclose(r.x, d["x"], atol=1e-5)
    ref = FIX["known"][name]["1e-8"]
    assert r.status == ref["status"]
    assert relerr(r.objective, ref["objective"]) <= 1e-8
    assert abs(r.iterations - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])


def test_infeasible_lp_is_reported(ctx):
    r = solve_pdhg(ctx, known_lp("infeasible"), tolerance=1e-8)
    assert r.status == "primal infeasible"
    assert r.iterations == FIX["known"]["infeasible"]["1e-8"]["iterations"]


def test_matrix_norm_of_diag_3_4(ctx):
    from anukulan_amd import estimate_matrix_norm
    assert abs(estimate_matrix_norm(ctx, known_lp("norm_diag"), 500, 1e-12) - 4.0) <= 1e-6


@ENGINES
@pytest.mark.parametrize("periods", [4, 12])
@pytest.mark.parametrize("tol", ["1e-4", "1e-8"])
def test_refinery_matches_reference(ctx, periods, tol, engine):
    fx = FIX["refinery"][str(periods)]
    lp = G.refinery_model(periods).to_lp()
    assert (lp.rows, lp.cols, lp.nnz) == (fx["rows"], fx["cols"], fx["nnz"])
    r = solve_pdhg(ctx, lp, tolerance=float(tol), small_lp_engine=engine)
    ref = fx[tol]
    print(f"\nrefinery{periods} tol {tol}: gpu {r.iterations} it obj {r.objective:.12e} | "
          f"ref {ref['iterations']} it obj {ref['objective']:.12e} | simplex "
          f"{fx['simplex_objective']:.12e}")
    assert r.status == ref["status"] == "optimal"
    assert relerr(r.objective, ref["objective"]) <= 1e-8
    assert abs(r.iterations - ref["iterations"]) <= 0.02 * ref["iterations"]
    assert abs(r.matrix_norm_estimate - ref["matrix_norm"]) <= 1e-12 * ref["matrix_norm"]
    # error against the FP64 simplex optimum, stated
    err = relerr(r.objective, fx["simplex_objective"])
    assert err <= (1e-3 if tol == "1e-4" else 1e-7)


@ENGINES
def test_graph_replay_is_bit_identical_to_direct_launches(ctx, engine):
    lp = G.refinery_model(4).to_lp()
    a = solve_pdhg(ctx, lp, tolerance=1e-6, use_graphs=1, small_lp_engine=engine)
    b = solve_pdhg(ctx, lp, tolerance=1e-6, use_graphs=0, small_lp_engine=engine)
    c = solve_pdhg(ctx, lp, tolerance=1e-6, use_graphs=1, small_lp_engine=engine)
    assert a.iterations == b.iterations == c.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.array_equal(a.x, c.x)


def test_iteration_limit_is_respected(ctx):
    r = solve_pdhg(ctx, known_lp("wyndor"), tolerance=1e-14, max_iterations=80)
    assert r.status == "iteration limit"
    assert r.iterations == 80


def test_polishing_respects_the_iteration_limit(ctx):
    r = solve_pdhg(ctx, known_lp("wyndor"), tolerance=1e-14, gap_tolerance=1e-2,
                   max_iterations=80)
    assert r.iterations == 80


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(constant_step_size=0),                       # adaptive step size (general loop)
    dict(restarts=0),
    dict(primal_weight_updates=0),
    dict(ruiz_iterations=0, pock_chambolle=0),
    dict(constant_step_size=0, adaptive_step_size=0, restarts=0, primal_weight_updates=0,
         ruiz_iterations=0, pock_chambolle=0),
    dict(halpern=0),                                  # averaged restarts (general loop)
    dict(halpern_restart=1),
])
def test_every_configuration_still_solves(ctx, cfg):
    r = solve_pdhg(ctx, known_lp("wyndor"), tolerance=1e-6, max_iterations=2000000, **cfg)
    assert r.status == "optimal"
    assert abs(r.objective - 36.0) <= 1e-4 * 36.0


def test_is_deterministic(ctx):
    lp = known_lp("wyndor")
    a = solve_pdhg(ctx, lp, tolerance=1e-8)
    b = solve_pdhg(ctx, lp, tolerance=1e-8)
    assert a.iterations == b.iterations and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x)


def test_a_warm_start_is_used(ctx):
    lp = known_lp("wyndor")
    cold = solve_pdhg(ctx, lp, tolerance=1e-8)
    warm = solve_pdhg(ctx, lp, tolerance=1e-8, warm_x=cold.x, warm_y=cold.y, ruiz_iterations=0,
                      pock_chambolle=0)
    assert warm.status == "optimal"
    assert warm.iterations < cold.iterations


def test_polishing_buys_feasibility_for_gap(ctx):
    r = solve_pdhg(ctx, known_lp("wyndor"), tolerance=1e-8, gap_tolerance=1e-2)
    assert r.status == "optimal"
    assert r.residual["relative_primal"] <= 1e-8 and r.residual["relative_dual"] <= 1e-8
    assert r.residual["relative_gap"] <= 1e-2
    assert abs(r.objective - 36.0) <= 0.5


def test_live_reference_agrees_on_options(ctx, ref):
    # Same LP, same non-default options, reference solve run live.
    nm = G.refinery_model(4)
    rlp = ref.RefLp.from_natural(nm)
    for opts in (dict(tolerance=1e-6), dict(tolerance=1e-6, gap_tolerance=1e-3),
                 dict(tolerance=1e-5, reflection=0.5), dict(tolerance=1e-6, detect_infeasibility=0)):
        want = rlp.solve(**opts)
        got = solve_pdhg(ctx, nm.to_lp(), **opts)
        assert got.status == want.status
        assert relerr(got.objective, want.objective) <= 1e-8, opts
        assert abs(got.iterations - want.iterations) <= 0.02 * want.iterations, opts
