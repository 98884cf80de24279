"""Batched case stacks on the GPU.

The batch has no counterpart in the reference (its case stack is solved one
LP at a time), so it is pinned by invariants: LP i of a batch must reach what
the reference's own solve_pdhg reaches on LP i, and what the plain GPU solve
of LP i reaches. Each lane's row sums and elementwise steps are the
reference's operations in the reference's order (bit-identical); only the
check reductions (a fixed tree instead of a serial loop) and ||K|| differ in
the last bits, and they only steer decisions. The bar: the same status,
the objective within 1e-8 relative, the iteration count within 2%.
"""
import numpy as np
import pytest

from anukulan_amd import AnukError, generators as G, solve_pdhg, solve_pdhg_batch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return abs(a - b) / max(1.0, abs(b))


def stack(periods, count=None):
    base, cases = G.case_stack(periods)
    if count is not None:
        cases = cases[:count]
    return base, cases, [G.case_model(base, c) for c in cases]


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_batch_matches_the_reference(ctx, ref, tol):
    base, cases, models = stack(4, 24)
    lps = [m.to_lp() for m in models]
    got = solve_pdhg_batch(ctx, lps, tolerance=tol)
    same = 0
    for m, g in zip(models, got):
        w = ref.RefLp.from_natural(m).solve(tolerance=tol)
        assert g.status == w.status, m.name
        assert relerr(g.objective, w.objective) <= 1e-8, (m.name, g.objective, w.objective)
        assert abs(g.iterations - w.iterations) <= 0.02 * w.iterations, (m.name, g.iterations,
                                                                         w.iterations)
        assert abs(g.matrix_norm_estimate - w.matrix_norm_estimate) <= 1e-12 * w.matrix_norm_estimate
        same += g.iterations == w.iterations
    print(f"\n{same} of {len(models)} cases took exactly the reference's iteration count")


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_batch_matches_plain_gpu_solves(ctx, tol):
    base, cases, models = stack(4, 12)
    lps = [m.to_lp() for m in models]
    got = solve_pdhg_batch(ctx, lps, tolerance=tol)
    for lp, g in zip(lps, got):
        p = solve_pdhg(ctx, lp, tolerance=tol, small_lp_engine=1)  # the plain fused kernels
        assert g.status == p.status == "optimal"
        assert relerr(g.objective, p.objective) <= 1e-8
        assert abs(g.iterations - p.iterations) <= 0.02 * p.iterations


def test_full_24_period_stack_against_the_fixture(ctx):
    # The 24-period base plan is case 0 of the stack: its batched solve must
    # reproduce the reference's fixture exactly at 1e-4.
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "pdhg_reference.json")))["refinery"]["24"]["1e-4"]
    base, cases, models = stack(24)
    got = solve_pdhg_batch(ctx, [m.to_lp() for m in models], tolerance=1e-4)
    assert all(g.status == "optimal" for g in got)
    assert abs(got[0].iterations - fx["iterations"]) <= 0.02 * fx["iterations"]
    assert relerr(got[0].objective, fx["objective"]) <= 1e-8


def test_full_stack_against_the_simplex_optima(ctx):
    # Every case of the 61-case stack, batched to rel KKT 1e-8, against the
    # FP64 optimum of the reference's simplex on that case
    # (tests/golden/pdhg_reference.json stack24, make_fixtures.py).
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "pdhg_reference.json")))["stack24"]
    base, cases, models = stack(24)
    assert len(models) == fx["cases"] and set(fx["simplex_status"]) == {0}
    got = solve_pdhg_batch(ctx, [m.to_lp() for m in models], tolerance=1e-8,
                           max_iterations=2_000_000, time_limit_seconds=1e9)
    assert all(g.status == "optimal" for g in got)
    errs = [relerr(g.objective, o) for g, o in zip(got, fx["simplex_objective"])]
    print(f"\n61 cases at 1e-8 vs simplex: max rel error {max(errs):.2e}, "
          f"median {sorted(errs)[30]:.2e}")
    assert max(errs) <= 1e-6


def stack_fixture():
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "pdhg_reference.json")))["stack24_cases"]


def check_against_reference(got, want, label, counts=False):
    """Per case: the reference's status and polish outcome, the objective
    within 1e-8 relative, the iteration count within 2%; with `counts`, also
    exactly the reference's restarts, polish attempts and polish iterations
    (the batched solver polishes beside its main periods, and must still take
    each case's polishing decisions as the reference does)."""
    same = 0
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.status == w["status"], (label, i, g.status, w["status"])
        assert g.polished == w["polished"], (label, i)
        if counts:
            assert (g.restarts, g.polish_attempts, g.polish_iterations) == (
                w["restarts"], w["polish_attempts"], w["polish_iterations"]), (label, i)
        assert relerr(g.objective, w["objective"]) <= 1e-8, (label, i, g.objective, w["objective"])
        assert abs(g.iterations - w["iterations"]) <= 0.02 * w["iterations"], (
            label, i, g.iterations, w["iterations"])
        same += g.iterations == w["iterations"]
    print(f"\n{label}: {same} of {len(want)} cases took exactly the reference's iteration count")
    return same


@pytest.mark.parametrize("tol", ["1e-4", "1e-8"])
def test_full_stack_per_case_against_the_reference(ctx, tol):
    # All 61 cases of the 24-period stack, batched, against the reference's
    # solve_pdhg on each case (tests/golden/pdhg_reference.json
    # stack24_cases, make_fixtures.py). At 1e-8 five cases stop at the
    # 200,000-iteration limit in the reference; they must stop there here too.
    fx = stack_fixture()
    base, cases, models = stack(24)
    assert len(models) == fx["cases"] == 61
    got = solve_pdhg_batch(ctx, [m.to_lp() for m in models], tolerance=float(tol))
    check_against_reference(got, fx[tol], f"batched {tol}", counts=True)


def test_full_stack_plain_per_case_against_the_reference(ctx):
    # The plain GPU solver on every case at 1e-4, and at 1e-8 on a sample
    # that holds two of the cases the reference stops at the iteration limit.
    fx = stack_fixture()
    base, cases, models = stack(24)
    got = [solve_pdhg(ctx, m.to_lp(), tolerance=1e-4) for m in models]
    check_against_reference(got, fx["1e-4"], "plain 1e-4")
    sample = [0, 1, 4, 7, 30, 60]
    got = [solve_pdhg(ctx, models[i].to_lp(), tolerance=1e-8) for i in sample]
    check_against_reference(got, [fx["1e-8"][i] for i in sample], "plain 1e-8 sample")


@pytest.mark.parametrize("periods", [4, 24])
def test_every_launch_shape_gives_the_same_bits(ctx, periods):
    # The tuner picks among waves per workgroup (1, 2, 4, 8), consecutive
    # rows per wave (1, 2, 4) and the matrix's layout (its CSR, or the
    # fixed-width copy of each row's first entries: 8 wide for K^T here, 32
    # for K), and once the live lanes fit 32 or 16 slots the one-row CSR
    # kernels pack 2 or 4 rows into a wave; a lane's arithmetic must not
    # depend on the shape, so every shape, packed or not, and the tuned
    # choice give the same iterates to the bit. (Eight cases of the stack
    # retire at different iterations, so the runs pass through spans of 16.)
    base, cases, models = stack(periods, 8)
    lps = [m.to_lp() for m in models]
    want = solve_pdhg_batch(ctx, lps, tolerance=1e-8, batch_waves=4, batch_rows=1,
                            batch_layout=1, batch_pack=1)
    for w, r, s in ((1, 1, 1), (2, 1, 1), (8, 1, 1), (1, 2, 1), (4, 2, 1), (8, 2, 1), (1, 4, 1),
                    (2, 4, 1), (4, 4, 1), (8, 4, 1), (1, 1, 2), (2, 1, 2), (4, 1, 2), (8, 1, 2),
                    (0, 0, 0)):
        got = solve_pdhg_batch(ctx, lps, tolerance=1e-8, batch_waves=w, batch_rows=r,
                               batch_layout=s)
        for g, h in zip(got, want):
            assert g.iterations == h.iterations, (w, r, s)
            assert np.array_equal(g.x, h.x) and np.array_equal(g.y, h.y), (w, r, s)
    report = ctx.tuning_report()
    print("\n" + report)
    assert "batch_primal_step" in report and "batch_dual_step" in report and "chosen=w" in report


def test_fixed_width_layout_with_rows_past_its_width(ctx):
    # A random LP whose K has rows of up to 41 entries and whose K^T has rows
    # of up to 97: both fixed-width copies are 32 wide, and the longer rows
    # continue in the CSR. Six cases (scaled objectives) batched with either
    # layout must give the same iterates to the bit.
    base = G.random_lp(150, 700, seed=7)
    models = [G.NaturalModel(**{**base.__dict__, "objective": base.objective * (1.0 + 0.01 * k)})
              for k in range(6)]
    lps = [m.to_lp() for m in models]
    want = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_layout=1)
    got = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_layout=2)
    assert all(h.status == "optimal" for h in want)
    for g, h in zip(got, want):
        assert g.status == h.status and g.iterations == h.iterations
        assert np.array_equal(g.x, h.x) and np.array_equal(g.y, h.y)


@pytest.mark.parametrize("cases", [6, 20])
def test_binned_long_rows_give_the_same_bits(ctx, cases):
    # A random LP whose rows run from 2 to 25 entries in K and to 35 in K^T:
    # with the span at 16 (6 cases) or 32 (20 cases) both steps run binned,
    # short rows packed and every row longer than 8 a wave of its own,
    # summed in one round of 16 or 32 gathers (K^T's rows past 32 in rounds
    # of eight). The iterates must be those of one row per wave to the bit.
    base = G.random_lp(300, 900, mean_col_nnz=4.0, seed=3)
    models = [G.NaturalModel(**{**base.__dict__, "objective": base.objective * (1.0 + 0.01 * k)})
              for k in range(cases)]
    lps = [m.to_lp() for m in models]
    want = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_pack=1)
    got = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_pack=0)
    assert all(h.status == "optimal" for h in want)
    for g, h in zip(got, want):
        assert g.status == h.status and g.iterations == h.iterations
        assert np.array_equal(g.x, h.x) and np.array_equal(g.y, h.y)


def test_packed_dual_step_gives_the_same_bits(ctx):
    # A random LP whose rows of K hold at most 6 entries (K^T up to 59): with
    # six cases the dual step runs packed, four rows to a wave, and must
    # give the iterates of one row per wave to the bit. (The refinery
    # stacks pack the primal side: rows of K^T hold at most 8 entries.)
    base = G.random_lp(2000, 800, mean_col_nnz=3.0, seed=5)
    models = [G.NaturalModel(**{**base.__dict__, "objective": base.objective * (1.0 + 0.01 * k)})
              for k in range(6)]
    lps = [m.to_lp() for m in models]
    want = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_pack=1)
    got = solve_pdhg_batch(ctx, lps, tolerance=1e-6, batch_waves=2, batch_rows=1, batch_pack=0)
    assert all(h.status == "optimal" for h in want)
    for g, h in zip(got, want):
        assert g.status == h.status and g.iterations == h.iterations
        assert np.array_equal(g.x, h.x) and np.array_equal(g.y, h.y)


def test_more_than_one_wave_of_cases(ctx):
    # 70 LPs run as two groups of lanes; duplicates must give identical answers.
    base, cases, models = stack(4, 7)
    lps = [m.to_lp() for m in models] * 10
    got = solve_pdhg_batch(ctx, lps, tolerance=1e-4)
    for i, g in enumerate(got):
        h = got[i % 7]
        assert g.iterations == h.iterations and np.array_equal(g.x, h.x)


def test_batch_iteration_limit_and_exit_polish(ctx, ref):
    base, cases, models = stack(4, 6)
    lps = [m.to_lp() for m in models]
    got = solve_pdhg_batch(ctx, lps, tolerance=1e-9, max_iterations=200)
    for m, g in zip(models, got):
        w = ref.RefLp.from_natural(m).solve(tolerance=1e-9, max_iterations=200)
        assert g.iterations == w.iterations == 200
        assert (g.status, g.polished, g.polish_attempts) == (w.status, w.polished, w.polish_attempts)
        assert relerr(g.objective, w.objective) <= 1e-8


def test_batch_rejects_options_it_does_not_run(ctx):
    base, cases, models = stack(4, 2)
    lps = [m.to_lp() for m in models]
    with pytest.raises(AnukError) as e:
        solve_pdhg_batch(ctx, lps, constant_step_size=0)
    assert e.value.code == 4
    other = G.refinery_model(5).to_lp()
    with pytest.raises(AnukError) as e:
        solve_pdhg_batch(ctx, [lps[0], other])
    assert e.value.code == 1


def test_batch_of_infeasible_and_feasible(ctx):
    # A Farkas certificate in one lane must not disturb the others.
    from golden.known_lps import natural
    inf = natural("infeasible")
    feas = G.NaturalModel(**{**inf.__dict__, "row_lower": np.array([1.5])})
    got = solve_pdhg_batch(ctx, [inf.to_lp(), feas.to_lp(), inf.to_lp()], tolerance=1e-8)
    assert [g.status for g in got] == ["primal infeasible", "optimal", "primal infeasible"]
    plain = solve_pdhg(ctx, feas.to_lp(), tolerance=1e-8)
    assert got[1].iterations == plain.iterations
    assert abs(got[1].objective - 1.5) <= 1e-6
