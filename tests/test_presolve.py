"""Presolve of case stacks, pinned by what it must preserve rather than by
how the reference reduces (a reduction of this library's own design):

* the reduced model's exact optimum (the reference simplex, oracle/_ref),
  plus its offset, is the original's optimum;
* the postsolved simplex point of the reduced model is an exact KKT point
  of the original model: primal and dual residuals and the gap at rounding
  level, computed by evaluate_residual (pdhg.cpp:107-174) on the original;
* a structural reduction is taken only when every case of a stack allows it,
  and the reduced cases of a stack share one matrix;
* infeasible and unbounded stacks are named with the case that proves it.

The models: the known-answer LPs of tests/test_pdhg.cpp, the refinery plan
and its case stack, refinery.mps, and random LPs with every reduction
planted in them (a feasible point and finite bounds make them solvable).
"""
import gzip
import os

import numpy as np
import pytest

from anukulan_amd import LP, AnukError, generators as G
from anukulan_amd.generators import NaturalModel, Case
from anukulan_amd.presolve import (model_row_duals, presolve, presolve_stack,
                                   standard_duals)

HERE = os.path.dirname(os.path.abspath(__file__))
INF = np.inf


def model_objective(nm, x):
    return float(np.dot(nm.objective, x)) + nm.objective_offset


def exact(ref, nm):
    """The reference simplex on nm: (x, model-row duals, model objective)."""
    lp = nm.to_lp()
    if lp.cols == 0:
        return np.zeros(0), np.zeros(nm.rows), nm.objective_offset
    s = ref.RefLp.from_natural(nm).simplex()
    assert s["status"] == 0, (nm.name, s["status"])
    return s["x"], model_row_duals(lp, s["y"], nm.rows), model_objective(nm, s["x"])


def kkt_on(nm, x, y_model):
    lp = nm.to_lp()
    return lp.evaluate_residual(x, standard_duals(lp, y_model))


def check_round_trip(ref, nm, ps, k=0, tol=1e-9):
    """Case k: the reduced optimum equals the original's, and the postsolved
    reduced optimum is a KKT point of the original."""
    red = ps.model(k)
    _, _, want = exact(ref, nm)
    rx, ry, got = exact(ref, red)
    assert abs(got - want) <= tol * max(1.0, abs(want)), (nm.name, k, got, want)
    # the reduced point is optimal for the reduced model ...
    if red.cols and red.rows:
        r = kkt_on(red, rx, ry)
        assert r["relative_primal"] <= tol and r["relative_dual"] <= tol, r
    # ... and maps to an optimal point of the original
    x, y = ps.postsolve(k, rx, ry)
    assert abs(model_objective(nm, x) - want) <= tol * max(1.0, abs(want))
    r = kkt_on(nm, x, y)
    assert r["relative_primal"] <= tol, (nm.name, k, r)
    assert r["relative_dual"] <= tol, (nm.name, k, r)
    assert r["relative_gap"] <= tol, (nm.name, k, r)
    return r


def planted_lp(seed, m=24, n=30):
    """A random LP (bounded columns, a known feasible point) with every
    reduction planted: empty rows, singleton rows, fixed and empty columns,
    a redundant row and side, and columns alone in equality rows, some of
    them free."""
    g = np.random.default_rng(seed)
    x0 = g.uniform(0.5, 4.5, n)
    trip = {}
    for i in range(m):
        for j in g.choice(n, size=int(g.integers(2, 6)), replace=False):
            trip[(i, int(j))] = float(g.choice([-1, 1]) * g.uniform(0.5, 3.0))
    rows = list(range(m))
    col_lo, col_hi = np.zeros(n), np.full(n, 5.0)
    cost = g.normal(size=n)
    rl, ru = [], []

    def activity(i):
        return sum(v * x0[j] for (r, j), v in trip.items() if r == i)

    for i in rows:
        a = activity(i)
        kind = i % 4
        rl.append(a if kind in (0, 2) else (-INF if kind == 1 else a - g.uniform(0.1, 2)))
        ru.append(a if kind == 0 else (a + g.uniform(0.1, 2) if kind != 2 else INF))
    nrow, ncol = m, n
    extra_cols = []

    def new_row(lo, hi):
        nonlocal nrow
        rl.append(lo)
        ru.append(hi)
        nrow += 1
        return nrow - 1

    def new_col(c, lo, hi, x):
        nonlocal ncol
        extra_cols.append((c, lo, hi, x))
        ncol += 1
        return ncol - 1

    # empty rows
    new_row(-1.0, 2.0)
    new_row(0.0, 0.0)
    # singleton rows on core columns (bounds around x0)
    for j in g.choice(n, 3, replace=False):
        a = float(g.choice([-2.0, 1.5]))
        i = new_row(a * x0[j] - abs(a) * 0.3, a * x0[j] + abs(a) * 0.3)
        trip[(i, int(j))] = a
    # fixed columns inside core rows
    for i in g.choice(m, 2, replace=False):
        v = float(g.uniform(1, 2))
        j = new_col(float(g.normal()), v, v, v)
        trip[(int(i), j)] = 1.0
        rl[int(i)] = rl[int(i)] + v if np.isfinite(rl[int(i)]) else rl[int(i)]
        ru[int(i)] = ru[int(i)] + v if np.isfinite(ru[int(i)]) else ru[int(i)]
    # empty columns: each sign of cost, and zero
    new_col(1.0, -2.0, 3.0, -2.0)
    new_col(-1.0, -2.0, 3.0, 3.0)
    new_col(0.0, 1.0, 3.0, 1.0)
    # a redundant row: both sides beyond the activity range
    cols = g.choice(n, 3, replace=False)
    i = new_row(-100.0, 100.0)
    for j in cols:
        trip[(i, int(j))] = 1.0
    # a row whose lower side is redundant but upper side is not
    i = new_row(-100.0, float(sum(x0[cols])) + 0.5)
    for j in cols:
        trip[(i, int(j))] = 1.0
    # rows that are nearly redundant and bind at the optimum: activity
    # ranges [0, 10] against a lower side of 0.2 (costs push the activity
    # down) and an upper side of 9.8 (costs push it up)
    for lo, hi, pair in ((0.2, INF, (1.0, 2.0)), (-INF, 9.8, (-1.0, -2.0))):
        i = new_row(lo, hi)
        for c in pair:
            trip[(i, new_col(c, 0.0, 5.0, 0.1))] = 1.0
    # column singletons in equality rows: bounded slack, free, and a
    # negative coefficient
    for c, lo, hi, a in ((float(g.normal()), 0.0, 10.0, 1.0), (float(g.normal()), -INF, INF, 2.0),
                         (float(g.normal()), 0.0, INF, -1.5)):
        cols = g.choice(n, 3, replace=False)
        s0 = 1.0 if np.isfinite(lo) else -0.7
        i = new_row(0.0, 0.0)
        for j in cols:
            trip[(i, int(j))] = 1.0
        j = new_col(c, lo, hi, s0)
        trip[(i, j)] = a
        b = float(sum(x0[cols]) + a * s0)
        rl[i] = ru[i] = b
    keys = sorted(trip)
    return NaturalModel(
        rows=nrow, cols=ncol, row_index=np.array([k[0] for k in keys], np.int32),
        col_index=np.array([k[1] for k in keys], np.int32),
        value=np.array([trip[k] for k in keys]),
        objective=np.concatenate([cost, [e[0] for e in extra_cols]]),
        row_lower=np.array(rl), row_upper=np.array(ru),
        col_lower=np.concatenate([col_lo, [e[1] for e in extra_cols]]),
        col_upper=np.concatenate([col_hi, [e[2] for e in extra_cols]]),
        maximize=bool(seed % 2), objective_offset=float(seed % 3), name=f"planted{seed}")


@pytest.mark.parametrize("name", ["wyndor", "equalities", "boxed_free", "ranged", "degenerate",
                                  "offset", "empty_matrix", "norm_diag"])
def test_known_lps_keep_their_optimum(ref, name):
    from golden.known_lps import natural
    nm = natural(name)
    ps = presolve(nm)
    assert ps.status == "reduced"
    check_round_trip(ref, nm, ps)


def test_known_lps_reduce_as_expected():
    from golden.known_lps import natural
    # wyndor: x <= 4 and 2y <= 12 are bounds
    ps = presolve(natural("wyndor"))
    assert (ps.rows, ps.cols, ps.counts["singleton_rows"]) == (1, 2, 2)
    # offset: x >= 4 is a bound, then x is in no row: everything goes
    ps = presolve(natural("offset"))
    assert (ps.rows, ps.cols) == (0, 0)
    assert ps.model().objective_offset == 11.0
    # empty_matrix: an empty row and two empty columns
    ps = presolve(natural("empty_matrix"))
    assert (ps.rows, ps.cols, ps.counts["empty_rows"], ps.counts["empty_columns"]) == (0, 0, 1, 2)
    assert ps.model().objective_offset == -5.0


@pytest.mark.parametrize("seed", range(12))
def test_planted_reductions_keep_the_optimum_and_kkt(ref, seed):
    nm = planted_lp(seed)
    ps = presolve(nm, slack_column_singletons=1)
    assert ps.status == "reduced"
    c = ps.counts
    assert c["empty_rows"] >= 2 and c["singleton_rows"] >= 3, ps.summary()
    assert c["fixed_columns"] >= 2 and c["empty_columns"] >= 3, ps.summary()
    assert c["slack_column_singletons"] >= 3, ps.summary()
    assert c["redundant_rows"] + c["free_rows"] >= 1 and c["redundant_sides"] >= 1, ps.summary()
    check_round_trip(ref, nm, ps)


@pytest.mark.parametrize("seed", range(4))
def test_each_rule_alone_keeps_the_optimum(ref, seed):
    nm = planted_lp(100 + seed)
    rules = ["empty_rows", "singleton_rows", "redundant_rows", "fixed_columns", "empty_columns",
             "slack_column_singletons"]
    for keep in rules:
        ps = presolve(nm, **{r: int(r == keep) for r in rules})
        assert ps.counts.get(keep, 0) > 0 or keep == "redundant_rows", (keep, ps.summary())
        assert ps.status == "reduced"
        check_round_trip(ref, nm, ps)


def test_no_rules_is_the_identity():
    nm = planted_lp(7)
    ps = presolve(nm, empty_rows=0, singleton_rows=0, redundant_rows=0, fixed_columns=0,
                  empty_columns=0, slack_column_singletons=0)
    assert (ps.rows, ps.cols, ps.nnz) == (nm.rows, nm.cols, nm.value.size)
    red = ps.model()
    a = LP.from_model(nm.rows, nm.cols, nm.row_index, nm.col_index, nm.value, nm.objective,
                      nm.row_lower, nm.row_upper, nm.col_lower, nm.col_upper, nm.maximize,
                      nm.objective_offset).export()
    b = red.to_lp().export()
    for k in ("start", "index", "value", "c", "q", "lower", "upper"):
        assert np.array_equal(a[k], b[k]), k
    x = np.arange(nm.cols, dtype=float)
    y = np.arange(nm.rows, dtype=float)
    xo, yo = ps.postsolve(0, x, y)
    assert np.array_equal(xo, x) and np.array_equal(yo, y)


def refinery_cases(periods, picks):
    base, cases = G.case_stack(periods)
    return base, [cases[i] for i in picks]


def test_refinery_plan_keeps_its_optimum(ref):
    nm = G.refinery_model(4)
    assert presolve(nm).counts["slack_column_singletons"] == 0   # off by default
    ps = presolve(nm, slack_column_singletons=1)
    assert ps.counts["slack_column_singletons"] == 14 * 4, ps.summary()
    check_round_trip(ref, nm, ps)


def test_case_stack_reduces_once_for_every_case(ref):
    base, cases = refinery_cases(4, [0, 1, 5, 25, 31, 37, 49, 55, 60])
    ps = presolve_stack(base, cases, slack_column_singletons=1)
    assert ps.status == "reduced" and ps.cases == len(cases) and ps.cols < base.cols
    lps = [m.to_lp() for m in ps.models()]
    k0 = lps[0].export()
    for lp in lps[1:]:
        e = lp.export()
        for key in ("start", "index", "value"):
            assert np.array_equal(e[key], k0[key])
    for k, case in enumerate(cases):
        check_round_trip(ref, G.case_model(base, case), ps, k)


def test_a_reduction_one_case_forbids_is_not_taken(ref):
    nm = planted_lp(3)
    base_case = Case("a", nm.objective.copy(), nm.row_lower.copy(), nm.row_upper.copy(),
                     nm.col_lower.copy(), nm.col_upper.copy())
    alone = presolve_stack(nm, [base_case], slack_column_singletons=1)
    # the same model twice, once with every fixed column unfixed and every
    # equality a ranged row: no column may be fixed, no slack substitution made
    other = Case("b", nm.objective.copy(), nm.row_lower.copy(), nm.row_upper.copy(),
                 nm.col_lower.copy(), nm.col_upper.copy())
    fixed = nm.col_lower == nm.col_upper
    other.col_upper[fixed] += 1.0
    eq = nm.row_lower == nm.row_upper
    other.row_upper[eq] += 0.5
    both = presolve_stack(nm, [base_case, other], slack_column_singletons=1)
    assert alone.counts["fixed_columns"] >= 2 and both.counts["fixed_columns"] == 0
    assert alone.counts["slack_column_singletons"] >= 3
    assert both.counts["slack_column_singletons"] == 0
    for k, case in enumerate([base_case, other]):
        check_round_trip(ref, G.case_model(nm, case), both, k)


@pytest.mark.parametrize("seed", range(4))
def test_a_stack_whose_cases_move_different_numbers(ref, seed):
    # Every case allows the same reductions, but each moves its own numbers:
    # fixed columns at a different value per case (shifting that case's row
    # bounds and offset), singleton rows with per-case bounds, per-case
    # costs (moved by the column substitutions) and right-hand sides.
    nm = planted_lp(200 + seed)
    g = np.random.default_rng(seed)
    fixed = np.flatnonzero(nm.col_lower == nm.col_upper)
    has_entries = np.bincount(nm.row_index, minlength=nm.rows) > 0
    cases = []
    for k in range(3):
        c = Case(f"c{k}", nm.objective * g.uniform(0.8, 1.2, nm.cols), nm.row_lower.copy(),
                 nm.row_upper.copy(), nm.col_lower.copy(), nm.col_upper.copy())
        shift = g.uniform(-0.3, 0.3, fixed.size)
        c.col_lower[fixed] += shift
        c.col_upper[fixed] += shift
        # widen every finite row side a little, differently per case
        c.row_lower[np.isfinite(c.row_lower)] -= g.uniform(0.0, 0.5, np.isfinite(c.row_lower).sum())
        c.row_upper[np.isfinite(c.row_upper)] += g.uniform(0.0, 0.5, np.isfinite(c.row_upper).sum())
        eq = (nm.row_lower == nm.row_upper) & has_entries   # an empty row keeps 0 = 0
        c.row_lower[eq] = c.row_upper[eq] = nm.row_lower[eq] + g.uniform(-0.2, 0.2, eq.sum())
        cases.append(c)
    ps = presolve_stack(nm, cases, slack_column_singletons=1)
    assert ps.status == "reduced"
    assert ps.counts["fixed_columns"] >= 2 and ps.counts["slack_column_singletons"] >= 1
    solved = 0
    for k, c in enumerate(cases):
        cm = G.case_model(nm, c)
        if ref.RefLp.from_natural(cm).simplex()["status"] != 0:
            continue   # a perturbation may leave a case without an optimum
        check_round_trip(ref, cm, ps, k)
        solved += 1
    assert solved >= 1


def test_infeasible_and_unbounded_name_the_case():
    from golden.known_lps import natural
    nm = natural("wyndor")
    good = Case("good", nm.objective.copy(), nm.row_lower.copy(), nm.row_upper.copy(),
                nm.col_lower.copy(), nm.col_upper.copy())
    # row 0 (x <= 4) with x >= 5: the singleton row's bound crosses the column's
    bad = Case("bad", nm.objective.copy(), nm.row_lower.copy(), nm.row_upper.copy(),
               np.array([5.0, 0.0]), nm.col_upper.copy())
    ps = presolve_stack(nm, [good, good, bad])
    assert ps.status == "infeasible" and ps.failing_case == 2, ps.message
    assert ps.message.startswith("case 2:")
    with pytest.raises(ValueError):
        ps.model(0)
    # an empty row that cannot hold
    em = natural("empty_matrix")
    bad = Case("bad", em.objective.copy(), np.array([1.0]), np.array([2.0]), em.col_lower.copy(),
               em.col_upper.copy())
    assert presolve_stack(em, [bad]).status == "infeasible"
    # a column in no row, cost -1, no upper bound
    unb = Case("unbounded", em.objective.copy(), em.row_lower.copy(), em.row_upper.copy(),
               em.col_lower.copy(), np.array([5.0, INF]))
    ps = presolve_stack(em, [unb])
    assert ps.status == "unbounded" and ps.failing_case == 0
    # a row whose activity cannot reach its lower side
    inf = presolve(natural("infeasible"))
    assert inf.status == "infeasible"


def test_malformed_input_is_rejected():
    from anukulan_amd import _lib
    from anukulan_amd._lib import lib
    from anukulan_amd.api import _dp, _ip
    from ctypes import byref, c_void_p
    h = c_void_p()
    start = np.array([0, 2], np.int32)
    index = np.array([1, 0], np.int32)   # columns not increasing
    value = np.array([1.0, 1.0])
    z = np.zeros(2)
    rc = lib.anuk_presolve_stack(1, 2, _ip(start), _ip(index), _dp(value), 1, _dp(z), _dp(z),
                                 _dp(z), _dp(z), _dp(z), None, None, byref(h))
    assert rc == 1 and not h.value
    index = np.array([0, 1], np.int32)
    value = np.array([1.0, 0.0])         # an explicit zero
    rc = lib.anuk_presolve_stack(1, 2, _ip(start), _ip(index), _dp(value), 1, _dp(z), _dp(z),
                                 _dp(z), _dp(z), _dp(z), None, None, byref(h))
    assert rc == 1 and "zero" in lib.anuk_last_error(None).decode()
    lo, hi = np.array([1.0, 0.0]), np.array([0.0, 0.0])   # an empty bound interval
    value = np.array([1.0, 1.0])
    rc = lib.anuk_presolve_stack(1, 2, _ip(start), _ip(index), _dp(value), 1, _dp(z), _dp(z),
                                 _dp(z), _dp(lo), _dp(hi), None, None, byref(h))
    assert rc == 1
    with pytest.raises(TypeError):
        presolve(planted_lp(0), no_such_rule=1)


def test_refinery_mps_keeps_its_optimum_beside_the_reference_presolve(ref):
    from anukulan_amd.mps import read_mps
    path = os.path.join(HERE, "golden", "refinery.mps.gz")
    text = gzip.open(path, "rt").read()
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".mps", delete=False) as f:
        f.write(text)
    try:
        nm = read_mps(f.name).to_natural()
    finally:
        os.unlink(f.name)
    ps = presolve(nm, slack_column_singletons=1)
    r = check_round_trip(ref, nm, ps, tol=1e-8)
    # the file's published optimum (bench/results/ipm_refinery.txt:19),
    # reached from the reduced model
    _, _, reduced_opt = exact(ref, ps.model())
    assert abs(reduced_opt - 2.426062266887e10) <= 1e-12 * 2.426062266887e10
    theirs = ref.RefPresolve(text)
    assert theirs.status == "reduced"
    # both reductions reach the original optimum; sizes are reported
    _, _, want = exact(ref, nm)
    red = ps.model()
    print(f"\nrefinery.mps {nm.rows}x{nm.cols}: ours {ps.rows}x{ps.cols} ({ps.summary()}); "
          f"reference {len(theirs.reduced['row_lower'])}x{len(theirs.reduced['objective'])}; "
          f"optimum {want:.12e}, residual after postsolve {r}")
    assert ps.cols < nm.cols


def test_every_readable_corpus_lp_keeps_its_verdict_and_optimum(ref):
    """The MPS corpus (hand-written texts covering every section and bound
    type, and seeded random models): every LP the reader accepts, presolved
    with every rule on, keeps the reference simplex's verdict, and an
    optimal one keeps its optimum and maps back to a KKT point."""
    from golden.mps_corpus import all_cases
    from anukulan_amd.mps import read_mps_text
    OPTIMAL, UNBOUNDED, INFEASIBLE = 0, 1, 2
    seen = {"optimal": 0, "infeasible": 0, "unbounded": 0, "reduced": 0}
    for label, text, opts in all_cases():
        kw = {}
        if "negative_up" in opts:
            kw["negative_up_bound"] = "cplex" if opts["negative_up"] else "highs"
        if "format" in opts:
            kw["format"] = {0: "auto", 1: "free", 2: "fixed"}[opts["format"]]
        try:
            m = read_mps_text(text, **kw)
        except AnukError:
            continue
        if m.hessian is not None or m.cols == 0:
            continue
        nm = m.to_natural()
        status = ref.RefLp.from_natural(nm).simplex()["status"]
        if status not in (OPTIMAL, UNBOUNDED, INFEASIBLE):
            continue
        ps = presolve(nm, slack_column_singletons=1)
        if status == OPTIMAL:
            assert ps.status == "reduced", (label, ps.message)
            check_round_trip(ref, nm, ps, tol=1e-8)
            seen["optimal"] += 1
        elif ps.status == "reduced":
            # the reduced model must reach the same verdict
            red = ps.model()
            if red.cols:
                assert ref.RefLp.from_natural(red).simplex()["status"] == status, label
            seen["reduced"] += 1
        else:
            # a verdict without solving: never on an LP with an optimum
            seen[ps.status] += 1
    print("\ncorpus presolve verdicts:", seen)
    assert seen["optimal"] >= 40
