"""Exact finish (anuk_exact_finish) on the CPU: the crossover of a point to a
starting basis, and the bounded primal simplex that finishes from it.

Pinned by the known optima of the reference's PDHG tests (known_lps) and
its infeasible LP; by the reference simplex's FP64 optima of the 24-period
case stack (tests/golden/pdhg_reference.json stack24), cold and from points
near the optimum; by the reference simplex itself, live, on seeded random
LPs of the shape its crossover tests use (oracle/_ref); and by the
properties those tests hold (test_crossover.cpp:72-170): every crossover
basis is one column per row, all distinct, and starting from it never
changes the optimum. End to end, the reference's refinery.mps presolved by
the reference's presolve, finished, and postsolved gives its published
optimum.
"""
import json
import os

import numpy as np
import pytest

from anukulan_amd import AnukError, generators as G
from anukulan_amd.finish import exact_finish
from golden.known_lps import KNOWN, natural

import _oracle

HERE = os.path.dirname(os.path.abspath(__file__))
STACK = json.load(open(os.path.join(HERE, "golden", "pdhg_reference.json")))["stack24"]


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("name", sorted(KNOWN))
def test_known_answers(name):
    lp = natural(name).to_lp()
    r = exact_finish([lp])[0]
    if name == "infeasible":
        assert r.status == "infeasible"
        return
    assert r.status == "optimal", r.message
    assert abs(r.objective - KNOWN[name]["optimum"]) <= 1e-9 * max(1.0, abs(KNOWN[name]["optimum"]))
    if KNOWN[name]["x"] is not None:
        assert np.allclose(r.x, KNOWN[name]["x"], atol=1e-9)
    assert r.primal_violation <= 1e-9 and r.dual_violation <= 1e-9


def test_unbounded():
    # min -x - y with x - y >= 0, both nonnegative: x = y grows without bound
    nm = G.NaturalModel(rows=1, cols=2, row_index=np.array([0, 0], np.int32),
                        col_index=np.array([0, 1], np.int32), value=np.array([1.0, -1.0]),
                        objective=np.array([-1.0, -1.0]), row_lower=np.array([0.0]),
                        row_upper=np.array([np.inf]), col_lower=np.zeros(2),
                        col_upper=np.array([np.inf, np.inf]))
    assert exact_finish([nm.to_lp()])[0].status == "unbounded"


def stack_lps(count):
    base, cases = G.case_stack(STACK["periods"])
    return [G.case_model(base, c).to_lp() for c in cases[:count]]


def test_stack_cases_reach_the_reference_simplex_optima():
    lps = stack_lps(6)
    for i, r in enumerate(exact_finish(lps)):
        assert r.status == "optimal", (i, r.message)
        assert rel(r.objective, STACK["simplex_objective"][i]) <= 1e-9, (i, r.objective)
        assert r.primal_violation <= 1e-6 and r.dual_violation <= 1e-6, i
        assert not r.started_warm


def test_crossover_from_near_optimal_points():
    # A point a little off the optimum (as a converged first-order point
    # is): the crossover basis installs, the finish lands on the same
    # optimum, and it takes fewer pivots than a cold start.
    lps = stack_lps(3)
    cold = exact_finish(lps)
    g = np.random.default_rng(7)
    xs, ys = [], []
    for lp, c in zip(lps, cold):
        e = lp.export()
        x = c.x + 1e-7 * (1.0 + np.abs(c.x)) * g.standard_normal(lp.cols)
        xs.append(np.clip(x, e["lower"], e["upper"]))
        ys.append(c.y + 1e-7 * (1.0 + np.abs(c.y)) * g.standard_normal(lp.rows))
    warm = exact_finish(lps, xs, ys)
    for i, (w, c) in enumerate(zip(warm, cold)):
        assert w.status == "optimal" and w.started_warm, i
        assert rel(w.objective, STACK["simplex_objective"][i]) <= 1e-9, i
        assert w.iterations < c.iterations, (i, w.iterations, c.iterations)
        b = w.basic
        assert len(b) == lps[i].rows and len(set(b.tolist())) == len(b)
        assert b.min() >= 0 and b.max() < lps[i].rows + lps[i].cols
        assert w.info["pushed"] + w.info["logicals_remaining"] == lps[i].rows


def test_primal_point_alone_and_cold_agree():
    lps = stack_lps(2)
    cold = exact_finish(lps)
    warm = exact_finish(lps, [c.x for c in cold])  # no dual: bounds by proximity
    for w, c in zip(warm, cold):
        assert w.status == "optimal" and rel(w.objective, c.objective) <= 1e-9


def random_lp(seed, rows, cols):
    """min c'x over rows a'x <= b that a witness point satisfies, 0 <= x <= 10
    (the shape of the reference crossover tests' LPs)."""
    g = np.random.default_rng(seed)
    witness = g.uniform(0.0, 5.0, cols)
    ri, ci, v, rhs = [], [], [], []
    for i in range(rows):
        act = 0.0
        for j in range(cols):
            if g.integers(3) == 0:
                continue
            a = float(g.uniform(-3.0, 3.0))
            ri.append(i)
            ci.append(j)
            v.append(a)
            act += a * witness[j]
        rhs.append(act + (0.0 if i % 2 == 0 else 5.0))
    return G.NaturalModel(rows=rows, cols=cols, row_index=np.array(ri, np.int32),
                          col_index=np.array(ci, np.int32), value=np.array(v),
                          objective=g.uniform(-5.0, 5.0, cols), row_lower=np.full(rows, -np.inf),
                          row_upper=np.array(rhs), col_lower=np.zeros(cols),
                          col_upper=np.full(cols, 10.0))


@pytest.mark.skipif(_oracle.ref is None, reason="oracle/_ref not built")
def test_matches_the_reference_simplex_live():
    names = {0: "optimal", 1: "unbounded", 2: "infeasible"}
    lps, want = [], []
    for seed in range(1, 41):
        nm = random_lp(seed, 12 + seed % 9, 16 + seed % 11)
        lps.append(nm.to_lp())
        want.append(_oracle.RefLp.from_natural(nm).simplex())
    for seed, (r, w) in enumerate(zip(exact_finish(lps), want), start=1):
        assert r.status == names.get(w["status"], "other"), seed
        if r.status == "optimal":
            e = lps[seed - 1].export()
            ref_obj = float(e["c"] @ w["x"]) * e["objective_scale"] + e["objective_offset"]
            assert abs(r.objective - ref_obj) <= 1e-9 * max(1.0, abs(ref_obj)), seed


@pytest.mark.skipif(_oracle.ref is None, reason="oracle/_ref not built")
def test_crossover_counts_are_the_references():
    # On the LP as given (no equilibration), the crossover applies the
    # reference's rules to the same point from the same all-logical basis:
    # the same candidates, and the same columns pushed into the basis.
    lps = stack_lps(3)
    cold = exact_finish(lps)
    g = np.random.default_rng(11)
    for lp, c in zip(lps, cold):
        e = lp.export()
        x = np.clip(c.x + 1e-6 * (1.0 + np.abs(c.x)) * g.standard_normal(lp.cols), e["lower"], e["upper"])
        y = c.y + 1e-6 * (1.0 + np.abs(c.y)) * g.standard_normal(lp.rows)
        mine = exact_finish([lp], [x], [y], ruiz_iterations=0)[0]
        ref = _oracle.RefLp.from_export(e).crossover(x, y)
        assert ref["ok"]
        assert mine.info["candidates"] == ref["candidates"]
        assert mine.info["pushed"] == ref["pushed"]
        assert mine.info["rejected_small_pivot"] == ref["rejected_small_pivot"]
        assert mine.status == "optimal" and rel(mine.objective, c.objective) <= 1e-9


def test_refinery_mps_presolved_finished_and_postsolved(ref):
    # The reference's data file end to end on the host: the reference's own
    # presolve (oracle/_ref, test infrastructure), the exact finish of the
    # reduced LP from all logicals, the reference's postsolve. The answer is
    # the file's published optimum 2.426062266887e10
    # (bench/results/ipm_refinery.txt:19), and it is feasible for the model
    # as read.
    import gzip
    from anukulan_amd.mps import read_mps_text
    raw = gzip.decompress(open(os.path.join(HERE, "golden", "refinery.mps.gz"), "rb").read())
    P = ref.RefPresolve(raw)
    assert P.status == "reduced"
    d = P.reduced
    rows_of = np.repeat(np.arange(d["rows"], dtype=np.int32), np.diff(d["start"]))
    reduced = G.NaturalModel(d["rows"], d["cols"], rows_of, d["index"], d["value"], d["objective"],
                             d["row_lower"], d["row_upper"], d["col_lower"], d["col_upper"],
                             d["maximize"], d["objective_offset"])
    r = exact_finish([reduced.to_lp()])[0]
    assert r.status == "optimal", r.message
    x = P.postsolve_primal(r.x)
    m = read_mps_text(raw)
    act = np.zeros(m.rows)
    np.add.at(act, m.row_index(), m.value * x[m.index])
    scale = 1.0 + np.max(np.abs(np.where(np.isfinite(m.row_upper), m.row_upper, 0.0)))
    assert np.max(np.maximum(m.row_lower - act, act - m.row_upper)) <= 1e-12 * scale
    assert np.all(x >= m.col_lower - 1e-9) and np.all(x <= m.col_upper + 1e-9)
    obj = float(m.objective @ x) + m.objective_offset
    assert abs(obj - 2.426062266887e10) <= 1e-12 * 2.426062266887e10
    assert abs(r.objective - obj) <= 1e-12 * abs(obj)


def test_stack_cases_from_the_base_basis_reach_the_optima():
    # The reference's case-stack workflow (app/anukulan_cli.cpp:1065-1230):
    # the base plan's final basis starts each what-if case. Price moves
    # leave that basis primal feasible; supply, demand and capacity moves can
    # make it infeasible, which the composite phase one repairs.
    base, cases = G.case_stack(STACK["periods"])
    picks = [0, 2, 19, 26, 29, 34, 41, 48, 53, 60]   # base; price; supply; demand; capacity
    lps = [G.case_model(base, cases[i]).to_lp() for i in picks]
    f0 = exact_finish(lps[:1])[0]
    assert f0.status == "optimal" and f0.state.shape == (lps[0].rows + lps[0].cols,)
    assert np.array_equal(np.flatnonzero(f0.state == 0), np.sort(f0.basic))
    warm = exact_finish(lps[1:], start=f0)
    cold = exact_finish(lps[1:])
    for k, (w, c) in enumerate(zip(warm, cold)):
        i = picks[k + 1]
        assert w.status == "optimal" and w.started_warm, (i, w.message)
        assert rel(w.objective, STACK["simplex_objective"][i]) <= 1e-9, (i, w.objective)
        assert w.primal_violation <= 1e-6 and w.dual_violation <= 1e-6, i
    assert sum(w.iterations for w in warm) * 3 < sum(c.iterations for c in cold)
    # a (basic, state) pair works as well as a FinishResult
    again = exact_finish(lps[1:2], start=(f0.basic, f0.state))[0]
    assert again.objective == warm[0].objective and again.iterations == warm[0].iterations


def test_a_start_state_the_bounds_cannot_hold_is_moved():
    lp = natural("wyndor").to_lp()
    f = exact_finish([lp])[0]
    state = f.state.copy()
    # every nonbasic column claims its upper bound, which wyndor's columns
    # and logicals do not have
    state[state != 0] = 2
    r = exact_finish([lp], start=(f.basic, state))[0]
    assert r.status == "optimal" and abs(r.objective - 36.0) <= 1e-9
    # an all-logical start is a valid start too
    logical = np.arange(lp.cols, lp.cols + lp.rows, dtype=np.int32)
    st = np.ones(lp.cols + lp.rows, np.int32)
    st[logical] = 0
    r = exact_finish([lp], start=(logical, st))[0]
    assert r.status == "optimal" and abs(r.objective - 36.0) <= 1e-9


def test_bad_start_bases():
    lp = natural("wyndor").to_lp()
    f = exact_finish([lp])[0]
    with pytest.raises(ValueError):
        exact_finish([lp], start=(f.basic[:-1], f.state))
    bad = f.state.copy()
    bad[0] = 7
    with pytest.raises(AnukError):
        exact_finish([lp], start=(f.basic, bad))
    twice = f.basic.copy()
    twice[-1] = twice[0]
    with pytest.raises(AnukError):
        exact_finish([lp], start=(twice, f.state))
    other = natural("equalities").to_lp()
    with pytest.raises(AnukError):
        exact_finish([lp, other], start=f)


def test_bad_arguments():
    lp = natural("wyndor").to_lp()
    with pytest.raises(AnukError):
        exact_finish([lp], primal_tolerance=0.0)
    with pytest.raises(ValueError):
        exact_finish([lp], [np.zeros(3)])
    with pytest.raises(AttributeError):
        exact_finish([lp], no_such_option=1)
    assert exact_finish([]) == []
