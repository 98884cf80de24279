"""Branch-and-bound node relaxations on the GPU (anuk_solve_node_relaxations).

Held to the reference tree's first-order node solver (solve_pdhg per node,
branch_and_bound.cpp:528-566): on the depth-4 and depth-6 frontiers of
tests/golden/nodes_reference.json (make_node_fixtures.py, the reference's
solve_pdhg on every node), each node's status, verdict and Farkas flag must
be the reference's, its objective and bound within 1e-8 relative and its
iterations within 2%; every bound must be a bound (at most the node LP's
exact optimum, to the dual tolerance the verdict allows). Beyond the
reference: nodes the batch cannot run go one by one through the plain
solver with the same answers, and a best-first tree that solves its open
nodes 64 at a time proves the optimum that enumerating every integer point
finds, as the reference's own tree test demands of its first-order tree
(test_branch_and_bound.cpp:185-205, which does not require a proof: a node
the first-order method stalls on is left unsolved there too).
"""
import itertools
import json
import os

import numpy as np
import pytest

from anukulan_amd import AnukError, LP, generators as G, solve_pdhg
from anukulan_amd.nodes import branch_frontier, solve_node_relaxations

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "nodes_reference.json")))["frontiers"]


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


def frontier(name):
    f = FIX[name]
    lp = G.refinery_model(f["periods"]).to_lp()
    e = lp.export()
    lo, hi = branch_frontier(e["lower"], e["upper"], f["columns"],
                             [float.fromhex(v) for v in f["values"]])
    return lp, lo, hi, f["nodes"]


@pytest.mark.parametrize("name", sorted(FIX))
def test_frontier_matches_the_reference(ctx, name):
    lp, lo, hi, want = frontier(name)
    got = solve_node_relaxations(ctx, lp, lo, hi)
    assert len(got) == len(want)
    exact = 0
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.relaxation.status == w["status"], i
        assert (g.optimal, g.proved_infeasible) == (w["optimal"], w["proved_infeasible"]), i
        it = w["iterations"]
        assert abs(g.relaxation.iterations - it) <= 0.02 * it, (i, g.relaxation.iterations, it)
        assert rel(g.relaxation.objective, float.fromhex(w["objective"])) <= 1e-8, i
        if w["optimal"]:
            assert rel(g.bound, float.fromhex(w["bound"])) <= 1e-8, (i, g.bound)
        else:
            assert g.bound == -np.inf
        exact += g.relaxation.iterations == it
    print(f"\n{exact} of {len(want)} nodes took exactly the reference's iteration count")


@pytest.mark.parametrize("name", sorted(FIX))
def test_every_bound_is_a_bound(ctx, name):
    # Weak duality: the dual objective of a dual-feasible point is at most
    # the node's optimum. The verdict admits a dual residual up to 1e-7 of
    # the objective's size, so the bound may exceed the optimum by that much.
    lp, lo, hi, want = frontier(name)
    for i, g in enumerate(solve_node_relaxations(ctx, lp, lo, hi)):
        opt = float.fromhex(want[i]["exact_objective"])  # +inf: the node is empty
        if g.optimal and np.isfinite(opt):
            assert g.bound <= opt + 1e-6 * max(1.0, abs(opt)), (i, g.bound, opt)
            assert g.bound >= opt - 1e-4 * max(1.0, abs(opt)), (i, g.bound, opt)


def test_nodes_match_plain_solves(ctx):
    lp, lo, hi, _ = frontier("refinery4_depth4")
    got = solve_node_relaxations(ctx, lp, lo[:6], hi[:6])
    e = lp.export()
    for i, g in enumerate(got):
        node = LP.from_standard(lp.rows, lp.cols, e["num_equalities"], e["start"], e["index"],
                                e["value"], e["c"], e["q"], lo[i], hi[i], e["objective_scale"],
                                e["objective_offset"])
        p = solve_pdhg(ctx, node)
        assert g.relaxation.status == p.status
        assert abs(g.relaxation.iterations - p.iterations) <= 0.02 * p.iterations
        assert rel(g.relaxation.objective, p.objective) <= 1e-8


def test_options_the_batch_does_not_run_go_one_by_one(ctx):
    # Adaptive step sizes are not in the fused loop: each node is then the
    # plain solver's, bit for bit.
    lp, lo, hi, _ = frontier("refinery4_depth4")
    e = lp.export()
    opts = dict(constant_step_size=0, adaptive_step_size=1, max_iterations=3000)
    got = solve_node_relaxations(ctx, lp, lo[:3], hi[:3], **opts)
    for i, g in enumerate(got):
        node = LP.from_standard(lp.rows, lp.cols, e["num_equalities"], e["start"], e["index"],
                                e["value"], e["c"], e["q"], lo[i], hi[i], e["objective_scale"],
                                e["objective_offset"])
        p = solve_pdhg(ctx, node, **opts)
        assert g.relaxation.iterations == p.iterations and np.array_equal(g.x, p.x)


def tiny(rows, cols, ri, ci, v, obj, rlo, rhi, clo, chi, maximize=False):
    return G.NaturalModel(rows=rows, cols=cols, row_index=np.array(ri, np.int32),
                          col_index=np.array(ci, np.int32), value=np.array(v, float),
                          objective=np.array(obj, float), row_lower=np.array(rlo, float),
                          row_upper=np.array(rhi, float), col_lower=np.array(clo, float),
                          col_upper=np.array(chi, float), maximize=maximize)


def test_infeasible_node_beside_feasible_ones(ctx):
    # x + y >= 3: infeasible in the box [0,1]^2, feasible in [0,2]^2.
    lp = tiny(1, 2, [0, 0], [0, 1], [1.0, 1.0], [1.0, 2.0], [3.0], [np.inf], [0, 0], [5, 5]).to_lp()
    lo = np.zeros((3, 2))
    hi = np.array([[1.0, 1.0], [2.0, 2.0], [1.0, 1.0]])
    got = solve_node_relaxations(ctx, lp, lo, hi, tolerance=1e-8)
    assert [g.proved_infeasible for g in got] == [True, False, True]
    assert [g.optimal for g in got] == [False, True, False]
    assert got[0].bound == -np.inf and got[2].bound == -np.inf
    assert abs(got[1].bound - 4.0) <= 1e-6  # x = 2, y = 1


def test_node_arguments_are_checked(ctx):
    lp = tiny(1, 2, [0, 0], [0, 1], [1.0, 1.0], [1.0, 2.0], [3.0], [np.inf], [0, 0], [5, 5]).to_lp()
    with pytest.raises(ValueError):
        solve_node_relaxations(ctx, lp, np.zeros((2, 3)), np.ones((2, 3)))
    assert solve_node_relaxations(ctx, lp, np.zeros((0, 2)), np.zeros((0, 2))) == []
    with pytest.raises(AnukError):
        solve_node_relaxations(ctx, lp, np.zeros((1, 2)), np.ones((1, 2)), max_iterations=-5)


def test_frontier_larger_than_one_call_of_the_batch(ctx):
    # 300 nodes go to the batch in chunks of 256 (and groups of 64 lanes):
    # every node's result must equal its own solve's, wherever it lands.
    lp = G.refinery_model(2).to_lp()
    e = lp.export()
    cols = [j for j in range(lp.cols) if np.isfinite(e["upper"][j]) and e["upper"][j] >= 4.0]
    lo = np.tile(e["lower"], (300, 1))
    hi = np.tile(e["upper"], (300, 1))
    for i in range(300):
        j = cols[i % len(cols)]
        hi[i, j] = np.floor(e["upper"][j] / (2.0 + i // len(cols)))
    got = solve_node_relaxations(ctx, lp, lo, hi)
    assert len(got) == 300
    for i in (0, 63, 64, 255, 256, 299):
        node = LP.from_standard(lp.rows, lp.cols, e["num_equalities"], e["start"], e["index"],
                                e["value"], e["c"], e["q"], lo[i], hi[i], e["objective_scale"],
                                e["objective_offset"])
        p = solve_pdhg(ctx, node)
        assert got[i].relaxation.status == p.status, i
        assert abs(got[i].relaxation.iterations - p.iterations) <= 0.02 * p.iterations, i
        assert rel(got[i].relaxation.objective, p.objective) <= 1e-8, i


# ---- a best-first tree over batched relaxations ---------------------------

def enumerate_optimum(e, box):
    """min c'x over every integer point of `box` satisfying the standard
    form's rows (the reference test's exhaustive check)."""
    m, n = len(e["q"]), len(e["c"])
    K = np.zeros((m, n))
    for i in range(m):
        for k in range(e["start"][i], e["start"][i + 1]):
            K[i, e["index"][k]] += e["value"][k]
    neq = e["num_equalities"]
    best = None
    for pt in itertools.product(*[range(int(lo), int(hi) + 1) for lo, hi in box]):
        x = np.array(pt, float)
        a = K @ x
        if np.all(np.abs(a[:neq] - e["q"][:neq]) <= 1e-9) and np.all(a[neq:] >= e["q"][neq:] - 1e-9):
            v = float(e["c"] @ x)
            best = v if best is None else min(best, v)
    return best


def batched_tree(ctx, lp, batch=64):
    """Best-first branch and bound on a pure-integer model whose relaxations
    are solved `batch` open nodes at a time. Prunes on the node verdict's
    bound only. A node with no verdict (neither a bound nor a Farkas
    certificate: the first-order method stalled on it) is left unsolved, as
    the reference tree leaves it, and the answer then carries no proof."""
    e = lp.export()
    n = lp.cols
    open_nodes = [(-np.inf, e["lower"].copy(), e["upper"].copy())]
    best = np.inf
    solved = unsolved = 0
    while open_nodes:
        open_nodes.sort(key=lambda t: t[0])
        take, open_nodes = open_nodes[:batch], open_nodes[batch:]
        res = solve_node_relaxations(ctx, lp, np.array([t[1] for t in take]),
                                     np.array([t[2] for t in take]), tolerance=1e-9,
                                     max_iterations=100000)
        solved += len(take)
        for (_, lo, hi), r in zip(take, res):
            if r.proved_infeasible:
                continue
            if not r.optimal:
                unsolved += 1
                continue
            if r.bound >= best - 1e-6 * max(1.0, abs(best)):
                continue
            x = r.x
            frac = [(abs(x[j] - np.round(x[j])), j) for j in range(n) if abs(x[j] - np.round(x[j])) > 1e-6]
            if not frac:
                best = min(best, float(e["c"] @ np.round(x)))
                continue
            j = max(frac)[1]
            v = x[j]
            down_hi = hi.copy()
            down_hi[j] = min(hi[j], np.floor(v))
            up_lo = lo.copy()
            up_lo[j] = max(lo[j], np.ceil(v))
            if lo[j] <= down_hi[j]:
                open_nodes.append((r.bound, lo, down_hi))
            if up_lo[j] <= hi[j]:
                open_nodes.append((r.bound, up_lo, hi))
    return best, solved, unsolved


MILPS = {
    # max 5x + 4y, 6x + 4y <= 24, x + 2y <= 6 : relaxation (3, 1.5), integer (4, 0) at 20
    "knapsack": tiny(2, 2, [0, 0, 1, 1], [0, 1, 0, 1], [6, 4, 1, 2], [5, 4], [-np.inf] * 2,
                     [24, 6], [0, 0], [6, 6], maximize=True),
    # min 3x + 2y + 4z, x + y + 2z >= 4.5, 2x + z >= 3, x - y = 0 ... fractional rows
    "covering": tiny(3, 3, [0, 0, 0, 1, 1, 2, 2], [0, 1, 2, 0, 2, 0, 1], [1, 1, 2, 2, 1, 1, -1],
                     [3, 2, 4], [4.5, 3, 0], [np.inf, np.inf, 0], [0, 0, 0], [5, 5, 5]),
    # 2x + 2y = 3 has no integer point; the relaxation does
    "no_integer_point": tiny(1, 2, [0, 0], [0, 1], [2, 2], [1, 1], [3], [3], [0, 0], [3, 3]),
}


def seeded_milp(seed):
    g = np.random.default_rng(seed)
    n, m = 4, 3
    ri, ci, v = [], [], []
    for i in range(m):
        for j in g.choice(n, size=3, replace=False):
            ri.append(i)
            ci.append(int(j))
            v.append(float(g.integers(1, 7)))
    x0 = g.integers(0, 4, size=n).astype(float)
    act = np.zeros(m)
    for i, j, a in zip(ri, ci, v):
        act[i] += a * x0[j]
    return tiny(m, n, ri, ci, v, list(-g.integers(1, 9, size=n).astype(float)),
                [-np.inf] * m, list(act + g.integers(0, 3, size=m) + 0.5), [0] * n, [3] * n)


@pytest.mark.parametrize("name", list(MILPS) + [f"seeded{s}" for s in range(4)])
def test_batched_tree_agrees_with_enumeration(ctx, name):
    nm = MILPS[name] if name in MILPS else seeded_milp(int(name[6:]))
    lp = nm.to_lp()
    e = lp.export()
    want = enumerate_optimum(e, list(zip(e["lower"], e["upper"])))
    best, solved, unsolved = batched_tree(ctx, lp)
    print(f"\n{name}: {solved} relaxations, {unsolved} without a verdict")
    if want is None:
        assert best == np.inf
    else:
        assert abs(best - want) <= 1e-6 * max(1.0, abs(want)), (best, want, solved)
