"""Sensitivities on the GPU: the adjoint of the PDHG fixed point.

No counterpart in the reference. Pinned by the exact basis sensitivities
B^-1 of the reference simplex's final basis and by finite differences
through reference simplex re-solves: central ones on the random LP and the
4-period stack, one-sided ones (+h and -h) on the 61-case, 24-period stack
(tests/golden/adjoint_reference.npz, written by
tests/golden/make_adjoint_fixtures.py; the checker itself is checked
against the finite differences in test_sensitivity.py). The GPU side starts
from its own PDHG solution at rel KKT 1e-8. Bar: 1e-5 relative,
|a - b| / max(1, |b|).
"""
import os

import numpy as np
import pytest

from anukulan_amd import LP, generators as G, pdhg_adjoint, solve_pdhg, solve_pdhg_batch

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "adjoint_reference.npz"))
FIELDS = ("c", "q", "lower", "upper")


def grads_of(a):
    return (a.grad_c, a.grad_q, a.grad_lower, a.grad_upper)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


@pytest.fixture(scope="module")
def random_lp(ctx):
    lp = G.random_lp(m=300, n=700, seed=5).to_lp()
    r = solve_pdhg(ctx, lp, tolerance=1e-8)
    assert r.status == "optimal"
    return lp, r


def test_adjoint_equals_basis_sensitivities(ctx, random_lp):
    lp, r = random_lp
    a = pdhg_adjoint(ctx, lp, r.x, r.y, w_x=FIX["random_w_x"], w_y=FIX["random_w_y"])
    errs = {f: rel(g, FIX[f"random_grad_{f}"]) for f, g in zip(FIELDS, grads_of(a))}
    print(f"\nadjoint vs B^-1: {errs}; {a.iterations} CG iterations, residuals "
          f"{a.residual_primal:.2e} / {a.residual_dual:.2e}; |F|={a.free_columns} |A|={a.binding_rows}")
    assert a.converged
    assert a.free_columns == a.binding_rows  # a nondegenerate vertex: B is square
    assert max(errs.values()) <= 1e-5


def test_adjoint_equals_finite_differences(ctx, random_lp):
    lp, r = random_lp
    a = pdhg_adjoint(ctx, lp, r.x, r.y, w_x=FIX["random_w_x"], w_y=FIX["random_w_y"])
    g = grads_of(a)
    got = np.array([g[f][i] for f, i in FIX["random_coords"]])
    fd = FIX["random_fd"]
    err = np.abs(got - fd) / np.maximum(1.0, np.abs(fd))
    print(f"\nadjoint vs central differences at {len(fd)} coordinates: max rel {err.max():.2e}")
    assert err.max() <= 1e-4


def test_fixed_point_iteration_reaches_the_same_point(ctx):
    # The literal transposed map contracts by sqrt(1 - 0.81 (s_min / s_max)^2)
    # per step on the active set (s: singular values of B~), so it is only
    # practical on a well-conditioned basis: the Wyndor LP's.
    from golden.known_lps import natural
    lp = natural("wyndor").to_lp()
    r = solve_pdhg(ctx, lp, tolerance=1e-10)
    rng = np.random.default_rng(3)
    wx, wy = rng.standard_normal(lp.cols), rng.standard_normal(lp.rows)
    cg = pdhg_adjoint(ctx, lp, r.x, r.y, w_x=wx, w_y=wy)
    fp = pdhg_adjoint(ctx, lp, r.x, r.y, w_x=wx, w_y=wy, method="fixed_point", max_iterations=20000)
    errs = [rel(a, b) for a, b in zip(grads_of(fp), grads_of(cg))]
    print(f"\nfixed point vs CG: {errs}; residuals {fp.residual_primal:.2e} / {fp.residual_dual:.2e}")
    assert max(errs) <= 1e-6


def test_lanes_are_independent(ctx):
    base, cases = G.case_stack(4)
    lps = [G.case_model(base, c).to_lp() for c in cases[:5]]
    rs = solve_pdhg_batch(ctx, lps, tolerance=1e-8)
    rng = np.random.default_rng(4)
    ws = [rng.standard_normal(lps[0].cols) for _ in lps]
    batch = pdhg_adjoint(ctx, lps, [r.x for r in rs], [r.y for r in rs], w_x=ws)
    for lp, r, w, b in zip(lps, rs, ws, batch):
        one = pdhg_adjoint(ctx, lp, r.x, r.y, w_x=w)
        for u, v in zip(grads_of(one), grads_of(b)):
            assert np.array_equal(u, v)


def stack_report(ctx, periods, count, key, fd):
    base, cases = G.case_stack(periods)
    models = [G.case_model(base, c) for c in cases[:count]]
    lps = [m.to_lp() for m in models]
    rs = solve_pdhg_batch(ctx, lps, tolerance=1e-8, max_iterations=2_000_000,
                          time_limit_seconds=1e9)
    assert all(r.status == "optimal" for r in rs), [r.status for r in rs]
    ws = [FIX[f"{key}_{i}_w_x"] for i in range(count)]
    adj = pdhg_adjoint(ctx, lps, [r.x for r in rs], [r.y for r in rs], w_x=ws)
    rows = []
    for i, a in enumerate(adj):
        g = grads_of(a)
        coords = FIX[f"{key}_{i}_coords"]
        got = np.array([g[f][j] for f, j in coords])
        basis = FIX[f"{key}_{i}_basis"]
        e_basis = float(np.max(np.abs(got - basis) / np.maximum(1.0, np.abs(basis))))
        agree = float(np.mean(np.abs(got - basis) <= 1e-5 * np.maximum(1.0, np.abs(basis))))
        line = (i, e_basis, agree, a.free_columns, a.binding_rows, a.converged,
                tuple(FIX[f"{key}_{i}_degeneracy"]))
        if fd:
            f = FIX[f"{key}_{i}_fd"]
            ok = np.isfinite(f)
            line += (float(np.max(np.abs(got[ok] - f[ok]) / np.maximum(1.0, np.abs(f[ok])))),)
        rows.append(line)
    print(f"\n{key}: case, max rel err vs B^-1, share within 1e-5, |F|, |A|, converged, "
          f"(basic at bound, nonbasic with zero reduced cost)" + (", max rel err vs FD" if fd else ""))
    for line in rows:
        print("  ", line)
    return rows


# The refinery plans are degenerate at their optimum (hundreds of basic
# columns at a bound, hundreds of nonbasic columns with zero reduced cost).
# Read off the PDHG point, the active-set system is not square there; the
# adjoint completes it to an optimal basis (the exact finish from the point)
# and solves that basis's system on the device (active_set AUTO).
def test_refinery_stack4_report(ctx):
    rows = stack_report(ctx, 4, 6, "stack4", fd=True)
    assert all(r[5] for r in rows)  # every adjoint solve converged


def unique_coordinates(i):
    """Coordinates of case i whose one-sided differences agree (to 1e-6
    relative): the derivative along them is the same from either side."""
    p, m = FIX[f"stack24_{i}_fd_plus"], FIX[f"stack24_{i}_fd_minus"]
    ok = np.isfinite(p) & np.isfinite(m)
    scale = np.maximum(1.0, np.maximum(np.abs(p), np.abs(m)))
    return ok & (np.abs(p - m) <= 1e-6 * scale), 0.5 * (p + m)


def test_refinery_stack24_against_basis_sensitivities(ctx):
    # The 61-case, 24-period stack batched to 1e-8, L = w'x on the purchase
    # and sale columns, 200 coordinates per case
    # (tests/golden/adjoint_reference.npz, make_adjoint_fixtures.py). Per
    # case, on every coordinate whose one-sided differences agree, the adjoint
    # must be within 1e-5 relative of the difference; and where the
    # reference simplex's basis sensitivity B^-1 also agrees with it, within
    # 1e-5 relative of B^-1. The other coordinates are where the derivative
    # is set-valued (a kink: the two sides differ), reported as a share.
    #
    # A basis at a degenerate vertex need not stay feasible in either
    # direction of a coordinate; its sensitivity there is then no one-sided
    # derivative, even where the two sides agree. Such a coordinate is
    # allowed only when the basis the adjoint completed (the exact finish
    # from the same pair, which is deterministic) is shown to be infeasible
    # both ways along it, and at most one in a thousand of them.
    import _sensitivity as S
    from anukulan_amd.finish import exact_finish
    base, cases = G.case_stack(24)
    lps = [G.case_model(base, c).to_lp() for c in cases]
    rs = solve_pdhg_batch(ctx, lps, tolerance=1e-8, max_iterations=2_000_000,
                          time_limit_seconds=1e9)
    assert all(r.status == "optimal" for r in rs)
    ws = [FIX[f"stack24_{i}_w_x"] for i in range(len(lps))]
    adj = pdhg_adjoint(ctx, lps, [r.x for r in rs], [r.y for r in rs], w_x=ws)
    lines, exceptions, unique_total, checked_b = [], [], 0, 0
    worst_fd = worst_b = 0.0
    for i, a in enumerate(adj):
        assert a.converged, i
        g = grads_of(a)
        coords = FIX[f"stack24_{i}_coords"]
        got = np.array([g[f][j] for f, j in coords])
        basis = FIX[f"stack24_{i}_basis"]
        u, fd = unique_coordinates(i)
        e_fd = np.abs(got - fd) / np.maximum(1.0, np.abs(fd))
        v = u & (np.abs(basis - fd) <= 1e-6 * np.maximum(1.0, np.abs(fd)))
        e_b = np.abs(got - basis) / np.maximum(1.0, np.abs(basis))
        off = u & ((e_fd > 1e-5) | (v & (e_b > 1e-5)))
        if off.any():
            fin = exact_finish([lps[i]], [rs[i].x], [rs[i].y])[0]
            e = lps[i].export()
            for k in np.flatnonzero(off):
                f, j = int(coords[k][0]), int(coords[k][1])
                up, down = S.basis_directions(e, fin.basic, fin.x, FIELDS[f], j)
                assert not up and not down, (i, FIELDS[f], j, got[k], fd[k], basis[k])
                exceptions.append((i, FIELDS[f], j, float(got[k]), float(fd[k]), float(basis[k])))
        good = u & ~off
        unique_total += int(u.sum())
        checked_b += int(v.sum())
        worst_fd = max(worst_fd, float(e_fd[good].max()))
        worst_b = max(worst_b, float(e_b[good & v].max()))
        lines.append((i, int(u.sum()), float(e_fd[good].max()), int(v.sum()),
                      float(e_b[good & v].max()), int(off.sum()), a.basis_completed,
                      a.iterations))
    total = 200 * len(adj)
    assert worst_fd <= 1e-5 and worst_b <= 1e-5
    assert len(exceptions) <= unique_total // 1000, exceptions
    print(f"\nstack24: {unique_total} of {total} coordinates have one-sided differences that "
          f"agree (set-valued share {1 - unique_total / total:.3f}); adjoint vs differences max "
          f"rel {worst_fd:.2e}; on the {checked_b} where B^-1 agrees too, adjoint vs B^-1 max "
          f"rel {worst_b:.2e}; {len(exceptions)} coordinates where the completed basis is "
          f"infeasible both ways (case, field, index, adjoint, difference, B^-1): {exceptions}")
    print("  case, unique coords, max rel vs FD, coords with B^-1, max rel vs B^-1, "
          "exceptions, basis-completed, CG iterations")
    for line in lines:
        print("  ", line)


def test_autograd_function(ctx, random_lp):
    torch = pytest.importorskip("torch")
    from anukulan_amd.autograd import pdhg_solution
    lp, _ = random_lp
    e = lp.export()
    c, q = torch.tensor(e["c"], requires_grad=True), torch.tensor(e["q"], requires_grad=True)
    lo, up = torch.tensor(e["lower"], requires_grad=True), torch.tensor(e["upper"], requires_grad=True)
    x, y = pdhg_solution(c, q, lo, up, lp, ctx, 1e-8)
    loss = (torch.from_numpy(FIX["random_w_x"]) * x).sum() + (torch.from_numpy(FIX["random_w_y"]) * y).sum()
    loss.backward()
    for t, f in ((c, "c"), (q, "q"), (lo, "lower"), (up, "upper")):
        assert rel(t.grad.numpy(), FIX[f"random_grad_{f}"]) <= 1e-5, f
