"""Exact basis sensitivities, the checker for the adjoint (test infrastructure).

Given a standard-form LP (its export) and the final basis of the reference
simplex (SimplexResult::final_basic, simplex.hpp:429; indices into the logical
form [K | -I] of simplex.hpp:14-36: j < n is column j, n + i the slack of
row i), the vertex is x_B = B^-1 (q - K_N x_N) with B = [K | -I][:, basic],
and the duals are y = B^-T c_B. So for L = w_x'x + w_y'y:

    dL/dq       = B^-T w_B                  (w_B: w_x on basic columns, 0 on slacks)
    dL/dbound_j = w_j - K[:, j]' B^-T w_B   (nonbasic column j, at that bound)
    dL/dc_B     = B^-1 w_y                  (basic columns; 0 elsewhere)

B is factorised with scipy's sparse LU (splu). Central finite differences
through re-solves with the reference simplex check this module itself.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu


def basis_matrix(e, basic):
    m, n = len(e["start"]) - 1, len(e["c"])
    k = sp.csr_matrix((e["value"], e["index"], e["start"]), shape=(m, n)).tocsc()
    cols = []
    for j in basic:
        if j < n:
            cols.append(k[:, int(j)])
        else:
            cols.append(sp.csc_matrix(([-1.0], ([int(j) - n], [0])), shape=(m, 1)))
    return sp.hstack(cols).tocsc(), k


def basis_sensitivity(e, basic, x, w_x=None, w_y=None, tol=1e-9):
    """Gradients (grad_c, grad_q, grad_lower, grad_upper) of w_x'x + w_y'y at
    the basis `basic` of the standard-form LP `e`, x the vertex."""
    m, n = len(e["start"]) - 1, len(e["c"])
    basic = np.asarray(basic, dtype=np.int64)
    B, k = basis_matrix(e, basic)
    lu = splu(B)
    gq, gc = np.zeros(m), np.zeros(n)
    gl, gu = np.zeros(n), np.zeros(n)
    is_basic = np.zeros(n + m, bool)
    is_basic[basic] = True
    if w_x is not None:
        w_b = np.array([w_x[j] if j < n else 0.0 for j in basic])
        mu = lu.solve(w_b, trans="T")
        gq = mu
        t = k.T @ mu
        lo, up = e["lower"], e["upper"]
        for j in range(n):
            if is_basic[j]:
                continue
            if np.isfinite(lo[j]) and abs(x[j] - lo[j]) <= tol * (1 + abs(lo[j])):
                gl[j] = w_x[j] - t[j]
            elif np.isfinite(up[j]) and abs(x[j] - up[j]) <= tol * (1 + abs(up[j])):
                gu[j] = w_x[j] - t[j]
    if w_y is not None:
        nu = lu.solve(np.asarray(w_y, dtype=np.float64))
        for pos, j in enumerate(basic):
            if j < n:
                gc[j] = nu[pos]
    return gc, gq, gl, gu


def perturbed(e, field, index, delta):
    f = dict(e)
    f[field] = np.array(e[field], dtype=np.float64, copy=True)
    f[field][index] += delta
    return f


def central_difference(ref, e, field, index, loss, rel_step=1e-6):
    """(L(theta + h) - L(theta - h)) / 2h through two reference simplex
    solves, h = rel_step * max(1, |theta|)."""
    h = rel_step * max(1.0, abs(float(e[field][index])))
    vals = []
    for s in (1.0, -1.0):
        sx = ref.RefLp.from_export(perturbed(e, field, index, s * h)).simplex()
        if sx["status"] != 0:
            return None
        vals.append(loss(sx["x"], sx["y"]))
    return (vals[0] - vals[1]) / (2.0 * h)


def basis_directions(e, basic, x, field, index, tol=1e-9):
    """Whether the basis `basic` at the vertex x stays primal feasible when
    coordinate (field, index) moves up and when it moves down: (up, down).
    Only then is its sensitivity a one-sided derivative in that direction.
    A move of q_i shifts x_B by B^-1 e_i; a move of the bound a nonbasic
    column sits at shifts x_B by -B^-1 K[:, j]; c moves no primal value.
    A basic variable at a bound (a degenerate one, the row logicals
    included: logical i is K x - q, in [0, inf) on inequality rows and 0 on
    equality rows) must not be pushed past it."""
    if field == "c":
        return True, True
    m, n = len(e["start"]) - 1, len(e["c"])
    basic = np.asarray(basic, dtype=np.int64)
    B, k = basis_matrix(e, basic)
    if field == "q":
        rhs = np.zeros(m)
        rhs[index] = 1.0
    else:
        rhs = -k[:, int(index)].toarray().ravel()
    d = splu(B).solve(rhs)
    kx = k @ x
    neq = e["num_equalities"]
    up_ok = down_ok = True
    for pos, j in enumerate(basic):
        if abs(d[pos]) <= 1e-12:
            continue
        if j < n:
            v, lo, hi = x[j], e["lower"][j], e["upper"][j]
        else:
            r = int(j) - n
            v, lo, hi = kx[r] - e["q"][r], 0.0, (0.0 if r < neq else np.inf)
        at_lo = np.isfinite(lo) and abs(v - lo) <= tol * (1 + abs(lo))
        at_hi = np.isfinite(hi) and abs(v - hi) <= tol * (1 + abs(hi))
        # moving up by t > 0 changes v by t d; down by -t d
        if (at_lo and d[pos] < 0) or (at_hi and d[pos] > 0):
            up_ok = False
        if (at_lo and d[pos] > 0) or (at_hi and d[pos] < 0):
            down_ok = False
    return up_ok, down_ok
