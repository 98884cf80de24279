"""Writes tests/golden/adjoint_reference.npz: exact basis sensitivities from
the reference simplex's final basis (tests/_sensitivity.py) and central
finite differences through reference simplex re-solves, for

  random   generators.random_lp(m=300, n=700, seed=5), nondegenerate at its
           optimum; L = w_x'x + w_y'y with w ~ N(0,1) everywhere; full
           gradients and finite differences at 200 coordinates.
  stack4   the first 6 cases of the 4-period refinery stack; L = w'x with
           w ~ N(0,1) on the product-sale and crude-purchase columns; full
           gradients and finite differences at 200 coordinates per case.
  stack24  all 61 cases of the 24-period stack (the 61-LP case stack), the
           same loss; the basis sensitivities at 200 coordinates per case,
           the degeneracy of each case's vertex (basic columns at a bound,
           nonbasic columns with zero reduced cost), and one-sided finite
           differences at those coordinates: (L(theta + h) - L(theta)) / h
           and (L(theta) - L(theta - h)) / h, h = 1e-6 max(1, |theta|), each
           L through a re-solve by the reference's simplex warm-started from
           the case's own optimal basis (dual simplex, falling back to the
           primal as solve_lp does), the way the reference's case-stack
           workflow re-solves a case (app/anukulan_cli.cpp:1065-1230).
           Where the two sides agree the derivative along that coordinate
           is the same from either side.

Everything is in standard-form coordinates (c, q, lower, upper). Run in the
container that has the reference sources (several minutes):

    make -C oracle && python tests/golden/make_adjoint_fixtures.py [--only stack24]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
import _sensitivity as S  # noqa: E402
from anukulan_amd import generators as G  # noqa: E402

FIELDS = ("c", "q", "lower", "upper")
RANDOM_SIZE = dict(m=300, n=700, seed=5)
COORDS = 200


def seeds_on(model, rng):
    w = np.zeros(model.cols)
    cols = [j for period in model.tags["buy"] + model.tags["sell"] for j in period]
    w[cols] = rng.standard_normal(len(cols))
    return w


def pick_coordinates(rng, grads, count):
    """count coordinates over (field, index): rows of q, columns of c, and
    the columns at a bound (where that bound's gradient lives)."""
    gc, gq, gl, gu = grads
    pool = [(1, i) for i in range(len(gq))] + [(0, j) for j in range(len(gc))]
    pool += [(2, j) for j in np.flatnonzero(gl)] + [(3, j) for j in np.flatnonzero(gu)]
    pick = rng.choice(len(pool), size=min(count, len(pool)), replace=False)
    return np.array([pool[p] for p in pick], dtype=np.int64)


def degeneracy(e, sx):
    n, m = len(e["c"]), len(e["q"])
    x, y = sx["x"], sx["y"]
    basic = np.zeros(n + m, bool)
    basic[sx["basic"]] = True
    lo, up = e["lower"], e["upper"]
    at_bound = (np.isfinite(lo) & (np.abs(x - lo) <= 1e-9 * (1 + np.abs(lo)))) | \
               (np.isfinite(up) & (np.abs(x - up) <= 1e-9 * (1 + np.abs(up))))
    start, index, value = e["start"], e["index"], e["value"]
    kty = np.zeros(n)
    np.add.at(kty, index, value * np.repeat(y, np.diff(start)))
    rc = e["c"] - kty
    return int((basic[:n] & at_bound).sum()), int((~basic[:n] & (np.abs(rc) < 1e-9)).sum())


def solve_case(model, w_x, w_y, rng, fd):
    rlp = _oracle.RefLp.from_natural(model)
    e = rlp.export()
    sx = rlp.simplex()
    assert sx["status"] == 0, model.name
    grads = S.basis_sensitivity(e, sx["basic"], sx["x"], w_x, w_y)
    coords = pick_coordinates(rng, grads, COORDS)
    out = {"objective": sx["objective"], "grads": grads, "coords": coords,
           "basis_values": np.array([grads[f][i] for f, i in coords]),
           "degeneracy": degeneracy(e, sx)}
    if fd:
        def loss(x, y):
            return float(w_x @ x + (0.0 if w_y is None else w_y @ y))
        vals = []
        for f, i in coords:
            d = S.central_difference(_oracle, e, FIELDS[f], int(i), loss)
            vals.append(np.nan if d is None else d)
        out["fd_values"] = np.array(vals)
    return out


def one_sided(rlp_e, basis, field, index, loss, l0, rel_step=1e-6):
    """(L(theta + h) - L(theta)) / h and (L(theta) - L(theta - h)) / h
    through warm re-solves from `basis` (nan when a re-solve fails)."""
    h = rel_step * max(1.0, abs(float(rlp_e[field][index])))
    out = []
    for s in (1.0, -1.0):
        sx = _oracle.RefLp.from_export(S.perturbed(rlp_e, field, index, s * h)).simplex_warm(
            dual=True, start=basis)
        if sx["status"] != 0:
            out.append(np.nan)
            continue
        out.append(s * (loss(sx["x"]) - l0) / h)
    return out


def stack24_case(i, base, case, w, coord_seed):
    model = G.case_model(base, case)
    rlp = _oracle.RefLp.from_natural(model)
    e = rlp.export()
    sx = rlp.simplex_warm(dual=False)  # the reference's default solve (primal, cold)
    assert sx["status"] == 0, model.name
    grads = S.basis_sensitivity(e, sx["basic"], sx["x"], w, None)
    coords = pick_coordinates(np.random.default_rng(coord_seed), grads, COORDS)
    basis = (sx["basic"], sx["var_status"])
    l0 = float(w @ sx["x"])
    fd = np.array([one_sided(e, basis, FIELDS[f], int(j), lambda x: float(w @ x), l0)
                   for f, j in coords])
    return {"coords": coords, "basis": np.array([grads[f][j] for f, j in coords]),
            "fd_plus": fd[:, 0], "fd_minus": fd[:, 1], "degeneracy": degeneracy(e, sx),
            "objective": sx["objective"]}


def stack24_section(data, threads=8):
    """The 61-case stack: weights from a seeded stream, one coordinate seed
    per case, cases solved in parallel (the ctypes calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    base, cases = G.case_stack(24)
    rng = np.random.default_rng(20261207)
    ws = [seeds_on(base, rng) for _ in cases]
    coord_seeds = rng.integers(0, 2**32, size=len(cases))
    with ThreadPoolExecutor(max_workers=threads) as pool:
        rs = list(pool.map(lambda k: stack24_case(k, base, cases[k], ws[k], int(coord_seeds[k])),
                           range(len(cases))))
    for i, r in enumerate(rs):
        data[f"stack24_{i}_w_x"] = ws[i]
        data[f"stack24_{i}_coords"] = r["coords"]
        data[f"stack24_{i}_basis"] = r["basis"]
        data[f"stack24_{i}_fd_plus"] = r["fd_plus"]
        data[f"stack24_{i}_fd_minus"] = r["fd_minus"]
        data[f"stack24_{i}_degeneracy"] = np.array(r["degeneracy"])
        data[f"stack24_{i}_objective"] = np.array(r["objective"])
        print("stack24", i, r["degeneracy"], flush=True)


def main():
    if _oracle.ref is None:
        sys.exit("oracle/_ref/libanukulan_ref.so is not built")
    path = os.path.join(HERE, "adjoint_reference.npz")
    if sys.argv[1:] == ["--only", "stack24"]:
        data = {k: v for k, v in np.load(path).items() if not k.startswith("stack24_")}
        stack24_section(data)
        np.savez_compressed(path, **data)
        print("updated", path)
        return
    data = {}
    rng = np.random.default_rng(20261205)
    model = G.random_lp(**RANDOM_SIZE)
    wx, wy = rng.standard_normal(model.cols), rng.standard_normal(model.rows)
    r = solve_case(model, wx, wy, rng, fd=True)
    data.update({"random_w_x": wx, "random_w_y": wy, "random_coords": r["coords"],
                 "random_fd": r["fd_values"], "random_basis": r["basis_values"],
                 "random_degeneracy": np.array(r["degeneracy"])})
    for name, g in zip(FIELDS, r["grads"]):
        data[f"random_grad_{name}"] = g
    print("random done", r["degeneracy"])

    base, cases = G.case_stack(4)
    for i, case in enumerate(cases[:6]):
        model = G.case_model(base, case)
        w = seeds_on(base, rng)
        r = solve_case(model, w, None, rng, fd=True)
        data[f"stack4_{i}_w_x"] = w
        data[f"stack4_{i}_coords"] = r["coords"]
        data[f"stack4_{i}_fd"] = r["fd_values"]
        data[f"stack4_{i}_basis"] = r["basis_values"]
        data[f"stack4_{i}_degeneracy"] = np.array(r["degeneracy"])
        for name, g in zip(FIELDS, r["grads"]):
            data[f"stack4_{i}_grad_{name}"] = g
        print("stack4", i, r["degeneracy"])

    stack24_section(data)
    np.savez_compressed(path, **data)
    print("wrote", path)


if __name__ == "__main__":
    main()
