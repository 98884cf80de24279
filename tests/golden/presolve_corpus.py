"""Small seeded models, written as MPS text, built so that every presolve
rule has something to act on: empty, singleton, forcing, redundant and
parallel rows; fixed, empty and dual-fixable columns; free and slack column
singletons in equality and inequality rows; doubleton equations; integer
columns with fractional bounds; and models presolve proves infeasible or
unbounded. Each model is paired with option sets (the defaults, one rule
switched off, coefficient tightening on) so the rules also run without their
neighbours.
"""
from __future__ import annotations

import numpy as np

INF = float("inf")


def _num(rng):
    """A coefficient: mostly small integers and halves, some full doubles."""
    u = rng.random()
    if u < 0.5:
        v = float(rng.integers(1, 6)) * (1 if rng.random() < 0.7 else -1)
    elif u < 0.7:
        v = float(rng.integers(1, 9)) / 2.0 * (1 if rng.random() < 0.6 else -1)
    else:
        v = float(rng.normal() * 3.0)
        if abs(v) < 0.05:
            v = 0.75
    return v


def random_model(rng):
    """(rows, cols, triplets, c, row_lo, row_up, col_lo, col_up, types, maximize)."""
    n = int(rng.integers(3, 14))
    # a point every column box will contain, so most models are feasible
    x0 = [float(rng.integers(0, 3)) for _ in range(n + 4)]
    trip = []
    rlo, rup = [], []
    cols = list(range(n))

    def add_row(entries, lo, up):
        i = len(rlo)
        for j, v in entries:
            trip.append((i, j, v))
        rlo.append(lo)
        rup.append(up)
        return i

    def bounds_for(entries, kind):
        act = sum(v * x0[j] for j, v in entries)  # activity at the interior point x0
        if kind == "E":
            return act, act
        if kind == "L":
            return -INF, act + float(rng.integers(0, 4))
        if kind == "G":
            return act - float(rng.integers(0, 4)), INF
        return act - float(rng.integers(1, 4)), act + float(rng.integers(1, 4))

    # ordinary rows
    for _ in range(int(rng.integers(2, 8))):
        k = int(rng.integers(1, min(5, n) + 1))
        ent = [(int(j), _num(rng)) for j in rng.choice(cols, size=k, replace=False)]
        add_row(ent, *bounds_for(ent, str(rng.choice(list("ELGR")))))
    specials = rng.random(9)
    if specials[0] < 0.5:  # a doubleton equality
        j1, j2 = (int(j) for j in rng.choice(cols, size=2, replace=False))
        ent = [(j1, _num(rng)), (j2, _num(rng))]
        add_row(ent, *bounds_for(ent, "E"))
    if specials[1] < 0.4:  # a singleton row
        ent = [(int(rng.integers(n)), _num(rng))]
        add_row(ent, *bounds_for(ent, str(rng.choice(list("ELGR")))))
    if specials[2] < 0.35 and len(rlo) > 0:  # a parallel copy of the first row
        base = [(j, v) for (i, j, v) in trip if i == 0]
        r = float(rng.choice([2.0, -1.0, 0.5, 3.0]))
        ent = [(j, v * r) for j, v in base]
        lo, up = rlo[0] * r, rup[0] * r
        if r < 0:
            lo, up = up, lo
        add_row(ent, lo - float(rng.integers(0, 2)) if np.isfinite(lo) else lo,
                up + float(rng.integers(0, 2)) if np.isfinite(up) else up)
    if specials[3] < 0.3:  # a forcing row: nonnegative terms held at or below zero
        k = int(rng.integers(1, 4))
        ent = [(int(j), abs(_num(rng))) for j in rng.choice(cols, size=k, replace=False)]
        add_row(ent, -INF, 0.0)
    if specials[4] < 0.3:  # a redundant row
        k = int(rng.integers(1, 3))
        ent = [(int(j), _num(rng)) for j in rng.choice(cols, size=k, replace=False)]
        add_row(ent, -1e6, 1e6)
    if specials[5] < 0.25:  # a row with no entries
        add_row([], -1.0 if rng.random() < 0.9 else 1.0, 2.0)
    m = len(rlo)

    clo, cup = [], []
    for j in range(n):
        u = rng.random()
        if u < 0.45:
            lo, up = 0.0, INF
        elif u < 0.65:
            lo = float(rng.integers(-3, 1))
            lo, up = lo, x0[j] + float(rng.integers(0, 6))
        elif u < 0.72:
            lo, up = x0[j], x0[j]  # fixed
        elif u < 0.82:
            lo, up = -INF, INF  # free
        elif u < 0.9:
            lo, up = -INF, x0[j] + float(rng.integers(0, 4))
        else:
            lo, up = 0.5 * float(rng.integers(-2, 1)), x0[j] + 0.5 * float(rng.integers(1, 8))
        clo.append(lo)
        cup.append(up)
    # some columns in no row at all
    for _ in range(int(rng.integers(0, 3))):
        clo.append(0.0 if rng.random() < 0.8 else -INF)
        cup.append(float(rng.integers(1, 5)) if rng.random() < 0.7 else INF)
    n_all = len(clo)
    c = [(_num(rng) if rng.random() < 0.75 else 0.0) for _ in range(n_all)]
    # mostly no push toward an infinite bound (a push there often proves the
    # model unbounded before anything else happens)
    for j in range(n_all):
        if rng.random() < 0.8:
            if np.isinf(clo[j]) and c[j] > 0:
                c[j] = -abs(c[j]) if not np.isinf(cup[j]) else 0.0
            if np.isinf(cup[j]) and c[j] < 0:
                c[j] = abs(c[j]) if not np.isinf(clo[j]) else 0.0
    # a free column alone in an equality row (a free column singleton)
    if specials[6] < 0.35:
        j = n_all
        c.append(_num(rng))
        clo.append(-INF)
        cup.append(INF)
        n_all += 1
        i = int(rng.integers(m)) if m else None
        if i is not None:
            trip.append((i, j, _num(rng)))
    types = [1 if rng.random() < 0.12 else 0 for _ in range(n_all)]
    maximize = bool(rng.random() < 0.3)
    return m, n_all, trip, c, rlo, rup, clo, cup, types, maximize


def _fmt(v):
    return repr(float(v))


def to_mps(model, name="P"):
    """Free-format MPS of a random_model tuple. Triplets on the same (row,
    column) are summed by both readers in file order."""
    m, n, trip, c, rlo, rup, clo, cup, types, maximize = model
    lines = [f"NAME          {name}"]
    if maximize:
        lines += ["OBJSENSE", "    MAX"]
    lines += ["ROWS", " N  obj"]
    kinds = []
    for i in range(m):
        lo, up = rlo[i], rup[i]
        if lo == up:
            k = "E"
        elif np.isinf(lo) and np.isinf(up):
            k = "N"
        elif np.isinf(lo):
            k = "L"
        elif np.isinf(up):
            k = "G"
        else:
            k = "R"
        kinds.append(k)
        lines.append(f" {'L' if k == 'R' else k}  r{i}")
    lines.append("COLUMNS")
    by_col = {}
    for i, j, v in trip:
        by_col.setdefault(j, []).append((i, v))
    in_int = False
    for j in range(n):
        if bool(types[j]) != in_int:
            in_int = bool(types[j])
            lines.append("    M  'MARKER'  " + ("'INTORG'" if in_int else "'INTEND'"))
        lines.append(f"    x{j}  obj  {_fmt(c[j])}")
        for i, v in by_col.get(j, []):
            lines.append(f"    x{j}  r{i}  {_fmt(v)}")
    if in_int:
        lines.append("    M  'MARKER'  'INTEND'")
    lines.append("RHS")
    for i in range(m):
        k = kinds[i]
        if k in "EG":
            lines.append(f"    rhs  r{i}  {_fmt(rlo[i])}")
        elif k in "LR":
            lines.append(f"    rhs  r{i}  {_fmt(rup[i])}")
    ranged = [i for i in range(m) if kinds[i] == "R"]
    if ranged:
        lines.append("RANGES")
        for i in ranged:
            lines.append(f"    rng  r{i}  {_fmt(rup[i] - rlo[i])}")
    lines.append("BOUNDS")
    for j in range(n):
        lo, up = clo[j], cup[j]
        if lo == up:
            lines.append(f" FX bnd x{j} {_fmt(lo)}")
            continue
        if np.isinf(lo) and np.isinf(up):
            lines.append(f" FR bnd x{j}")
            continue
        if np.isinf(lo):
            lines.append(f" MI bnd x{j}")
        elif lo != 0.0:
            lines.append(f" LO bnd x{j} {_fmt(lo)}")
        if not np.isinf(up):
            lines.append(f" UP bnd x{j} {_fmt(up)}")
    lines.append("ENDATA")
    return "\n".join(lines) + "\n"


OPTION_SETS = [
    {},
    {"doubleton_equations": 0},
    {"bound_tightening": 0},
    {"dual_fixing": 0, "inequality_column_singletons": 0},
    {"slack_column_singletons": 0, "free_column_singletons": 0},
    {"forcing_rows": 0, "redundant_rows": 0},
    {"duplicate_rows": 0, "singleton_rows": 0},
    {"coefficient_tightening": 1},
    {"max_rounds": 1},
]


def cases(count=400, seed=20261208):
    """(label, MPS text, presolve options) for `count` seeded models."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        model = random_model(rng)
        opts = OPTION_SETS[k % len(OPTION_SETS)]
        out.append((f"p{k}", to_mps(model, f"P{k}"), dict(opts)))
    return out
