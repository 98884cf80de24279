"""Seeded synthetic LPs for the PDHG workloads.

refinery_model   a multi-period refinery plan: crude purchase and storage,
                 a CDU with two operating modes and crude-segregated cuts,
                 FCC / hydrotreater / reformer conversion, blending into
                 eight products under sulphur and octane specifications, and
                 product storage linking each period to the next.
case_stack       the base plan plus 60 single-datum what-if cases on it.
random_lp        a large random LP, feasible and bounded by construction.

Every generator is a pure function of its arguments (numpy PCG64 seeded
from them). Models are returned in natural form (row bounds, column bounds,
objective, triplets) for LP.from_model, which applies the reference's
standard-form rules.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List

import numpy as np

INF = np.inf


@dataclass
class NaturalModel:
    rows: int
    cols: int
    row_index: np.ndarray
    col_index: np.ndarray
    value: np.ndarray
    objective: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    maximize: bool = False
    objective_offset: float = 0.0
    name: str = ""
    tags: dict = field(default_factory=dict)

    def to_lp(self):
        from .api import LP
        return LP.from_model(self.rows, self.cols, self.row_index, self.col_index, self.value,
                             self.objective, self.row_lower, self.row_upper, self.col_lower,
                             self.col_upper, self.maximize, self.objective_offset)


class _Builder:
    def __init__(self):
        self.ri: List[int] = []
        self.ci: List[int] = []
        self.v: List[float] = []
        self.cost: List[float] = []
        self.clo: List[float] = []
        self.cup: List[float] = []
        self.rlo: List[float] = []
        self.rup: List[float] = []

    def col(self, cost=0.0, lo=0.0, up=INF) -> int:
        self.cost.append(cost)
        self.clo.append(lo)
        self.cup.append(up)
        return len(self.cost) - 1

    def row(self, lo, up) -> int:
        self.rlo.append(lo)
        self.rup.append(up)
        return len(self.rlo) - 1

    def add(self, r, c, v):
        if v != 0.0:
            self.ri.append(r)
            self.ci.append(c)
            self.v.append(float(v))

    def model(self, name, tags) -> NaturalModel:
        return NaturalModel(
            rows=len(self.rlo), cols=len(self.cost), row_index=np.array(self.ri, np.int32),
            col_index=np.array(self.ci, np.int32), value=np.array(self.v),
            objective=np.array(self.cost), row_lower=np.array(self.rlo),
            row_upper=np.array(self.rup), col_lower=np.array(self.clo),
            col_upper=np.array(self.cup), name=name, tags=tags)


CUTS = ["LPG", "LIGHT_NAPHTHA", "HEAVY_NAPHTHA", "KEROSENE", "GASOIL", "RESIDUE"]
PRODUCTS = ["LPG", "GASOLINE95", "GASOLINE91", "PETCHEM_NAPHTHA", "JET", "DIESEL", "FUEL_OIL",
            "LSFO"]
UNITS = ["FCC", "HDT", "REFORMER"]
# Where each straight-run cut may go: product names or unit names.
CUT_DEST = {
    "LPG": ["LPG"],
    "LIGHT_NAPHTHA": ["GASOLINE95", "GASOLINE91", "PETCHEM_NAPHTHA", "REFORMER"],
    "HEAVY_NAPHTHA": ["GASOLINE91", "PETCHEM_NAPHTHA", "REFORMER"],
    "KEROSENE": ["JET", "DIESEL", "HDT"],
    "GASOIL": ["DIESEL", "FUEL_OIL", "HDT", "FCC"],
    "RESIDUE": ["FUEL_OIL", "LSFO", "FCC"],
}
# Unit outputs: (unit, stream, yield per unit feed, sulphur %, octane, destinations)
UNIT_OUT = [
    ("FCC", "FCC_GASOLINE", 0.55, 0.08, 92.0, ["GASOLINE95", "GASOLINE91"]),
    ("FCC", "LIGHT_CYCLE_OIL", 0.25, 1.20, 0.0, ["DIESEL", "FUEL_OIL", "LSFO"]),
    ("FCC", "SLURRY", 0.10, 2.50, 0.0, ["FUEL_OIL", "LSFO"]),
    ("HDT", "TREATED_KERO", 0.48, 0.01, 0.0, ["JET", "DIESEL"]),
    ("HDT", "TREATED_DIESEL", 0.49, 0.01, 0.0, ["DIESEL", "LSFO"]),
    ("REFORMER", "REFORMATE", 0.85, 0.0005, 98.0, ["GASOLINE95", "GASOLINE91"]),
]
CUT_SULPHUR = np.array([0.01, 0.03, 0.08, 0.25, 0.90, 2.20])   # x crude sulphur
CUT_OCTANE = {"LIGHT_NAPHTHA": 72.0, "HEAVY_NAPHTHA": 58.0}
SULPHUR_MAX = {"GASOLINE95": 0.05, "GASOLINE91": 0.08, "JET": 0.20, "DIESEL": 0.15,
               "FUEL_OIL": 3.50, "LSFO": 0.50}
OCTANE_MIN = {"GASOLINE95": 95.0, "GASOLINE91": 91.0}
UNIT_COST = {"FCC": 3.0, "HDT": 2.0, "REFORMER": 2.5}
CDU_MODES = 2


@dataclass
class RefineryData:
    """Every random number the refinery model reads, drawn once per seed."""
    periods: int
    crude_cost: np.ndarray      # periods x 6, U[40,90]
    crude_supply: np.ndarray    # periods x 6, U[50,200]
    crude_sulphur: np.ndarray   # 6, U[0.1,3.0] %
    yields: np.ndarray          # 6 crudes x modes x 6 cuts, Dirichlet(1,...,1)
    unit_capacity: np.ndarray   # periods x 3 (FCC, HDT, REFORMER), U[100,400]
    cdu_capacity: np.ndarray    # periods, U[500,900]
    price: np.ndarray           # periods x 8, U[50,120] (sale revenue per unit)
    demand: np.ndarray          # periods x 8, U[20,150]


def refinery_data(periods: int, seed: int) -> RefineryData:
    g = np.random.default_rng(np.random.PCG64([seed, periods, 0x5EF1]))
    return RefineryData(
        periods=periods,
        crude_cost=g.uniform(40.0, 90.0, (periods, 6)),
        crude_supply=g.uniform(50.0, 200.0, (periods, 6)),
        crude_sulphur=g.uniform(0.1, 3.0, 6),
        yields=g.dirichlet(np.ones(6), (6, CDU_MODES)),
        unit_capacity=g.uniform(100.0, 400.0, (periods, 3)),
        cdu_capacity=g.uniform(500.0, 900.0, periods),
        price=g.uniform(50.0, 120.0, (periods, 8)),
        demand=g.uniform(20.0, 150.0, (periods, 8)),
    )


def refinery_model(periods: int = 12, seed: int = 20261203, data: RefineryData = None) -> NaturalModel:
    """Minimise crude + operating + holding cost minus sales revenue."""
    d = data if data is not None else refinery_data(periods, seed)
    b = _Builder()
    P = {p: i for i, p in enumerate(PRODUCTS)}
    prev_crude = [None] * 6
    prev_prod = [None] * 8
    tags = {"buy": [], "sell": [], "cdu_cap_row": [], "unit_cap_row": []}
    for t in range(periods):
        # columns of the period
        buy = [b.col(d.crude_cost[t, c], 0.0, d.crude_supply[t, c]) for c in range(6)]
        store = [b.col(0.2, 0.0, 300.0) for _ in range(6)]
        run = [[b.col(0.0) for _ in range(CDU_MODES)] for _ in range(6)]
        cut_flow = {}   # (c, mode, k, dest) -> col
        for c in range(6):
            for mo in range(CDU_MODES):
                for k, cut in enumerate(CUTS):
                    for dest in CUT_DEST[cut]:
                        cost = UNIT_COST.get(dest, 0.0)
                        cut_flow[(c, mo, k, dest)] = b.col(cost)
        out_flow = {}   # (stream index, dest) -> col
        for s, (_u, _name, _y, _sul, _oct, dests) in enumerate(UNIT_OUT):
            for dest in dests:
                out_flow[(s, dest)] = b.col(0.0)
        sell = [b.col(-d.price[t, p], 0.0, d.demand[t, p]) for p in range(8)]
        pstore = [b.col(0.3, 0.0, 200.0) for _ in range(8)]
        tags["buy"].append(buy)
        tags["sell"].append(sell)

        # crude tanks: carried-in + bought - charged - carried-out = 0
        for c in range(6):
            r = b.row(0.0, 0.0)
            if prev_crude[c] is not None:
                b.add(r, prev_crude[c], 1.0)
            b.add(r, buy[c], 1.0)
            for mo in range(CDU_MODES):
                b.add(r, run[c][mo], -1.0)
            b.add(r, store[c], -1.0)
        # CDU capacity
        r = b.row(-INF, d.cdu_capacity[t])
        tags["cdu_cap_row"].append(r)
        for c in range(6):
            for mo in range(CDU_MODES):
                b.add(r, run[c][mo], 1.0)
        # each segregated cut: yield * charge = sum of its dispositions
        for c in range(6):
            for mo in range(CDU_MODES):
                for k, cut in enumerate(CUTS):
                    r = b.row(0.0, 0.0)
                    b.add(r, run[c][mo], d.yields[c, mo, k])
                    for dest in CUT_DEST[cut]:
                        b.add(r, cut_flow[(c, mo, k, dest)], -1.0)
        # conversion unit capacity (feed) and output balances
        unit_rows = []
        for u, unit in enumerate(UNITS):
            r = b.row(-INF, d.unit_capacity[t, u])
            unit_rows.append(r)
            for (c, mo, k, dest), col in cut_flow.items():
                if dest == unit:
                    b.add(r, col, 1.0)
        tags["unit_cap_row"].append(unit_rows)
        for s, (unit, _name, y, _sul, _oct, dests) in enumerate(UNIT_OUT):
            r = b.row(0.0, 0.0)
            for (c, mo, k, dest), col in cut_flow.items():
                if dest == unit:
                    b.add(r, col, y)
            for dest in dests:
                b.add(r, out_flow[(s, dest)], -1.0)
        # product inventory balance: carried-in + blended - sold - carried-out = 0
        inputs = {p: [] for p in PRODUCTS}   # (col, sulphur, octane)
        for (c, mo, k, dest), col in cut_flow.items():
            if dest in P:
                oct_ = CUT_OCTANE.get(CUTS[k], 0.0)
                inputs[dest].append((col, d.crude_sulphur[c] * CUT_SULPHUR[k], oct_))
        for (s, dest), col in out_flow.items():
            _u, _name, _y, sul, oct_, _d = UNIT_OUT[s]
            inputs[dest].append((col, sul, oct_))
        for p, prod in enumerate(PRODUCTS):
            r = b.row(0.0, 0.0)
            if prev_prod[p] is not None:
                b.add(r, prev_prod[p], 1.0)
            for col, _s, _o in inputs[prod]:
                b.add(r, col, 1.0)
            b.add(r, sell[p], -1.0)
            b.add(r, pstore[p], -1.0)
        # blend specifications, linear in the blended volumes
        for prod, smax in SULPHUR_MAX.items():
            r = b.row(0.0, INF)
            for col, sul, _o in inputs[prod]:
                b.add(r, col, smax - sul)
        for prod, omin in OCTANE_MIN.items():
            r = b.row(0.0, INF)
            for col, _s, oct_ in inputs[prod]:
                b.add(r, col, oct_ - omin)
        prev_crude = store
        prev_prod = pstore
    m = b.model(f"refinery{periods}", tags)
    return m


@dataclass
class Case:
    name: str
    objective: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray


def case_stack(periods: int = 24, seed: int = 20261203) -> (NaturalModel, List[Case]):
    """The base plan and 60 one-datum moves on it.

    24 price moves (each crude, +-3% and +-6%, every period), 6 supply moves
    (one crude's supply halved in every period), 16 demand moves (each
    product, +-10%, every period), 12 capacity cuts (CDU, FCC or hydrotreater
    to 80% in one of periods 1-4) and 2 reformer cuts (80% in period 1 or 2),
    which brings the count to 60. K is the same in every case.
    """
    base = refinery_model(periods, seed)
    cases = [Case("base", base.objective.copy(), base.row_lower.copy(), base.row_upper.copy(),
                  base.col_lower.copy(), base.col_upper.copy())]

    def derive(name):
        return Case(name, base.objective.copy(), base.row_lower.copy(), base.row_upper.copy(),
                    base.col_lower.copy(), base.col_upper.copy())

    buy, sell = base.tags["buy"], base.tags["sell"]
    for c in range(6):
        for pct in (-6, -3, 3, 6):
            k = derive(f"crude{c}_price{pct:+d}%")
            for t in range(periods):
                k.objective[buy[t][c]] *= 1.0 + pct / 100.0
            cases.append(k)
    for c in range(6):
        k = derive(f"crude{c}_supply_halved")
        for t in range(periods):
            k.col_upper[buy[t][c]] *= 0.5
        cases.append(k)
    for p in range(8):
        for pct in (-10, 10):
            k = derive(f"{PRODUCTS[p]}_demand{pct:+d}%")
            for t in range(periods):
                k.col_upper[sell[t][p]] *= 1.0 + pct / 100.0
            cases.append(k)
    for unit, rows in (("CDU", [[r] for r in base.tags["cdu_cap_row"]]),
                       ("FCC", [[u[0]] for u in base.tags["unit_cap_row"]]),
                       ("HDT", [[u[1]] for u in base.tags["unit_cap_row"]])):
        for t in range(4):
            k = derive(f"{unit}_capacity_cut20%_period{t + 1}")
            for r in rows[t]:
                k.row_upper[r] *= 0.8
            cases.append(k)
    for t in range(2):
        k = derive(f"REFORMER_capacity_cut20%_period{t + 1}")
        k.row_upper[base.tags["unit_cap_row"][t][2]] *= 0.8
        cases.append(k)
    assert len(cases) == 61
    return base, cases


def case_model(base: NaturalModel, case: Case) -> NaturalModel:
    return replace(base, objective=case.objective, row_lower=case.row_lower,
                   row_upper=case.row_upper, col_lower=case.col_lower, col_upper=case.col_upper,
                   name=f"{base.name}:{case.name}")


def zipf_exponent_for_mean(target: float, lo: int = 2, hi: int = 5000) -> float:
    """Exponent a with E[k] = target for P(k) ~ k^-a on [lo, hi] (bisection)."""
    k = np.arange(lo, hi + 1, dtype=np.float64)

    def mean(a):
        w = k ** -a
        return float((k * w).sum() / w.sum())

    a_lo, a_hi = 1.01, 4.0
    for _ in range(100):
        mid = 0.5 * (a_lo + a_hi)
        if mean(mid) > target:
            a_lo = mid
        else:
            a_hi = mid
    return 0.5 * (a_lo + a_hi)


def random_lp(m: int = 400_000, n: int = 1_000_000, mean_col_nnz: float = 8.0,
              seed: int = 20261204) -> NaturalModel:
    """Random LP, feasible and bounded by construction.

    Column counts are Zipf on [2, 5000] with the exponent chosen so their
    mean is `mean_col_nnz` (about 1.95 for 8); rows are drawn uniformly
    without repeats per column; values N(0,1) redrawn below |a| = 0.05.
    x0 ~ U[0,10]; the first half of the rows are equalities A x = A x0, the
    rest A x <= A x0 + s with s ~ U[0,1]. The objective is c = A'y0 + s0 with
    s0 ~ U[0,1], y0 ~ N(0,1) on equality rows and -|N(0,1)| on the <= rows, so
    y0 is dual feasible and the LP is bounded. Bounds: l = 0, u = inf on 80%
    of columns and U[10,100] on the rest (above x0, so x0 stays feasible).
    """
    g = np.random.default_rng(np.random.PCG64([seed, m, n]))
    a = zipf_exponent_for_mean(mean_col_nnz)
    ks = np.arange(2, 5001, dtype=np.float64)
    p = ks ** -a
    p /= p.sum()
    counts = g.choice(np.arange(2, 5001), size=n, p=p)
    counts = np.minimum(counts, m)
    total = int(counts.sum())
    cols = np.repeat(np.arange(n, dtype=np.int64), counts)
    rows = g.integers(0, m, size=total, dtype=np.int64)
    # one entry per (row, col): drop repeats
    key = cols * m + rows
    key = np.unique(key)
    cols = (key // m).astype(np.int32)
    rows = (key % m).astype(np.int32)
    vals = g.standard_normal(key.size)
    small = np.abs(vals) < 0.05
    while small.any():
        vals[small] = g.standard_normal(int(small.sum()))
        small = np.abs(vals) < 0.05
    x0 = g.uniform(0.0, 10.0, n)
    ax0 = np.zeros(m)
    np.add.at(ax0, rows, vals * x0[cols])
    neq = m // 2
    slack = g.uniform(0.0, 1.0, m - neq)
    row_lower = np.concatenate([ax0[:neq], np.full(m - neq, -INF)])
    row_upper = np.concatenate([ax0[:neq], ax0[neq:] + slack])
    y0 = np.concatenate([g.standard_normal(neq), -np.abs(g.standard_normal(m - neq))])
    aty0 = np.zeros(n)
    np.add.at(aty0, cols, vals * y0[rows])
    c = aty0 + g.uniform(0.0, 1.0, n)
    col_upper = np.full(n, INF)
    boxed = g.uniform(0.0, 1.0, n) < 0.2
    col_upper[boxed] = np.maximum(g.uniform(10.0, 100.0, int(boxed.sum())), x0[boxed] + 1e-9)
    return NaturalModel(rows=m, cols=n, row_index=rows, col_index=cols, value=vals, objective=c,
                        row_lower=row_lower, row_upper=row_upper, col_lower=np.zeros(n),
                        col_upper=col_upper, name=f"random{m}x{n}",
                        tags={"zipf_exponent": a, "nnz": int(key.size)})


def uneven_stack(count: int = 512, seed: int = 20261206, max_rows: int = 50_000) -> List[NaturalModel]:
    """Independent LPs of uneven size, each with its own constraint matrix:
    count // 2 refinery plans with period counts drawn uniformly from
    {4, 12, 24, 52} (each with its own data seed), and the rest random LPs
    (random_lp) with m ~ logU[1e3, max_rows], n = 2.5 m and a mean of 8
    nonzeros per column, shuffled together."""
    g = np.random.default_rng(np.random.PCG64([seed, count]))
    models = []
    for i in range(count // 2):
        periods = int(g.choice([4, 12, 24, 52]))
        models.append(refinery_model(periods, seed=seed + 1 + i))
    for i in range(count - count // 2):
        m = int(round(float(np.exp(g.uniform(np.log(1e3), np.log(max_rows))))))
        models.append(random_lp(m=m, n=int(2.5 * m), seed=seed + 100_000 + i))
    order = g.permutation(len(models))
    return [models[i] for i in order]
