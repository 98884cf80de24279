"""Known-answer LPs of the reference's PDHG tests, stated in natural form.

Each entry restates one embedded test model of tests/test_pdhg.cpp as rows,
columns and bounds (row_lower <= A x <= row_upper, col bounds, objective,
sense, offset) together with the optimum that test asserts. INF marks an
absent bound. Data only.
"""
import math

INF = math.inf

KNOWN = {
    # test_pdhg.cpp:38-65 - max 3x + 5y, x <= 4, 2y <= 12, 3x + 2y <= 18 -> 36 at (2, 6)
    "wyndor": dict(rows=3, cols=2, triplets=[(0, 0, 1.0), (1, 1, 2.0), (2, 0, 3.0), (2, 1, 2.0)],
                   objective=[3.0, 5.0], row_lower=[-INF] * 3, row_upper=[4.0, 12.0, 18.0],
                   col_lower=[0.0, 0.0], col_upper=[INF, INF], maximize=True, offset=0.0,
                   optimum=36.0, x=[2.0, 6.0]),
    # test_pdhg.cpp:67-88 - min x + y, x + y = 10, x - y = 2 -> 10 at (6, 4)
    "equalities": dict(rows=2, cols=2, triplets=[(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, -1.0)],
                       objective=[1.0, 1.0], row_lower=[10.0, 2.0], row_upper=[10.0, 2.0],
                       col_lower=[0.0, 0.0], col_upper=[INF, INF], maximize=False, offset=0.0,
                       optimum=10.0, x=[6.0, 4.0]),
    # test_pdhg.cpp:90-113 - min -x - y, x + y <= 10, 1 <= x <= 3, y free -> -10
    "boxed_free": dict(rows=1, cols=2, triplets=[(0, 0, 1.0), (0, 1, 1.0)],
                       objective=[-1.0, -1.0], row_lower=[-INF], row_upper=[10.0],
                       col_lower=[1.0, -INF], col_upper=[3.0, INF], maximize=False, offset=0.0,
                       optimum=-10.0, x=None),
    # test_pdhg.cpp:115-133 - min x + 2y, 5 <= x + y <= 8 -> 5 at (5, 0)
    "ranged": dict(rows=1, cols=2, triplets=[(0, 0, 1.0), (0, 1, 1.0)], objective=[1.0, 2.0],
                   row_lower=[5.0], row_upper=[8.0], col_lower=[0.0, 0.0], col_upper=[INF, INF],
                   maximize=False, offset=0.0, optimum=5.0, x=[5.0, 0.0]),
    # test_pdhg.cpp:135-167 - max x + y, x + y <= 2, x <= 1, y <= 1 -> 2 at (1, 1)
    "degenerate": dict(rows=3, cols=2, triplets=[(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (2, 1, 1.0)],
                       objective=[1.0, 1.0], row_lower=[-INF] * 3, row_upper=[2.0, 1.0, 1.0],
                       col_lower=[0.0, 0.0], col_upper=[INF, INF], maximize=True, offset=0.0,
                       optimum=2.0, x=[1.0, 1.0]),
    # test_pdhg.cpp:169-184 - min x + 7, x >= 4 -> 11
    "offset": dict(rows=1, cols=1, triplets=[(0, 0, 1.0)], objective=[1.0], row_lower=[4.0],
                   row_upper=[INF], col_lower=[0.0], col_upper=[INF], maximize=False, offset=7.0,
                   optimum=11.0, x=[4.0]),
    # test_pdhg.cpp:292-313 - a row with no coefficients; x in [0,5] cost 1, y in [0,5] cost -1 -> -5
    "empty_matrix": dict(rows=1, cols=2, triplets=[], objective=[1.0, -1.0], row_lower=[-1.0],
                         row_upper=[INF], col_lower=[0.0, 0.0], col_upper=[5.0, 5.0],
                         maximize=False, offset=0.0, optimum=-5.0, x=[0.0, 5.0]),
    # test_pdhg.cpp:206-224 - rows 3x >= 1 and 4y >= 1: ||K|| = 4
    "norm_diag": dict(rows=2, cols=2, triplets=[(0, 0, 3.0), (1, 1, 4.0)], objective=[1.0, 1.0],
                      row_lower=[1.0, 1.0], row_upper=[INF, INF], col_lower=[0.0, 0.0],
                      col_upper=[INF, INF], maximize=False, offset=0.0, optimum=7.0 / 12.0,
                      x=[1.0 / 3.0, 0.25], norm=4.0),
    # A primal infeasible LP for the Farkas test: x + y >= 3 with x, y in [0, 1].
    "infeasible": dict(rows=1, cols=2, triplets=[(0, 0, 1.0), (0, 1, 1.0)], objective=[1.0, 1.0],
                       row_lower=[3.0], row_upper=[INF], col_lower=[0.0, 0.0], col_upper=[1.0, 1.0],
                       maximize=False, offset=0.0, optimum=None, x=None),
}


def natural(name):
    """The LP as a generators.NaturalModel."""
    import numpy as np
    from anukulan_amd.generators import NaturalModel
    d = KNOWN[name]
    t = d["triplets"]
    return NaturalModel(
        rows=d["rows"], cols=d["cols"],
        row_index=np.array([a for a, _, _ in t], np.int32),
        col_index=np.array([b for _, b, _ in t], np.int32),
        value=np.array([v for _, _, v in t], np.float64),
        objective=np.array(d["objective"], np.float64),
        row_lower=np.array(d["row_lower"], np.float64),
        row_upper=np.array(d["row_upper"], np.float64),
        col_lower=np.array(d["col_lower"], np.float64),
        col_upper=np.array(d["col_upper"], np.float64),
        maximize=d["maximize"], objective_offset=d["offset"], name=name)
