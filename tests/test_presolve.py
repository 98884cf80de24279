"""Presolve and postsolve (anuk_presolve_*), on the CPU.

Held to the reference's presolve and PostsolveStack (presolve.cpp via
oracle/_ref): on every model of tests/golden/presolve_corpus.py (400 seeded
models built so that every rule fires, under nine option sets) and on the
reference's refinery.mps and small.mps, the status, message, counts and
dual-exactness must be equal, the reduced model bit for bit, and the primal
and dual postsolve of seeded vectors bit for bit. Against the committed
digests (tests/golden/presolve_reference.json, make_presolve_fixtures.py),
and against the reference build itself where oracle/_ref is present (1,500
models there). Beyond the reference: a postsolved exact optimum of the
reduced refinery model is feasible for the original and has its optimum.
"""
import gzip
import json
import os

import numpy as np
import pytest

from anukulan_amd import AnukError
from anukulan_amd.mps import read_mps_text
from anukulan_amd.presolve import presolve, presolve_mps_text
from golden.make_presolve_fixtures import READER_ONLY, h, vectors
from golden.mps_corpus import digest
from golden.presolve_corpus import cases

import _oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIX = json.load(open(os.path.join(GOLD, "presolve_reference.json")))
CASES = cases()


def data(name) -> bytes:
    return gzip.decompress(open(os.path.join(GOLD, name + ".mps.gz"), "rb").read())


def model_dict(m):
    return {"ok": True, "error": "", "warnings": m.warnings, "rows": m.rows, "cols": m.cols,
            "name": m.name, "maximize": m.maximize, "objective_offset": m.objective_offset,
            "free_rows_dropped": m.free_rows_dropped, "objective_nonzeros": m.objective_nonzeros,
            "fixed_format": m.fixed_format, "row_names": m.row_names, "col_names": m.col_names,
            "start": m.start, "index": m.index, "value": m.value, "objective": m.objective,
            "row_lower": m.row_lower, "row_upper": m.row_upper, "col_lower": m.col_lower,
            "col_upper": m.col_upper, "col_type": m.col_type, "hessian": m.hessian}


def record(P, seed) -> dict:
    """Our presolve's outcome in the fixture's form."""
    out = {"status": P.status, "message": P.message, "dual_exact": P.dual_exact,
           "counts": {k: (float(v).hex() if isinstance(v, float) else int(v))
                      for k, v in P.counts.items()}}
    if P.status != "reduced":
        return out
    red = digest(model_dict(P.reduced))
    for k in READER_ONLY:
        red.pop(k, None)
    out["reduced"] = red
    rx, ry, rc, oc = vectors(seed, P.reduced.rows, P.reduced.cols, P.original_cols)
    out["primal"] = h(P.postsolve_primal(rx))
    out["dual"] = h(P.postsolve_dual(ry, rc, oc))
    return out


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_presolve_matches_reference_digest(k):
    label, text, opts = CASES[k]
    assert record(presolve_mps_text(text, **opts), 1000 + k) == FIX["cases"][label]


@pytest.mark.parametrize("k,name", [(0, "refinery"), (1, "small")])
def test_reference_data_files(k, name):
    assert record(presolve_mps_text(data(name)), 5000 + k) == FIX["files"][name]


@pytest.mark.skipif(_oracle.ref is None, reason="oracle/_ref not built")
def test_presolve_matches_reference_live():
    for k, (label, text, opts) in enumerate(cases(1500)):
        mine = record(presolve_mps_text(text, **opts), 1000 + k)
        R = _oracle.RefPresolve(text, **opts)
        want = {"status": R.status, "message": R.message, "dual_exact": R.dual_exact,
                "counts": {c: (float(v).hex() if isinstance(v, float) else int(v))
                           for c, v in R.counts.items()}}
        if R.status == "reduced":
            red = digest(R.reduced)
            for key in READER_ONLY:
                red.pop(key, None)
            want["reduced"] = red
            rows, cols = R.reduced["rows"], R.reduced["cols"]
            oc_n, or_n = cols + R.counts["cols_removed"], rows + R.counts["rows_removed"]
            rx, ry, rc, oc = vectors(1000 + k, rows, cols, oc_n)
            want["primal"] = h(R.postsolve_primal(rx))
            want["dual"] = h(R.postsolve_dual(ry, rc, oc, or_n))
        assert mine == want, label


def test_refinery_reduction_is_the_references_published_one():
    # bench/results/ipm_refinery.txt:4-14: rows 1380 -> 1380, columns 4368 ->
    # 3072, nonzeros 10888 -> 9592, 2 rounds, 1296 slack singletons, 5138
    # bounds cut of which 4128 were infinite, the largest new one 1.5e6.
    P = presolve_mps_text(data("refinery"))
    assert (P.original_rows, P.reduced.rows, P.original_cols, P.reduced.cols) == (1380, 1380, 4368, 3072)
    assert (P.original_nnz, P.reduced.nnz) == (10888, 9592)
    c = P.counts
    assert (c["rounds"], c["slack_column_singletons"], c["bounds_tightened"],
            c["bounds_made_finite"], c["largest_new_finite_bound"]) == (2, 1296, 5138, 4128, 1.5e6)


@pytest.mark.skipif(_oracle.ref is None, reason="oracle/_ref not built")
def test_postsolved_optimum_is_the_original_optimum():
    # The reduced refinery model solved exactly (the reference's simplex),
    # then postsolved: feasible for the original model to rounding, and its
    # objective is the original optimum (2.426062266887e10,
    # bench/results/ipm_refinery.txt:19).
    raw = data("refinery")
    P = presolve_mps_text(raw)
    e = P.reduced.to_lp().export()
    rlp = _oracle.RefLp.from_export(e)
    sx = rlp.simplex()
    assert sx["status"] == 0
    x = P.postsolve_primal(sx["x"])
    m = read_mps_text(raw)
    act = np.zeros(m.rows)
    np.add.at(act, m.row_index(), m.value * x[m.index])
    scale = 1.0 + np.max(np.abs(np.where(np.isfinite(m.row_upper), m.row_upper, 0.0)))
    viol = np.maximum(m.row_lower - act, act - m.row_upper)
    assert np.max(viol) <= 1e-9 * scale
    assert np.all(x >= m.col_lower - 1e-9) and np.all(x <= m.col_upper + 1e-9)
    obj = float(m.objective @ x) + m.objective_offset
    assert abs(obj - 2.426062266887e10) <= 1e-9 * 2.426062266887e10


def test_infeasible_and_unbounded_are_statuses():
    base = "NAME T\nROWS\n N obj\n L r0\nCOLUMNS\n x obj {c} r0 1\nRHS\n rhs r0 {b}\nBOUNDS\n {bnd}\nENDATA\n"
    P = presolve_mps_text(base.format(c=1, b=-1, bnd="UP bnd x 5"))  # x <= -1 with x >= 0
    assert P.status == "infeasible" and P.reduced is None and P.message
    P = presolve_mps_text(base.format(c=-1, b=5, bnd="MI bnd x"))  # max x over x <= 5... min -x
    assert P.status in ("reduced", "unbounded")
    P = presolve_mps_text("NAME U\nROWS\n N obj\nCOLUMNS\n x obj -1\nENDATA\n")
    assert P.status == "unbounded"
    with pytest.raises(ValueError):
        P.postsolve_primal([])


def test_in_memory_models_and_options():
    from anukulan_amd import generators as G
    nm = G.refinery_model(4)
    P = presolve(nm)
    assert P.status == "reduced" and P.reduced.cols <= nm.cols
    off = presolve(nm, slack_column_singletons=0, bound_tightening=0, dual_fixing=0,
                   free_column_singletons=0, doubleton_equations=0)
    assert off.reduced.cols >= P.reduced.cols
    with pytest.raises(AttributeError):
        presolve(nm, no_such_rule=1)
    with pytest.raises(AnukError):
        presolve(nm, max_rounds=-1)
