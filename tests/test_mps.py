"""The MPS reader and writer (anuk_mps_read / anuk_mps_write), on the CPU.

The reader is held to the reference's reader (read_mps_stream,
src/anukulan/mps_reader.cpp) field by field, on every text of
tests/golden/mps_corpus.py and on the reference's data files: against the
committed digests of what the reference's reader made of them
(tests/golden/mps_reference.json, written by make_mps_fixtures.py), and
against the reference build itself when oracle/_ref is present. A failed read
is compared by its message.
"""
import gzip
import json
import os

import numpy as np
import pytest

from anukulan_amd import AnukError, generators as G
from anukulan_amd.mps import MpsModel, read_mps, read_mps_text, write_mps
from golden.mps_corpus import all_cases, digest, lp_digest

import _oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIX = json.load(open(os.path.join(GOLD, "mps_reference.json")))
CASES = all_cases()
FORMAT_NAMES = {0: "auto", 1: "free", 2: "fixed"}


def data(name) -> bytes:
    return gzip.decompress(open(os.path.join(GOLD, name + ".mps.gz"), "rb").read())


def keywords(opts) -> dict:
    """The corpus's reader options as read_mps_text keywords."""
    kw = {}
    for k, v in opts.items():
        if k == "format":
            kw["format"] = FORMAT_NAMES[v]
        elif k == "negative_up":
            kw["negative_up_bound"] = "cplex" if v else "highs"
        else:
            kw[k] = v
    return kw


def ours(text, **opts) -> dict:
    """Our reader's result in the form ref_read_mps returns."""
    try:
        m = read_mps_text(text, **keywords(opts))
    except AnukError as e:
        assert e.code in (6, 7), e
        return {"ok": False, "error": e.message, "warnings": []}
    return {"ok": True, "error": "", "warnings": m.warnings, "rows": m.rows, "cols": m.cols,
            "name": m.name, "maximize": m.maximize, "objective_offset": m.objective_offset,
            "free_rows_dropped": m.free_rows_dropped, "objective_nonzeros": m.objective_nonzeros,
            "fixed_format": m.fixed_format, "row_names": m.row_names, "col_names": m.col_names,
            "start": m.start, "index": m.index, "value": m.value, "objective": m.objective,
            "row_lower": m.row_lower, "row_upper": m.row_upper, "col_lower": m.col_lower,
            "col_upper": m.col_upper, "col_type": m.col_type, "hessian": m.hessian}


def comparable(d: dict) -> dict:
    """A failed read is compared by its message (our reader does not return
    the warnings gathered before the failure)."""
    return {"ok": False, "error": d["error"]} if not d["ok"] else d


@pytest.mark.parametrize("label,text,opts", CASES, ids=[c[0] for c in CASES])
def test_reader_matches_reference_digest(label, text, opts):
    assert comparable(digest(ours(text, **opts))) == comparable(FIX["cases"][label])


@pytest.mark.skipif(_oracle.ref is None, reason="oracle/_ref not built")
def test_reader_matches_reference_reader_live():
    for label, text, opts in CASES:
        want = comparable(digest(_oracle.ref_read_mps(text, **opts)))
        assert comparable(digest(ours(text, **opts))) == want, label


@pytest.mark.parametrize("name", ["refinery", "small"])
def test_reference_data_files(name):
    raw = data(name)
    assert digest(ours(raw)) == FIX["files"][name]
    if _oracle.ref is not None:
        assert digest(ours(raw)) == digest(_oracle.ref_read_mps(raw))


def test_reader_from_a_file(tmp_path):
    p = tmp_path / "refinery.mps"
    p.write_bytes(data("refinery"))
    m = read_mps(str(p))
    assert (m.rows, m.cols, m.nnz, m.maximize) == (1380, 4368, 10888, True)
    with pytest.raises(AnukError) as e:
        read_mps(str(tmp_path / "absent.mps"))
    assert e.value.code == 7 and "cannot open" in e.value.message


def test_standard_form_of_refinery_matches_reference():
    lp = read_mps_text(data("refinery")).to_lp()
    assert lp_digest(lp.export()) == FIX["standard_form"]["refinery"]


def test_quadratic_file_is_not_an_lp():
    text = [c[1] for c in CASES if c[0] == "quadobj"][0]
    m = read_mps_text(text)
    assert m.hessian is not None
    with pytest.raises(ValueError):
        m.to_lp()


# ---- the writer ----------------------------------------------------------------

FIELDS = ("start", "index", "value", "objective", "row_lower", "row_upper", "col_lower",
          "col_upper", "col_type")


def assert_same_model(a: MpsModel, b: MpsModel, names=True):
    assert (a.rows, a.cols, a.maximize) == (b.rows, b.cols, b.maximize)
    assert np.float64(a.objective_offset).tobytes() == np.float64(b.objective_offset).tobytes()
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype and x.shape == y.shape and x.tobytes() == y.tobytes(), f
    if names:
        assert a.row_names == b.row_names and a.col_names == b.col_names


@pytest.mark.parametrize("name", ["refinery", "small"])
def test_round_trip_of_the_reference_files(tmp_path, name):
    m = read_mps_text(data(name))
    write_mps(str(tmp_path / "copy.mps"), m)
    back = read_mps(str(tmp_path / "copy.mps"))
    assert_same_model(m, back)
    assert back.warnings == [] and back.name == m.name


def test_round_trip_of_every_readable_corpus_model(tmp_path):
    done = 0
    for label, text, opts in CASES:
        r = ours(text, **opts)
        if not r["ok"] or r["hessian"] is not None:
            continue
        m = read_mps_text(text, **keywords(opts))
        if any(" " in n for n in m.row_names + m.col_names):
            continue  # names with spaces exist only in fixed format
        if np.any((m.row_lower == -np.inf) & (m.row_upper == np.inf)):
            continue  # a free row is written as an N row, which the reader drops
        ranged = np.isfinite(m.row_lower) & np.isfinite(m.row_upper) & (m.row_lower != m.row_upper)
        p = tmp_path / f"{label}.mps"
        write_mps(str(p), m)
        back = read_mps(str(p))
        if not ranged.any():
            assert_same_model(m, back)
        else:
            # a ranged row keeps both bounds when some range reproduces them
            assert np.array_equal(back.row_upper[~ranged], m.row_upper[~ranged]), label
            assert np.allclose(back.row_lower, m.row_lower) and np.allclose(back.row_upper, m.row_upper)
        done += 1
    assert done > 150


def test_round_trip_of_generated_models(tmp_path):
    base, cases = G.case_stack(4)
    models = [G.refinery_model(4, seed=3), G.random_lp(m=300, n=700, seed=9),
              G.case_model(base, cases[5])]
    for k, nm in enumerate(models):
        p = tmp_path / f"g{k}.mps"
        write_mps(str(p), nm)
        back = read_mps(str(p))
        a = nm.to_lp().export()
        b = back.to_lp().export()
        for f in ("start", "index", "value", "c", "q", "lower", "upper", "row_model", "row_sign"):
            assert a[f].tobytes() == b[f].tobytes(), (k, f)
        assert a["objective_offset"] == b["objective_offset"]


def test_ranged_rows_and_every_bound_kind_round_trip(tmp_path):
    rng = np.random.default_rng(17)
    m, n = 60, 40
    ri = rng.integers(0, m, 300)
    ci = rng.integers(0, n, 300)
    v = rng.normal(size=300)
    lo = rng.normal(size=m) * 100
    up = lo + rng.uniform(0.0, 50.0, m)
    kinds = rng.integers(0, 5, m)
    lo = np.where(kinds == 1, -np.inf, lo)
    up = np.where(kinds == 2, np.inf, up)
    up = np.where(kinds == 3, lo, up)
    lo = np.where(kinds == 4, -np.inf, lo)
    up = np.where(kinds == 4, np.inf, up)  # a free row: written as N, dropped on reading
    clo = rng.normal(size=n)
    cup = clo + rng.uniform(0.0, 5.0, n)
    ck = rng.integers(0, 6, n)
    clo = np.where(ck == 1, -np.inf, np.where(ck == 2, 0.0, clo))
    cup = np.where(ck == 3, np.inf, np.where(ck == 4, clo, cup))
    clo = np.where(ck == 5, -np.inf, clo)
    cup = np.where(ck == 5, np.inf, cup)
    types = rng.integers(0, 4, n).astype(np.int32)
    nm = G.NaturalModel(rows=m, cols=n, row_index=ri, col_index=ci, value=v,
                        objective=rng.normal(size=n), row_lower=lo, row_upper=up, col_lower=clo,
                        col_upper=cup, maximize=True, objective_offset=-3.25, name="mixed")
    p = tmp_path / "mixed.mps"
    write_mps(str(p), nm, col_type=types)
    back = read_mps(str(p))
    keep = kinds != 4
    assert back.rows == int(keep.sum()) and back.free_rows_dropped == int((~keep).sum())
    assert back.col_type.tolist() == types.tolist()
    assert back.maximize and back.objective_offset == -3.25
    assert np.array_equal(back.col_lower, clo) and np.array_equal(back.col_upper, cup)
    exact = (back.row_lower == lo[keep]) & (back.row_upper == up[keep])
    print(f"\n{int(exact.sum())} of {int(keep.sum())} kept rows have both bounds exactly")
    assert exact.all()


def test_writer_rejects_what_mps_cannot_state(tmp_path):
    nm = G.random_lp(m=5, n=8, seed=1)
    with pytest.raises(AnukError) as e:
        write_mps(str(tmp_path / "x.mps"), nm, row_names=[f"r {i}" for i in range(5)])
    assert e.value.code == 1 and "white space" in e.value.message
    with pytest.raises(AnukError):
        write_mps(str(tmp_path / "x.mps"), nm, col_names=["same"] * 8)
    with pytest.raises(AnukError):
        write_mps(str(tmp_path / "x.mps"), nm, col_names=["MARKER"] + [f"c{j}" for j in range(7)])
    with pytest.raises(AnukError) as e:
        write_mps(str(tmp_path / "no" / "such" / "dir.mps"), nm)
    assert e.value.code == 7


def test_writer_avoids_names_the_reader_would_misread(tmp_path):
    # a column named like the bound set name, a row named like the objective
    nm = G.random_lp(m=3, n=4, seed=2)
    nm.col_upper[:] = np.inf
    nm.col_lower[0] = -np.inf  # FR / MI bound lines: valueless, three tokens
    nm.col_upper[0] = np.inf
    p = tmp_path / "names.mps"
    write_mps(str(p), nm, row_names=["OBJ", "r1", "r2"], col_names=["BND", "x1", "x2", "x3"])
    back = read_mps(str(p))
    assert back.row_names == ["OBJ", "r1", "r2"] and back.col_names == ["BND", "x1", "x2", "x3"]
    assert back.col_lower[0] == -np.inf and back.col_upper[0] == np.inf
