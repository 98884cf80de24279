"""Host-side checks that need no GPU.

The C ABI library must load and export every entry point its header
declares; the host LP layout it builds (standard form, scaling, residual)
must be the reference's, bit for bit; its options must default to the
reference's PdhgOptions; errors must come back as status codes; and the
synthetic workloads must be deterministic and of the stated shape.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from anukulan_amd import AnukError, LP, Context, default_options, generators as G
from anukulan_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "anukulan_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(anuk_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 38, names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert sorted(set(names) - set(_lib.SIGNATURES)) == []


def test_library_is_the_in_tree_build():
    assert os.path.dirname(os.path.abspath(_lib.LIB_PATH)) == os.path.join(ROOT, "anukulan_amd")


def test_context_without_a_device_is_a_status_not_a_crash():
    try:
        c = Context(0)
    except AnukError as e:
        assert e.code == 2
        assert "device" in str(e).lower() or "hip" in str(e).lower()
        return
    c.close()
    pytest.skip("a GPU is visible here")


def small_models():
    from golden.known_lps import KNOWN, natural
    out = [natural(n) for n in KNOWN]
    out.append(G.refinery_model(3))
    base, cases = G.case_stack(4)
    out += [G.case_model(base, c) for c in cases[:5]]
    return out


def test_standard_form_is_the_reference(ref):
    for m in small_models():
        a = m.to_lp().export()
        b = ref.RefLp.from_natural(m).export()
        for k in ("start", "index", "value", "c", "q", "lower", "upper", "row_model", "row_sign"):
            assert np.array_equal(a[k], b[k]), (m.name, k)
        assert a["num_equalities"] == b["num_equalities"], m.name
        assert (a["objective_scale"], a["objective_offset"]) == (
            b["objective_scale"], b["objective_offset"]), m.name


def test_scaling_is_the_reference(ref):
    for m in small_models()[-6:]:
        lp, rl = m.to_lp(), ref.RefLp.from_natural(m)
        for ruiz, pc in ((10, True), (3, False), (0, True)):
            s, rs, cs = lp.scaled(ruiz, pc)
            t, trs, tcs = rl.scaled(ruiz, pc)
            assert np.array_equal(rs, trs) and np.array_equal(cs, tcs), m.name
            a, b = s.export(), t.export()
            for k in ("start", "index", "value", "c", "q", "lower", "upper"):
                assert np.array_equal(a[k], b[k]), (m.name, k)


def test_evaluate_residual_is_the_reference(ref):
    rng = np.random.default_rng(11)
    for m in small_models()[-3:]:
        lp, rl = m.to_lp(), ref.RefLp.from_natural(m)
        x, y = rng.uniform(0, 3, lp.cols), rng.uniform(-1, 1, lp.rows)
        assert lp.evaluate_residual(x, y) == rl.evaluate_residual(x, y)


def test_default_options_are_the_references(ref):
    mine = default_options()
    theirs = _lib.PdhgOptions()
    ref.ref.ref_pdhg_default_options(ctypes.byref(theirs))
    for name, kind in _lib.PdhgOptions._fields_:
        a, b = getattr(mine, name), getattr(theirs, name)
        if issubclass(kind, ctypes._Pointer) or kind is ctypes.c_void_p:
            assert not a and not b, name  # warm starts and stop flag: absent
        elif isinstance(a, float) and math.isnan(a):
            assert math.isnan(b), name
        else:
            assert a == b, name


def test_bad_input_is_an_error_status():
    # start not nondecreasing
    with pytest.raises(AnukError) as e:
        LP.from_standard(2, 2, 0, [0, 2, 1], [0, 1], [1.0, 1.0], [1, 1], [1, 1], [0, 0],
                         [np.inf, np.inf])
    assert e.value.code == 1
    # column index out of range
    with pytest.raises(AnukError) as e:
        LP.from_standard(1, 2, 0, [0, 1], [5], [1.0], [1, 1], [1], [0, 0], [np.inf, np.inf])
    assert e.value.code == 1
    # triplet outside the matrix
    with pytest.raises(AnukError) as e:
        LP.from_model(1, 1, [3], [0], [1.0], [1.0], [0.0], [1.0], [0.0], [1.0])
    assert e.value.code == 1
    assert "row" in lib_error().lower() or "range" in lib_error().lower() or lib_error()


def lib_error():
    return _lib.lib.anuk_last_error(None).decode()


def test_refinery_generator_is_deterministic_and_shaped():
    a, b = G.refinery_model(12), G.refinery_model(12)
    for k in ("row_index", "col_index", "value", "objective", "row_lower", "row_upper",
              "col_lower", "col_upper"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    lp = a.to_lp()
    assert (lp.rows, lp.cols, lp.nnz) == (1248, 3228, 10786)
    assert G.refinery_model(12, seed=1).objective.tolist() != a.objective.tolist()
    lp24 = G.refinery_model(24).to_lp()
    assert (lp24.rows, lp24.cols, lp24.nnz) == (2496, 6456, 21586)


def test_case_stack_moves_one_datum_each():
    base, cases = G.case_stack(6)
    assert len(cases) == 61 and cases[0].name == "base"
    assert len({c.name for c in cases}) == 61
    for c in cases[1:]:
        # one datum moves: one array changes, in at most one entry per period
        changed = {f: int(np.count_nonzero(getattr(c, f) != getattr(base, f)))
                   for f in ("objective", "row_lower", "row_upper", "col_lower", "col_upper")}
        moved = [f for f, k in changed.items() if k]
        assert len(moved) == 1 and 1 <= changed[moved[0]] <= 6, (c.name, changed)
    # K is shared by every case
    k0 = G.case_model(base, cases[0]).to_lp().export()
    k7 = G.case_model(base, cases[7]).to_lp().export()
    for k in ("start", "index", "value"):
        assert np.array_equal(k0[k], k7[k])


def test_random_lp_generator_shape_small():
    m = G.random_lp(m=2000, n=5000, seed=3)
    assert m.rows == 2000 and m.cols == 5000
    nnz = m.value.size
    assert 6.0 * 5000 <= nnz <= 10.0 * 5000
    assert np.all(np.abs(m.value) >= 0.05)
    assert np.array_equal(G.random_lp(m=2000, n=5000, seed=3).value, m.value)
    lp = m.to_lp()
    assert lp.rows == 2000 and lp.num_equalities == 1000
