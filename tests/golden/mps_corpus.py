"""MPS texts that exercise the reader, and the digest both readers are
compared by.

The hand-written cases cover each rule of the reference's reader
(src/anukulan/mps_reader.cpp): every section, row type and bound type, the
RANGES signs, both negative-UP conventions, markers, fixed format with names
that hold spaces, the format decision, Fortran exponents, DOS line ends, the
quadratic sections, and each error and warning. `fuzz_cases` adds seeded
variants: random models written in free or fixed format, then damaged (a
token dropped, a line repeated or moved, a value garbled, case changed).
"""
from __future__ import annotations

import hashlib

import numpy as np


def fixed(*fields):
    """One fixed-format line: fields at 0-based columns 1, 4, 14, 24, 39, 49."""
    cols = (1, 4, 14, 24, 39, 49)
    line = ""
    for f, c in zip(fields, cols):
        if f is None:
            continue
        line = line.ljust(c) + f
    return line


BASE = """NAME          SMALL
ROWS
 N  cost
 L  cap
 G  need
 E  flow
COLUMNS
    a         cost      3.5        cap       1.0
    a         need      2.0
    b         cost      -1.25      cap       4.0
    b         flow      1.0        need      0.5
    c         flow      -1.0       cost      0.0
RHS
    rhs       cap       40.0       need      6.0
    rhs       flow      0.0
ENDATA
"""

HAND = [
    ("base", BASE, {}),
    ("offset", BASE.replace("    rhs       flow      0.0", "    rhs       flow      0.0        cost      -12.5"), {}),
    ("objsense_header", BASE.replace("NAME          SMALL", "NAME          SMALL\nOBJSENSE MAXIMIZE"), {}),
    ("objsense_body", BASE.replace("ROWS", "OBJSENSE\n    MAX\nROWS"), {}),
    ("objsens_min", BASE.replace("ROWS", "OBJSENS\n  min\nROWS"), {}),
    ("ranges", BASE.replace("ENDATA", "RANGES\n    rng  cap  -5.0  need  3.0\n    rng  flow  2.5\nENDATA"), {}),
    ("ranges_eq_negative", BASE.replace("ENDATA", "RANGES\n    rng  flow  -2.0\nENDATA"), {}),
    ("ranges_eq_zero", BASE.replace("ENDATA", "RANGES\n    flow  0.0\nENDATA"), {}),
    ("ranges_objective_ignored", BASE.replace("ENDATA", "RANGES\n    r  cost  9\nENDATA"), {}),
    ("ranges_duplicate", BASE.replace("ENDATA", "RANGES\n    r  cap  1\n    r  cap  2\nENDATA"), {}),
    ("bounds_all", BASE.replace("ENDATA", """BOUNDS
 UP bnd       a         10.0
 LO bnd       a         -2.0
 FX bnd       b         1.5
 FR bnd       c
ENDATA"""), {}),
    ("bounds_mi_pl", BASE.replace("ENDATA", "BOUNDS\n MI bnd a\n PL bnd a\n UP bnd b 7\nENDATA"), {}),
    ("bounds_integer", BASE.replace("ENDATA", "BOUNDS\n BV bnd a\n LI bnd b -3\n UI bnd b 9\n SC bnd c 4\nENDATA"), {}),
    ("bounds_si", BASE.replace("ENDATA", "BOUNDS\n SI bnd c 12\nENDATA"), {}),
    ("bounds_no_set_name", BASE.replace("ENDATA", "BOUNDS\n UP a 10.0\n FR b\n MI c\nENDATA"), {}),
    ("bounds_three_token_bare_known_column", BASE.replace("ENDATA", "BOUNDS\n FR a b\nENDATA"), {}),
    ("bounds_three_token_bare_set", BASE.replace("ENDATA", "BOUNDS\n FR bnd b\nENDATA"), {}),
    ("bounds_new_column", BASE.replace("ENDATA", "BOUNDS\n UP bnd fresh 3\nENDATA"), {}),
    ("bounds_duplicate", BASE.replace("ENDATA", "BOUNDS\n UP bnd a 3\n UP bnd a 4\n LO bnd a 1\n FX bnd a 2\nENDATA"), {}),
    ("bounds_infinite", BASE.replace("ENDATA", "BOUNDS\n UP bnd a 1e30\n LO bnd b -1e31\n UP bnd c 2e29\nENDATA"), {}),
    ("bounds_infinite_threshold", BASE.replace("ENDATA", "BOUNDS\n UP bnd a 1e20\n LO bnd b -5e20\nENDATA"),
     {"infinity_threshold": 1e20}),
    ("negative_up_highs", BASE.replace("ENDATA", "BOUNDS\n UP bnd a -4\n MI bnd a\nENDATA"), {}),
    ("negative_up_invalid", BASE.replace("ENDATA", "BOUNDS\n UP bnd a -4\nENDATA"), {}),
    ("negative_up_cplex", BASE.replace("ENDATA", "BOUNDS\n UP bnd a -4\n UP bnd b 2\nENDATA"), {"negative_up": 1}),
    ("negative_up_cplex_integer", BASE.replace("ENDATA", "BOUNDS\n UI bnd a 5\n UP bnd b -1\nENDATA").replace(
        "    b         cost", "    M1 'MARKER' 'INTORG'\n    b         cost").replace(
        "    c         flow", "    M2 'MARKER' 'INTEND'\n    c         flow"), {"negative_up": 1}),
    ("markers", BASE.replace("    b         cost", "    MARKER                 'MARKER'                 'INTORG'\n    b         cost").replace(
        "    c         flow", "    MARKER                 'MARKER'                 'INTEND'\n    c         flow"), {}),
    ("marker_unrecognised", BASE.replace("    c         flow", "    MARKER  'MARKER'  'SOS1'\n    c         flow"), {}),
    ("marker_unquoted", BASE.replace("    b         cost", "    M MARKER INTORG\n    b         cost"), {}),
    ("second_n_row", BASE.replace(" L  cap", " N  spare\n L  cap").replace(
        "    a         need      2.0", "    a         need      2.0        spare     8.0"), {}),
    ("duplicate_row_name", BASE.replace(" E  flow", " E  flow\n G  cap"), {}),
    ("duplicate_rhs", BASE.replace("    rhs       flow      0.0", "    rhs       flow      0.0\n    rhs       cap       41.0"), {}),
    ("rhs_no_set_name", BASE.replace("    rhs       cap       40.0       need      6.0", "    cap  40.0  need  6.0"), {}),
    ("rhs_on_dropped_row", BASE.replace(" L  cap", " N  spare\n L  cap").replace(
        "    rhs       flow      0.0", "    rhs       flow      0.0        spare     5"), {}),
    ("fortran_exponent", BASE.replace("3.5", "0.35D1").replace("40.0", "4.0d+1"), {}),
    ("dos_line_ends", BASE.replace("\n", "\r\n"), {}),
    ("comments_and_blank", BASE.replace("COLUMNS", "* a comment\n\nCOLUMNS\n*another"), {}),
    ("no_final_newline", BASE.rstrip("\n"), {}),
    ("no_endata", BASE.replace("ENDATA\n", ""), {}),
    ("after_endata_ignored", BASE + "GARBAGE here\n", {}),
    ("tabs", BASE.replace("    a         cost", "\ta\tcost"), {}),
    ("lower_case_sections", BASE.replace("ROWS", "rows").replace("COLUMNS", "columns").replace("RHS\n", "rhs\n"), {}),
    ("no_n_row", "NAME X\nROWS\n L r\nCOLUMNS\n x r 1\nRHS\n rhs r 2\nENDATA\n", {}),
    ("empty", "", {}),
    ("only_name", "NAME ONLY\n", {}),
    ("name_body_ignored", "NAME\n    something here\nROWS\n N obj\nENDATA\n", {}),
    ("quadobj", BASE.replace("ENDATA", "QUADOBJ\n    a  a  2.0\n    a  b  -1.0\n    b  b  4.0\nENDATA"), {}),
    ("qsection_alias", BASE.replace("ENDATA", "QSECTION\n    a  b  1.0\n    c  c  3.0\nENDATA"), {}),
    ("qmatrix_full", BASE.replace("ENDATA", "QMATRIX\n    a  a  2.0\n    a  b  -1.0\n    b  a  -1.0\n    b  b  4.0\nENDATA"), {}),
    ("quadobj_both_triangles", BASE.replace("ENDATA", "QUADOBJ\n    a  b  1.0\n    b  a  1.0\nENDATA"), {}),
    ("qmatrix_asymmetric", BASE.replace("ENDATA", "QMATRIX\n    a  b  1.0\n    b  a  2.0\nENDATA"), {}),
    ("qmatrix_one_sided", BASE.replace("ENDATA", "QMATRIX\n    a  b  1.0\nENDATA"), {}),
    ("quad_new_column", BASE.replace("ENDATA", "QUADOBJ\n    z  z  1.0\nENDATA"), {}),
    ("duplicate_entries_summed", BASE.replace("    c         flow      -1.0       cost      0.0",
                                              "    c         flow      -1.0       flow      0.25\n    c  flow  0.125  cost  1\n    c  cost  2"), {}),
    ("entries_cancel", BASE.replace("    c         flow      -1.0       cost      0.0",
                                    "    c         flow      -1.0       flow      1.0"), {}),
    ("column_repeats_later", BASE.replace("RHS", "    a  flow  7.0\nRHS"), {}),
    # errors
    ("err_unknown_section", BASE.replace("RHS", "RHSX"), {}),
    ("err_unknown_row_type", BASE.replace(" G  need", " X  need"), {}),
    ("err_row_type_two_letters", BASE.replace(" G  need", " GE need"), {"format": 1}),
    ("err_rows_short", BASE.replace(" G  need", " G"), {"format": 1}),
    ("err_unknown_row_in_columns", BASE.replace("    a         need      2.0", "    a         nope      2.0"), {}),
    ("err_bad_value", BASE.replace("-1.25", "-1.2.5"), {}),
    ("err_nan_value", BASE.replace("-1.25", "nan"), {}),
    ("inf_value_accepted", BASE.replace("40.0", "inf"), {}),
    ("err_columns_even", BASE.replace("    a         need      2.0", "    a         need"), {}),
    ("err_rhs_unknown_row", BASE.replace("need      6.0", "nope      6.0"), {}),
    ("err_rhs_bad", BASE.replace("need      6.0", "need      6x"), {}),
    ("err_rhs_too_short", BASE.replace("    rhs       flow      0.0", "    rhs"), {}),
    ("err_ranges_unknown", BASE.replace("ENDATA", "RANGES\n    r  nope  1\nENDATA"), {}),
    ("err_ranges_short", BASE.replace("ENDATA", "RANGES\n    r\nENDATA"), {}),
    ("err_bound_type", BASE.replace("ENDATA", "BOUNDS\n XX bnd a 1\nENDATA"), {}),
    ("err_bound_needs_value", BASE.replace("ENDATA", "BOUNDS\n UP a\nENDATA"), {}),
    ("err_bound_too_many", BASE.replace("ENDATA", "BOUNDS\n FR a b c d\nENDATA"), {}),
    ("err_bound_bare_alone", BASE.replace("ENDATA", "BOUNDS\n FR\nENDATA"), {}),
    ("err_bound_bad_value", BASE.replace("ENDATA", "BOUNDS\n UP bnd a one\nENDATA"), {}),
    ("err_data_before_header", "  stray data\nNAME X\n", {}),
    ("err_empty_bound_interval", BASE.replace("ENDATA", "BOUNDS\n LO bnd a 5\n UP bnd a 4\nENDATA"), {}),
    ("err_quad_shape", BASE.replace("ENDATA", "QUADOBJ\n    a  b\nENDATA"), {}),
    ("err_quad_value", BASE.replace("ENDATA", "QUADOBJ\n    a  b  x\nENDATA"), {}),
    ("err_range_nan_row", BASE.replace("40.0", "inf").replace("ENDATA", "RANGES\n    r  cap  inf\nENDATA"), {}),
]

# Fixed format: names with spaces, fields in their columns.
FIXED = "\n".join([
    "NAME          FIXEDMODEL",
    "ROWS",
    fixed("N", "COST ROW"),
    fixed("L", "CAP 1"),
    fixed("G", "NEED 2"),
    "COLUMNS",
    fixed(None, "X ONE", "COST ROW", "1.5", "CAP 1", "2.0"),
    fixed(None, "X ONE", "NEED 2", "1.0"),
    fixed(None, "Y TWO", "CAP 1", "3.0", "COST ROW", "-2.0"),
    "RHS",
    fixed(None, "RHS SET", "CAP 1", "12.0", "NEED 2", "1.0"),
    "RANGES",
    fixed(None, "RNG", "CAP 1", "4.0"),
    "BOUNDS",
    fixed("UP", "BND SET", "X ONE", "8.0"),
    fixed("MI", "BND SET", "Y TWO"),
    "ENDATA", ""])
FIXED_PLAIN = "\n".join([
    "NAME          FIXEDPLAIN",
    "ROWS",
    fixed("N", "COST"),
    fixed("E", "BAL"),
    "COLUMNS",
    fixed(None, "X", "COST", "1.0", "BAL", "1.0"),
    fixed(None, "Y", "COST", "2.0", "BAL", "1.0"),
    "RHS",
    fixed(None, "RHS", "BAL", "3.0"),
    "BOUNDS",
    fixed("UP", "BND", "X", "2.5"),
    "ENDATA", ""])

HAND += [
    ("fixed_auto", FIXED, {}),
    ("fixed_forced", FIXED, {"format": 2}),
    ("fixed_text_forced_free", FIXED, {"format": 1}),
    ("fixed_plain_auto", FIXED_PLAIN, {}),
    ("fixed_plain_forced", FIXED_PLAIN, {"format": 2}),
    ("free_forced_fixed", BASE, {"format": 2}),
    ("fixed_value_spills", FIXED.replace(fixed(None, "RHS SET", "CAP 1", "12.0", "NEED 2", "1.0"),
                                         fixed(None, "RHS SET", "CAP 1", "12.0", "NEED 2", "1.0000000001"))
     , {}),
    ("fixed_short_line", FIXED.replace(fixed("MI", "BND SET", "Y TWO"), fixed("MI", "BND SET")), {}),
]

# More than 64 warnings: the cap and its marker line.
_many = "\n".join(f"    c  cap  {k}.0" for k in range(70))
HAND.append(("warning_cap", BASE.replace("ENDATA", "RHS\n" + _many + "\nENDATA"), {}))


def _random_model_text(rng, fixed_format):
    nr = int(rng.integers(1, 7))
    nc = int(rng.integers(1, 8))
    rows = [f"r{i}" + (" x" if fixed_format and rng.random() < 0.3 else "") for i in range(nr)]
    cols = [f"c{j}" + (" y" if fixed_format and rng.random() < 0.3 else "") for j in range(nc)]
    kinds = rng.choice(list("LGEN"), size=nr, p=[0.35, 0.3, 0.3, 0.05])

    def val():
        v = rng.choice([rng.normal() * 10.0 ** int(rng.integers(-3, 5)), float(rng.integers(-9, 10)), 0.0])
        # fixed-format values fit their 12-column field
        s = f"{float(v):.6g}" if fixed_format else repr(float(v))
        if rng.random() < 0.1:
            s = s.replace("e", "D")
        return s

    def line(*f):
        if fixed_format:
            return fixed(*f)
        return "    " + "  ".join(x for x in f if x is not None)

    out = ["NAME          R" + str(int(rng.integers(1000)))]
    if rng.random() < 0.3:
        out += ["OBJSENSE", "    MAX" if rng.random() < 0.5 else "    MIN"]
    out.append("ROWS")
    out.append(fixed("N", "obj") if fixed_format else " N  obj")
    for r, k in zip(rows, kinds):
        out.append(fixed(str(k), r) if fixed_format else f" {k}  {r}")
    out.append("COLUMNS")
    integer = False
    for c in cols:
        if rng.random() < 0.15:
            integer = not integer
            out.append(line(None, "MARKER", "'MARKER'", None, "'INTORG'" if integer else "'INTEND'"))
        targets = ["obj"] + list(rng.choice(rows, size=int(rng.integers(0, nr + 1)), replace=False))
        for t in targets:
            out.append(line(None, c, t, val()))
    out.append("RHS")
    for r in rows:
        if rng.random() < 0.6:
            out.append(line(None, "rhs", r, val()))
    if rng.random() < 0.3:
        out.append(line(None, "rhs", "obj", val()))
    if rng.random() < 0.4:
        out.append("RANGES")
        for r in rows:
            if rng.random() < 0.5:
                out.append(line(None, "rng", r, val()))
    if rng.random() < 0.7:
        out.append("BOUNDS")
        for c in cols:
            if rng.random() < 0.6:
                t = str(rng.choice(["UP", "LO", "FX", "FR", "MI", "PL", "BV", "LI", "UI", "SC", "SI"]))
                bare = t in ("FR", "MI", "PL", "BV")
                v = str(abs(float(rng.integers(0, 20)))) if rng.random() < 0.8 else val()
                out.append((fixed(t, "bnd", c) if bare else fixed(t, "bnd", c, v)) if fixed_format
                           else (f" {t} bnd {c}" if bare else f" {t} bnd {c} {v}"))
    if rng.random() < 0.15:
        out.append("QUADOBJ")
        for _ in range(int(rng.integers(1, 4))):
            a, b = rng.choice(cols, size=2)
            out.append(line(None, a, b, val()))
    out.append("ENDATA")
    return out


def _damage(lines, rng):
    lines = list(lines)
    for _ in range(int(rng.integers(1, 3))):
        if not lines:
            break
        k = int(rng.integers(len(lines)))
        op = int(rng.integers(7))
        if op == 0:
            toks = lines[k].split()
            if toks:
                del toks[int(rng.integers(len(toks)))]
                lines[k] = "    " + "  ".join(toks)
        elif op == 1:
            lines.insert(k, lines[k])
        elif op == 2:
            j = int(rng.integers(len(lines)))
            lines[k], lines[j] = lines[j], lines[k]
        elif op == 3:
            lines[k] = lines[k].replace("0", "O", 1)
        elif op == 4:
            lines[k] = lines[k].lower()
        elif op == 5:
            lines[k] = lines[k] + "  extra"
        else:
            lines[k] = lines[k].replace(" ", "\t", 2)
    return lines


def fuzz_cases(count=240, seed=20261207):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        fixed_format = rng.random() < 0.3
        lines = _random_model_text(rng, fixed_format)
        if rng.random() < 0.5:
            lines = _damage(lines, rng)
        text = "\n".join(lines) + ("\n" if rng.random() < 0.9 else "")
        if rng.random() < 0.1:
            text = text.replace("\n", "\r\n")
        opts = {}
        if rng.random() < 0.2:
            opts["negative_up"] = 1
        if rng.random() < 0.15:
            opts["format"] = int(rng.integers(0, 3))
        out.append((f"fuzz{i}", text, opts))
    return out


def all_cases():
    return HAND + fuzz_cases()


def _h(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:24]


def digest(r: dict) -> dict:
    """JSON-able summary of a reader result in the form tests/_oracle.py's
    ref_read_mps returns; floats by their exact bits."""
    d = {"ok": r["ok"], "error": r["error"], "warnings": list(r["warnings"])}
    if not r["ok"]:
        return d
    d.update(rows=int(r["rows"]), cols=int(r["cols"]), name=r["name"], maximize=bool(r["maximize"]),
             objective_offset=float(r["objective_offset"]).hex(),
             free_rows_dropped=int(r["free_rows_dropped"]),
             objective_nonzeros=int(r["objective_nonzeros"]), fixed_format=bool(r["fixed_format"]),
             row_names=_h(np.frombuffer("\n".join(r["row_names"]).encode(), np.uint8)),
             col_names=_h(np.frombuffer("\n".join(r["col_names"]).encode(), np.uint8)))
    for f, t in (("start", np.int32), ("index", np.int32), ("value", np.float64),
                 ("objective", np.float64), ("row_lower", np.float64), ("row_upper", np.float64),
                 ("col_lower", np.float64), ("col_upper", np.float64), ("col_type", np.int32)):
        d[f] = _h(np.asarray(r[f], t))
    q = r.get("hessian")
    d["hessian"] = None if q is None else [_h(np.asarray(q[0], np.int32)),
                                           _h(np.asarray(q[1], np.int32)),
                                           _h(np.asarray(q[2], np.float64))]
    return d


def lp_digest(e: dict) -> dict:
    """Digest of a standard form as LP.export / RefLp.export return it."""
    out = {k: _h(e[k]) for k in ("start", "index", "value", "c", "q", "lower", "upper",
                                 "row_model", "row_sign")}
    out.update(num_equalities=int(e["num_equalities"]),
               objective_scale=float(e["objective_scale"]).hex(),
               objective_offset=float(e["objective_offset"]).hex())
    return out
