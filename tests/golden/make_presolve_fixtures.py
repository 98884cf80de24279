"""Writes tests/golden/presolve_reference.json from the reference build.

What the reference's presolve (presolve.cpp:1215-1317, via oracle/_ref) and
its PostsolveStack make of every model of tests/golden/presolve_corpus.py
and of the reference's data files refinery.mps and small.mps: status,
message, counts, whether the dual inverts exactly, a digest of the reduced
model, and digests of the primal and dual postsolve of seeded vectors. Also
the reference's solve_pdhg on the presolved refinery.mps (CPU backend, 8
threads, which gives the serial backend's bits) at 1e-4 and 1e-6.

    make -C oracle && python tests/golden/make_presolve_fixtures.py
"""
import gzip
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from golden.mps_corpus import digest  # noqa: E402
from golden.presolve_corpus import cases  # noqa: E402

READER_ONLY = ("objective_nonzeros", "free_rows_dropped", "fixed_format", "warnings")


def h(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()[:24]


def vectors(seed, rows, cols, original_cols):
    """The seeded vectors each case's postsolve is checked on."""
    g = np.random.default_rng(seed)
    return (g.normal(size=cols), g.normal(size=rows), g.normal(size=cols),
            g.normal(size=original_cols))


def record(R, seed):
    out = {"status": R.status, "message": R.message, "dual_exact": R.dual_exact,
           "counts": {k: (float(v).hex() if isinstance(v, float) else int(v))
                      for k, v in R.counts.items()}}
    if R.status != "reduced":
        return out
    red = digest(R.reduced)
    for k in READER_ONLY:
        red.pop(k, None)
    out["reduced"] = red
    rows, cols = R.reduced["rows"], R.reduced["cols"]
    original_cols = cols + R.counts["cols_removed"]
    original_rows = rows + R.counts["rows_removed"]
    rx, ry, rc, oc = vectors(seed, rows, cols, original_cols)
    out["primal"] = h(R.postsolve_primal(rx))
    out["dual"] = h(R.postsolve_dual(ry, rc, oc, original_rows))
    return out


def main():
    if _oracle.ref is None:
        sys.exit("oracle/_ref/libanukulan_ref.so is not built")
    out = {"cases": {}, "files": {}, "pdhg": {}}
    for k, (label, text, opts) in enumerate(cases()):
        out["cases"][label] = record(_oracle.RefPresolve(text, **opts), 1000 + k)
    for k, name in enumerate(("refinery", "small")):
        raw = gzip.decompress(open(os.path.join(HERE, name + ".mps.gz"), "rb").read())
        out["files"][name] = record(_oracle.RefPresolve(raw), 5000 + k)
    raw = gzip.decompress(open(os.path.join(HERE, "refinery.mps.gz"), "rb").read())
    lp = _oracle.RefPresolve(raw).lp()
    out["pdhg"]["refinery_presolved"] = {"rows": lp.rows, "cols": lp.cols, "nnz": lp.nnz}
    for tol in ("1e-4", "1e-6"):
        s = lp.solve(threads=8, tolerance=float(tol), max_iterations=2_000_000,
                     time_limit_seconds=1e9)
        out["pdhg"]["refinery_presolved"][tol] = {
            "status": s.status, "iterations": s.iterations, "objective": s.objective,
            "restarts": s.restarts, "polish_attempts": s.polish_attempts}
    path = os.path.join(HERE, "presolve_reference.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path, len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
