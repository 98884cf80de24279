"""Writes the MPS fixtures from the reference build and the reference's data.

* tests/golden/refinery.mps.gz, small.mps.gz: the reference's
  data/refinery/refinery.mps and small.mps (inputs, gzip-compressed);
* tests/golden/mps_reference.json: what the reference's reader
  (read_mps_stream, via oracle/_ref) makes of every text in
  tests/golden/mps_corpus.py and of the two files, as digests; the standard
  form the reference builds from refinery.mps (to_standard_form); and the
  reference's solve_pdhg on it (CPU backend, 8 threads, which gives the
  serial backend's bits) at 1e-4 and 1e-6, and on the LP the reference's
  own presolve makes of it.

    make -C oracle && python tests/golden/make_mps_fixtures.py
"""
import gzip
import json
import os
import sys


HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from golden.mps_corpus import all_cases, digest, lp_digest  # noqa: E402

DATA = "/root/reference/data/refinery"


def main():
    if _oracle.ref is None:
        sys.exit("oracle/_ref/libanukulan_ref.so is not built")
    out = {"cases": {}, "files": {}}
    for label, text, opts in all_cases():
        out["cases"][label] = digest(_oracle.ref_read_mps(text, **opts))
    for name in ("refinery", "small"):
        raw = open(os.path.join(DATA, name + ".mps"), "rb").read()
        with open(os.path.join(HERE, name + ".mps.gz"), "wb") as f:
            f.write(gzip.compress(raw, compresslevel=9, mtime=0))
        out["files"][name] = digest(_oracle.ref_read_mps(raw))
    raw = open(os.path.join(DATA, "refinery.mps"), "rb").read()
    rlp = _oracle.RefLp(_oracle.ref.ref_lp_from_mps(raw))
    out["standard_form"] = {"refinery": lp_digest(rlp.export())}
    out["pdhg"] = {"refinery": {}}
    for tol in ("1e-4", "1e-6"):
        s = rlp.solve(threads=8, tolerance=float(tol), max_iterations=2_000_000,
                      time_limit_seconds=1e9)
        out["pdhg"]["refinery"][tol] = {"status": s.status, "iterations": s.iterations,
                                        "objective": s.objective, "restarts": s.restarts,
                                        "polish_attempts": s.polish_attempts}
    # The optimum by the reference's interior point method with presolve
    # (bench/results/ipm_refinery.txt:19).
    out["pdhg"]["refinery"]["optimum"] = 2.426062266887e+10
    # The reference's solve_pdhg on the LP its own presolve (default
    # PresolveOptions) makes of refinery.mps: the configuration the
    # reference's refinery figure is quoted with (docs/RESULTS.md:1834-1838).
    pre = _oracle.RefPresolve(raw)
    plp = pre.lp()
    out["pdhg"]["refinery_presolved"] = {"shape": [plp.rows, plp.cols, plp.nnz],
                                         "rows_removed": pre.counts["rows_removed"],
                                         "cols_removed": pre.counts["cols_removed"]}
    for tol in ("1e-4", "1e-6"):
        s = plp.solve(threads=8, tolerance=float(tol), max_iterations=2_000_000,
                      time_limit_seconds=1e9)
        out["pdhg"]["refinery_presolved"][tol] = {
            "status": s.status, "iterations": s.iterations, "objective": s.objective,
            "restarts": s.restarts, "polish_attempts": s.polish_attempts}
    path = os.path.join(HERE, "mps_reference.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path, len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
