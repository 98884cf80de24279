"""Writes tests/golden/pdhg_reference.json from the reference build.

Runs the reference's own solve_pdhg (CPU backend, shipped defaults) and its
simplex, from oracle/_ref/libanukulan_ref.so, on the known-answer LPs and on
the seeded refinery plans, and records what parity is judged against:
status, iterations, restarts, polish counts, objective, ||K|| and the
simplex optimum. Run in the container that has the reference sources:

    make -C oracle && python tests/golden/make_fixtures.py [--only SECTION]

SECTION is one of random, stack24, stack24_cases, random_1e-8, uneven512,
stack24_presolved.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from golden.known_lps import KNOWN, natural  # noqa: E402
from anukulan_amd import generators as G  # noqa: E402


def record(rlp, tol, **opts):
    s = rlp.solve(tolerance=tol, **opts)
    return {"status": s.status, "iterations": s.iterations, "restarts": s.restarts,
            "polish_attempts": s.polish_attempts, "polish_iterations": s.polish_iterations,
            "polished": s.polished, "objective": s.objective,
            "matrix_norm": s.matrix_norm_estimate, "final_step_size": s.final_step_size,
            "final_primal_weight": s.final_primal_weight}


# (rows, cols, tolerances, CPU threads). The threaded CPU backend gives the
# serial backend's bits at any thread count (threaded_backend.cpp), so the
# full-size LP is solved on all cores.
RANDOM = ((2000, 5000, ("1e-4", "1e-8"), 1), (10000, 25000, ("1e-4",), 1),
          (400000, 1000000, ("1e-4",), 8))


def random_section():
    """Seeded random LPs (generators.random_lp), up to the 1M-column LP."""
    out = {}
    for m, n, tols, threads in RANDOM:
        rlp = _oracle.RefLp.from_natural(G.random_lp(m=m, n=n))
        out[f"{m}x{n}"] = {"seed": 20261204, "rows": rlp.rows, "cols": rlp.cols, "nnz": rlp.nnz,
                           **{t: record(rlp, float(t), threads=threads) for t in tols}}
    return out


def stack_section(periods=24, threads=8):
    """The FP64 optimum of every case of the case stack by the reference's
    simplex (the error every batched and plain result is stated against)."""
    from concurrent.futures import ThreadPoolExecutor
    base, cases = G.case_stack(periods)
    lps = [_oracle.RefLp.from_natural(G.case_model(base, c)) for c in cases]
    with ThreadPoolExecutor(max_workers=threads) as pool:
        sx = list(pool.map(lambda lp: lp.simplex(), lps))
    return {"periods": periods, "cases": len(cases),
            "simplex_status": [int(r["status"]) for r in sx],
            "simplex_objective": [r["objective"] for r in sx]}


def stack_cases_section(periods=24, threads=8):
    """The reference's solve_pdhg on every case of the case stack at 1e-4 and
    1e-8 (status, iterations, objective, ...): the per-case parity bar of the
    batched and plain GPU solves."""
    from concurrent.futures import ThreadPoolExecutor
    base, cases = G.case_stack(periods)
    lps = [_oracle.RefLp.from_natural(G.case_model(base, c)) for c in cases]
    out = {"periods": periods, "cases": len(cases)}
    for tol in ("1e-4", "1e-8"):
        with ThreadPoolExecutor(max_workers=threads) as pool:
            out[tol] = list(pool.map(lambda lp: record(lp, float(tol)), lps))
    return out


def stack_presolved_section(periods=24, threads=8):
    """The reference's solve_pdhg at 1e-4 on every case of the case stack
    as this library's stack presolve reduces it (with the column
    substitutions, anukulan_amd.presolve): the per-case parity bar of the
    batched GPU solve of the reduced stack."""
    from concurrent.futures import ThreadPoolExecutor
    from anukulan_amd.presolve import presolve_stack
    base, cases = G.case_stack(periods)
    ps = presolve_stack(base, cases, slack_column_singletons=1)
    lps = [_oracle.RefLp.from_natural(m) for m in ps.models()]
    with ThreadPoolExecutor(max_workers=threads) as pool:
        runs = list(pool.map(lambda lp: record(lp, 1e-4), lps))
    return {"periods": periods, "cases": len(cases), "reduced": [ps.rows, ps.cols, int(ps.nnz)],
            "1e-4": runs}


def random_tight_section():
    """The random LP generator at 1e-8 (threaded CPU backend, serial bits):
    40k x 100k, which the reference solves, and the 1M-column LP, where the
    reference stops at its default time limit (the record says where)."""
    out = {}
    for m, n, threads in ((40000, 100000, 8), (400000, 1000000, 8)):
        rlp = _oracle.RefLp.from_natural(G.random_lp(m=m, n=n))
        out[f"{m}x{n}"] = {"seed": 20261204, "rows": rlp.rows, "cols": rlp.cols, "nnz": rlp.nnz,
                           "1e-8": record(rlp, 1e-8, threads=threads)}
        print(m, n, out[f"{m}x{n}"], flush=True)
    return out


def uneven_sample_section(threads=8):
    """The reference's solve_pdhg at 1e-4 on every 32nd LP of the 512-LP
    uneven stack (generators.uneven_stack): 16 LPs, refinery plans and
    random LPs, to hold the full-size stack's LPs to."""
    from concurrent.futures import ThreadPoolExecutor
    models = G.uneven_stack(512)
    picks = list(range(0, 512, 32))

    def one(i):
        rlp = _oracle.RefLp.from_natural(models[i])
        return {"index": i, "name": models[i].name, "rows": rlp.rows, "cols": rlp.cols,
                "nnz": rlp.nnz, "1e-4": record(rlp, 1e-4)}
    with ThreadPoolExecutor(max_workers=threads) as pool:
        return {"count": 512, "lps": list(pool.map(one, picks))}


SECTIONS = {"random": random_section, "stack24": stack_section, "uneven512": uneven_sample_section,
            "stack24_cases": stack_cases_section, "random_1e-8": random_tight_section,
            "stack24_presolved": stack_presolved_section}


def main():
    if _oracle.ref is None:
        sys.exit("oracle/_ref/libanukulan_ref.so is not built")
    path = os.path.join(HERE, "pdhg_reference.json")
    if len(sys.argv) == 3 and sys.argv[1] == "--only" and sys.argv[2] in SECTIONS:
        out = json.load(open(path))
        out[sys.argv[2]] = SECTIONS[sys.argv[2]]()
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        print("updated", path)
        return
    out = {"known": {}, "refinery": {}, "random": random_section(), "stack24": stack_section(),
           "stack24_cases": stack_cases_section(), "random_1e-8": random_tight_section(),
           "uneven512": uneven_sample_section(), "stack24_presolved": stack_presolved_section()}
    for name in KNOWN:
        rlp = _oracle.RefLp.from_natural(natural(name))
        out["known"][name] = {"1e-8": record(rlp, 1e-8), "1e-4": record(rlp, 1e-4),
                              "norm": rlp.matrix_norm(500, 1e-12)}
    for periods in (4, 12, 24):
        rlp = _oracle.RefLp.from_natural(G.refinery_model(periods))
        sx = rlp.simplex()
        out["refinery"][str(periods)] = {
            "seed": 20261203, "rows": rlp.rows, "cols": rlp.cols, "nnz": rlp.nnz,
            "1e-4": record(rlp, 1e-4), "1e-8": record(rlp, 1e-8),
            "simplex_status": sx["status"], "simplex_objective": sx["objective"]}
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
