#!/usr/bin/env python3
"""PDHG throughput on one MI355X: the 61-case refinery stack to 1e-4 rel KKT.

One step = solving every case of the stack (the 24-period base plan and its
60 one-datum moves) to relative KKT 1e-4 with the shipped PDHG defaults,
inputs built on the host beforehand. `value` is cases solved per second over
the timed steps. The dominant kernel's roofline comes from a separate
profiled pass over the same cases (HIP events around every launch on the
library's stream; graphs off so each launch is visible), and the CPU
baseline is the reference's own solve_pdhg (oracle/_ref, one thread) on a
bounded sample of the same cases.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cases/s to 1e-4 rel KKT, batched vs plain PDHG on 1 MI355X; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(name, n, m, nnz):
    """Algorithmic HBM bytes of one launch: matrix (8 B value + 4 B index per
    nonzero, 4 B per row start), the gathered vector once, and the dense
    vectors the epilogue reads and writes (8 B each)."""
    if name == "primal_step":        # K^T y, then x, c, l, u, anchor -> x', x_bar
        return 12 * nnz + 4 * (n + 1) + 8 * m + 56 * n
    if name == "primal_step_check":  # ... and dx
        return 12 * nnz + 4 * (n + 1) + 8 * m + 64 * n
    if name == "dual_step":          # K x_bar, then y, q, k_x, anchors -> y', k'
        return 12 * nnz + 4 * (m + 1) + 8 * n + 56 * m
    if name == "dual_step_check":    # ... and dy, k_dx
        return 12 * nnz + 4 * (m + 1) + 8 * n + 72 * m
    if name == "spmv":               # plain K x or K^T y
        return None
    return None


def parse_profile(text):
    rows = {}
    for line in text.splitlines()[1:]:
        parts = line.split()
        if len(parts) < 6:
            continue
        name = parts[0]
        rows[name] = {"calls": int(parts[1]), "timed": int(parts[2]), "mean_us": float(parts[3]),
                      "total_ms": float(parts[4])}
    return rows


def build_stack(periods):
    from anukulan_amd import generators as G
    base, cases = G.case_stack(periods)
    lps = [G.case_model(base, c).to_lp() for c in cases]
    return base, cases, lps


def cpu_baseline(base, cases, tol, budget_s):
    """The reference build's solve_pdhg, one thread, on cases in stack order
    until `budget_s` of CPU time is spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import _oracle
    except Exception as e:  # the checker is test infrastructure; report its absence
        return {"value": None, "unit": "cases/s", "cores": 1, "kind": "reference",
                "sample": f"unavailable: {e}"}
    if _oracle.ref is None:
        return {"value": None, "unit": "cases/s", "cores": 1, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    from anukulan_amd import generators as G
    done, iters, t0 = 0, 0, time.perf_counter()
    while done < len(cases) and time.perf_counter() - t0 < budget_s:
        rlp = _oracle.RefLp.from_natural(G.case_model(base, cases[done]))
        s0 = time.perf_counter()
        r = rlp.solve(threads=1, tolerance=tol)
        done += 1
        iters += r.iterations
        if done == 1:
            t0 = s0  # time solves only, not model construction
    elapsed = time.perf_counter() - t0
    return {"value": done / elapsed, "unit": "cases/s", "cores": 1, "kind": "reference",
            "sample": f"first {done} of {len(cases)} cases, reference solve_pdhg on the CPU "
                      f"backend, {iters} iterations in {elapsed:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--periods", type=int, default=24)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the case stack shards by case across GPUs, which is not built here)",
              file=sys.stderr)
        sys.exit(2)

    from anukulan_amd import Context, solve_pdhg

    base, cases, lps = build_stack(args.periods)
    lp0 = lps[0]
    n, m, nnz = lp0.cols, lp0.rows, lp0.nnz
    ctx = Context(0)

    def one_step():
        res = []
        for lp in lps:
            res.append(solve_pdhg(ctx, lp, tolerance=args.tol))
        return res

    for _ in range(args.warmup):
        one_step()
    ctx.synchronize()
    t0 = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results = one_step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    solved = sum(1 for r in results if r.status == "optimal")
    total_iters = sum(r.iterations for r in results)
    value = len(lps) * args.steps / elapsed

    roofline = None
    kernels = {}
    if not args.no_profile:
        ctx.set_profiling(True)
        for lp in lps:
            solve_pdhg(ctx, lp, tolerance=args.tol, use_graphs=0)
        kernels = parse_profile(ctx.profile_report())
        ctx.set_profiling(False)
        ranked = sorted(((k, v) for k, v in kernels.items() if kernel_bytes(k, n, m, nnz)),
                        key=lambda kv: -kv[1]["total_ms"])
        if ranked:
            name, k = ranked[0]
            b = kernel_bytes(name, n, m, nnz)
            achieved = b / (k["mean_us"] * 1e-6) / 1e9
            roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                        "bytes_per_launch": b, "mean_launch_us": k["mean_us"],
                        "launches_timed": k["timed"]}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(base, cases, args.tol, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "cases/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"refinery case stack: {args.periods}-period base plan + 60 "
                               f"what-if cases, each solved to rel KKT {args.tol:g}",
                   "solver": "plain PDHG, one case at a time (fused kernels, graph per check period)",
                   "rows": m, "cols": n, "nnz": nnz, "cases": len(lps),
                   "cases_optimal": solved, "iterations_per_step": total_iters,
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
        "kernels": {k: {"mean_us": v["mean_us"], "calls": v["calls"]} for k, v in kernels.items()},
    }
    if cpu and cpu.get("value"):
        out["speedup_vs_cpu_baseline"] = round(value / cpu["value"], 2)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
