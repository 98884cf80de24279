#!/usr/bin/env python3
"""PDHG throughput on one MI355X: the 61-case refinery stack to 1e-4 rel KKT.

One step = solving every case of the stack (the 24-period base plan and its
60 one-datum moves) to relative KKT 1e-4 with the shipped PDHG defaults,
inputs built on the host beforehand. `value` is cases per second of the
batched solver (all cases in one set of launches, sharing K) over the timed
steps. Beside it: the same stack solved one case at a time by the plain GPU
solver (`plain`), a profiled pass that times every kernel launch with HIP
events the dispatch records (hipExtLaunchKernel) and divides each launch's
algorithmic bytes by its time (`roofline`, dominant kernel), the
1M-column random LP's SpMV bandwidth and plain solve (`c1`), the uneven
512-LP stack solved in one set of launches (`c3`: LPs/s, p50/p99
time-to-solve), a 64-node branch-and-bound frontier's relaxations solved
in one batch (`nodes`: nodes/s against plain and reference solves), the
stack finished exactly (`finish`: batched PDHG to a seed tolerance, then
crossover and simplex per case on the host cores, against the reference
simplex), and the
reference's own solve_pdhg (oracle/_ref, one thread) on a bounded sample of
the cases (`cpu_baseline`), and the same on every case, one case per thread
over the host's cores (`cpu_baseline_all_cores`).

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


def parse_profile(text):
    rows = {}
    for line in text.splitlines()[1:]:
        parts = line.split()
        if len(parts) < 7:
            continue
        rows[parts[0]] = {"calls": int(parts[1]), "timed": int(parts[2]),
                          "mean_us": float(parts[3]), "total_ms": float(parts[4]),
                          "gbs": float(parts[6])}
    return rows


def host_info():
    """What the CPU figures ran on: cores the OS reports, cores this process
    may use, the CPU model, and the pool size used (the usable cores, capped
    at 16: a one-GPU box's share of its host)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model,
            "pool_threads": min(16, usable),
            "pool_cap_reason": "a one-GPU box's share of the host is 16 cores"}


def ref_oracle():
    """The reference build (test infrastructure) for the CPU-baseline legs,
    or None."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import _oracle
    except Exception:
        return None
    return _oracle if _oracle.ref is not None else None


def c0_section(ctx, args):
    """The 12-period refinery plan, plain FP64 PDHG at rel KKT 1e-4 and 1e-8:
    iterations, time to solve and iterations/s on the GPU (median of
    args.c0_reps solves after one warm-up), the objective's error against
    the reference simplex's FP64 optimum, and beside it the reference's
    solve_pdhg on the host at t=1, t=default_thread_count() (threaded
    backend) and t=the pool size."""
    from anukulan_amd import generators as G, solve_pdhg
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "pdhg_reference.json")))
    fx = fx["refinery"]["12"]
    model = G.refinery_model(12)
    lp = model.to_lp()
    opt = fx["simplex_objective"]
    out = {"workload": f"12-period refinery plan ({lp.rows} x {lp.cols}, {lp.nnz} nnz), plain "
                       f"FP64 PDHG", "simplex_optimum": opt, "by_tolerance": {}}
    ref = None if args.no_cpu_baseline else ref_oracle()
    hi = host_info()
    for tol in ("1e-4", "1e-8"):
        solve_pdhg(ctx, lp, tolerance=float(tol))
        runs = []
        for _ in range(args.c0_reps):
            t0 = time.perf_counter()
            r = solve_pdhg(ctx, lp, tolerance=float(tol))
            runs.append((time.perf_counter() - t0, r))
        runs.sort(key=lambda t: t[0])
        sec, r = runs[len(runs) // 2]
        rec = {"status": r.status, "iterations": r.iterations,
               "reference_iterations": fx[tol]["iterations"],
               "seconds": round(sec, 4), "lps_per_s": round(1.0 / sec, 2),
               "iterations_per_s": round(r.iterations / sec, 1),
               "objective": r.objective,
               "rel_error_vs_simplex": abs(r.objective - opt) / max(1.0, abs(opt)),
               "rel_diff_vs_reference_objective":
                   abs(r.objective - fx[tol]["objective"]) / max(1.0, abs(fx[tol]["objective"]))}
        if ref is not None:
            R = ref.RefLp.from_natural(model)
            host = {}
            for label, t in (("t1", 1), ("t_default", min(16, ref.ref_default_thread_count())),
                             ("t_pool", hi["pool_threads"])):
                t0 = time.perf_counter()
                w = R.solve(threads=t, tolerance=float(tol))
                hs = time.perf_counter() - t0
                host[label] = {"threads": t, "seconds": round(hs, 4),
                               "iterations": w.iterations,
                               "iterations_per_s": round(w.iterations / hs, 1)}
            rec["cpu_reference"] = host
            rec["gpu_vs_cpu_t1"] = round(host["t1"]["seconds"] / sec, 2)
        out["by_tolerance"][tol] = rec
    return out


def build_stack(periods, count):
    from anukulan_amd import generators as G
    base, cases = G.case_stack(periods)
    cases = cases[:count]
    return base, cases, [G.case_model(base, c).to_lp() for c in cases]


def cpu_baseline(base, cases, tol, budget_s):
    """The reference build's solve_pdhg, one thread, on cases in stack order
    until `budget_s` of solve time is spent."""
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
    done, iters, spent = 0, 0, 0.0
    while done < len(cases) and spent < budget_s:
        rlp = _oracle.RefLp.from_natural(G.case_model(base, cases[done]))
        t0 = time.perf_counter()
        r = rlp.solve(threads=1, tolerance=tol)
        spent += time.perf_counter() - t0
        done += 1
        iters += r.iterations
    return {"value": round(done / spent, 4), "unit": "cases/s", "cores": 1, "kind": "reference",
            "sample": f"first {done} of {len(cases)} cases solved by the reference's solve_pdhg "
                      f"(CPU backend, one thread): {iters} iterations in {spent:.1f} s"}


def cpu_baseline_parallel(base, cases, tol, threads):
    """Every case solved by the reference build's solve_pdhg, one case per
    thread over `threads` threads (the reference's CPU backend keeps no shared
    mutable state, and the ctypes calls release the GIL). The LPs are built
    before the clock starts; the time is the wall clock of the whole pool."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import _oracle
    except Exception as e:
        return {"value": None, "unit": "cases/s", "cores": threads, "kind": "reference",
                "sample": f"unavailable: {e}"}
    if _oracle.ref is None:
        return {"value": None, "unit": "cases/s", "cores": threads, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    from anukulan_amd import generators as G
    lps = [_oracle.RefLp.from_natural(G.case_model(base, c)) for c in cases]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        rs = list(pool.map(lambda lp: lp.solve(threads=1, tolerance=tol), lps))
    wall = time.perf_counter() - t0
    iters = sum(r.iterations for r in rs)
    return {"value": round(len(lps) / wall, 4), "unit": "cases/s", "cores": threads,
            "kind": "reference",
            "sample": f"all {len(lps)} cases solved by the reference's solve_pdhg, one case per "
                      f"thread on {threads} threads: {iters} iterations in {wall:.1f} s wall"}


def cpu_baseline_threaded(base, cases, tol, budget_s):
    """The reference build's solve_pdhg on its threaded CPU backend at
    t = default_thread_count() (hw/2, threading.cpp:162; capped at the pool
    size), one case at a time in stack order until `budget_s` of solve time
    is spent."""
    ref = ref_oracle()
    if ref is None:
        return None
    from anukulan_amd import generators as G
    t = min(host_info()["pool_threads"], ref.ref_default_thread_count())
    done, iters, spent = 0, 0, 0.0
    while done < len(cases) and spent < budget_s:
        rlp = ref.RefLp.from_natural(G.case_model(base, cases[done]))
        t0 = time.perf_counter()
        r = rlp.solve(threads=t, tolerance=tol)
        spent += time.perf_counter() - t0
        done += 1
        iters += r.iterations
    return {"value": round(done / spent, 4), "unit": "cases/s", "cores": t, "kind": "reference",
            "sample": f"first {done} of {len(cases)} cases solved one at a time by the reference's "
                      f"solve_pdhg on its threaded backend with {t} threads: {iters} iterations "
                      f"in {spent:.1f} s"}


def parse_tuning(text):
    """Lines of Context.tuning_report(): kernel, rows, nnz, the mean launch
    time of each shape (batched steps: w<waves per workgroup>r<rows per
    wave>, "s" for the fixed-width row copy; stack products: plan or
    balanced block order, "+nt" for non-temporal loads) and the shape
    chosen."""
    out = []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        rec = {"kernel": parts[0], "us_by_shape": {}}
        for p in parts[1:]:
            k, v = p.split("=")
            if k.startswith("us@"):
                rec["us_by_shape"][k[3:]] = float(v)
            elif k == "chosen":
                rec[k] = v
            else:
                rec[k] = int(v)
        out.append(rec)
    return out


def simplex_optima(periods, count):
    """The reference simplex's FP64 optimum of each case of the stack, as
    committed in tests/golden/pdhg_reference.json (stack24), or None."""
    path = os.path.join(ROOT, "tests", "golden", "pdhg_reference.json")
    if not os.path.exists(path):
        return None
    import json as _json
    st = _json.load(open(path)).get(f"stack{periods}")
    if st is None or len(st["simplex_objective"]) < count:
        return None
    return st["simplex_objective"][:count]


def error_vs_optimum(results, optima):
    """Relative objective error of each result against its FP64 optimum:
    |obj - opt| / max(1, |opt|); max and median over the cases."""
    if optima is None:
        return None
    errs = sorted(abs(r.objective - o) / max(1.0, abs(o)) for r, o in zip(results, optima))
    return {"max_rel": errs[-1], "median_rel": errs[len(errs) // 2],
            "optimum": "reference simplex (tests/golden/pdhg_reference.json stack24)"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/pmc_traffic.txt), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#") and parts[0] == kernel:
            return float(parts[1])
    return None


def c1_section(ctx, args):
    """The 1M-column random LP (m=400k, n=1M, ~8M nnz): SpMV bandwidth of
    K x and K'y on the device, then a plain solve to rel KKT 1e-4 and a
    profiled run of its fused steps. Bytes per launch are algorithmic."""
    import numpy as np
    from anukulan_amd import Matrix, generators as G, solve_pdhg

    t0 = time.perf_counter()
    lp = G.random_lp(m=args.c1_rows, n=args.c1_cols).to_lp()
    e = lp.export()
    k = Matrix(lp.rows, lp.cols, e["start"], e["index"], e["value"])
    kt = k.transpose()
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(5)
    x, y = ctx.vector(rng.uniform(-1, 1, lp.cols)), ctx.allocate(lp.rows)
    yv, xv = ctx.vector(rng.uniform(-1, 1, lp.rows)), ctx.allocate(lp.cols)
    ctx.prepare(k)
    ctx.prepare(kt)
    ctx.multiply(k, x, y)
    ctx.multiply_transpose(kt, yv, xv)
    ctx.synchronize()
    spmv = {}
    timed_products = () if args.no_profile else (
        ("K x", lambda: ctx.multiply(k, x, y)), ("K'y", lambda: ctx.multiply_transpose(kt, yv, xv)))
    for name, fn in timed_products:
        ctx.set_profiling(True)
        for _ in range(args.c1_reps):
            fn()
        row = parse_profile(ctx.profile_report())["spmv"]
        ctx.set_profiling(False)
        spmv[name] = {"mean_us": row["mean_us"], "achieved_GBs": round(row["gbs"], 1),
                      "frac": round(row["gbs"] / HBM_PEAK_GBS, 4), "launches": row["timed"]}
    # Control: the same rows (same lengths, same values) with each row's
    # columns contiguous near its diagonal, so x's lines are reused from
    # cache. The kernel and the byte count are the same; only the gather
    # pattern differs, which separates what the kernel streams from what
    # the random pattern costs.
    lens = np.diff(e["start"]).astype(np.int64)
    base = np.minimum((np.arange(lp.rows, dtype=np.int64) * lp.cols) // lp.rows,
                      lp.cols - np.maximum(lens, 1))
    within = np.arange(int(lens.sum()), dtype=np.int64) - np.repeat(e["start"][:-1].astype(np.int64), lens)
    band = Matrix(lp.rows, lp.cols, e["start"], (np.repeat(base, lens) + within).astype(np.int32),
                  e["value"])
    ctx.prepare(band)
    ctx.multiply(band, x, y)
    if not args.no_profile:
        ctx.set_profiling(True)
        for _ in range(args.c1_reps):
            ctx.multiply(band, x, y)
        row = parse_profile(ctx.profile_report())["spmv"]
        ctx.set_profiling(False)
        spmv["K x, banded control"] = {"mean_us": row["mean_us"],
                                        "achieved_GBs": round(row["gbs"], 1),
                                        "frac": round(row["gbs"] / HBM_PEAK_GBS, 4),
                                        "launches": row["timed"]}
    # Control: K's own index stream with one entry per row, so the product
    # makes exactly K x's gathers of x, in the same order, with no row sums
    # (and writes one output per entry). Its time is what the gathers alone
    # cost on this pattern.
    one = Matrix(lp.nnz, lp.cols, np.arange(lp.nnz + 1, dtype=np.int32), e["index"], e["value"])
    yg = ctx.allocate(lp.nnz)
    ctx.prepare(one)
    ctx.multiply(one, x, yg)
    if not args.no_profile:
        ctx.set_profiling(True)
        for _ in range(args.c1_reps):
            ctx.multiply(one, x, yg)
        row = parse_profile(ctx.profile_report())["spmv"]
        ctx.set_profiling(False)
        spmv["K x, gather control"] = {
            "mean_us": row["mean_us"], "launches": row["timed"],
            "gathers_per_s": round(lp.nnz / (1e-6 * row["mean_us"]), 0),
            "what": "K's index stream as one entry per row: the same gathers of x, no row sums, "
                    "8 B written per entry"}
    else:
        # Launched plainly, so an external profiler sees the same products.
        for _ in range(args.c1_reps):
            ctx.multiply(k, x, y)
            ctx.multiply_transpose(kt, yv, xv)
            ctx.multiply(band, x, y)
            ctx.multiply(one, x, yg)
        ctx.synchronize()
    yg.free()
    host = None
    ref = None if args.no_cpu_baseline else ref_oracle()
    if ref is not None:
        # The reference backend's multiply on the host (serial, and its
        # threaded backend over the pool), and a bounded run of its
        # solve_pdhg on the threaded backend for iterations/s.
        hi = host_info()
        xh = rng.uniform(-1, 1, lp.cols)
        bytes_kx = 12.0 * lp.nnz + 4.0 * (lp.rows + 1) + 8.0 * lp.cols + 8.0 * lp.rows
        host = {"spmv_K_x": {}}
        for label, t in (("t1", 1), ("t_pool", hi["pool_threads"])):
            sec = ref.ref_time_multiply(e["start"], e["index"], e["value"], xh, threads=t,
                                        reps=args.c1_host_reps)
            host["spmv_K_x"][label] = {"threads": t, "mean_us": round(1e6 * sec, 1),
                                       "achieved_GBs": round(bytes_kx / sec / 1e9, 2)}
        R = ref.RefLp.from_export(e)
        t0 = time.perf_counter()
        w = R.solve(threads=hi["pool_threads"], tolerance=1e-4,
                    max_iterations=args.c1_host_iterations)
        hs = time.perf_counter() - t0
        host["solve_sample"] = {"threads": hi["pool_threads"], "iterations": w.iterations,
                                "seconds": round(hs, 2),
                                "iterations_per_s": round(w.iterations / hs, 1),
                                "note": "reference solve_pdhg capped at "
                                        f"{args.c1_host_iterations} iterations, setup included"}
    for v in (x, y, yv, xv):
        v.free()
    ctx.release()

    t0 = time.perf_counter()
    r = solve_pdhg(ctx, lp, tolerance=1e-4, time_limit_seconds=args.c1_seconds,
                   use_graphs=0 if args.no_graphs else 1)
    solve_s = time.perf_counter() - t0
    kern = {}
    if not args.no_profile:
        ctx.set_profiling(True)
        solve_pdhg(ctx, lp, tolerance=1e-4, max_iterations=min(max(r.iterations, 80), 2000),
                   use_graphs=0)
        kern = parse_profile(ctx.profile_report())
        ctx.set_profiling(False)
    steps = {}
    for name in ("primal_step", "dual_step"):
        if name not in kern:
            continue
        mean_us = kern[name]["mean_us"]
        nbytes = kern[name]["gbs"] * 1e3 * mean_us  # algorithmic bytes per launch
        combine = kern.get(name + "_combine")
        if combine is not None:
            # With column panels the step is two dispatches, each timed by
            # its own event pair: the panel products and the combine.
            mean_us += combine["mean_us"]
        gbs = nbytes / (1e3 * mean_us) if mean_us > 0 else 0.0
        steps[name] = {"mean_us": round(mean_us, 3), "achieved_GBs": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if combine is not None:
            steps[name]["combine_us"] = combine["mean_us"]
    it_s = r.phases.get("iterating", 0.0) + r.phases.get("checks", 0.0)
    return {"workload": f"random LP m={lp.rows} n={lp.cols} nnz={lp.nnz} (Zipf column counts, "
                        f"seed 20261204), plain PDHG to rel KKT 1e-4",
            "build_s": round(build_s, 2), "spmv": spmv, "fused_steps": steps,
            "cpu_reference": host,
            "solve": {"status": r.status, "iterations": r.iterations,
                      "polish_attempts": r.polish_attempts,
                      "polish_iterations": r.polish_iterations, "restarts": r.restarts,
                      "objective": r.objective, "seconds": round(solve_s, 3),
                      "iterations_per_s": round(r.iterations / it_s, 1) if it_s > 0 else None,
                      "phases_s": {k2: round(v, 4) for k2, v in r.phases.items()}}}


def c3_section(ctx, args):
    """The uneven stack: count independent LPs (half refinery plans of 4-52
    periods, half random LPs with m ~ logU[1e3, 5e4], n = 2.5 m), each with its
    own K, solved to rel KKT 1e-4 in one set of launches over their
    block-diagonal stack; beside it, plain GPU solves of the same LPs one at a
    time on a bounded sample."""
    import numpy as np
    from anukulan_amd import generators as G, solve_pdhg, solve_pdhg_multi

    t0 = time.perf_counter()
    models = G.uneven_stack(args.c3_count)
    lps = [m.to_lp() for m in models]
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    rs = solve_pdhg_multi(ctx, lps, tolerance=1e-4, use_graphs=0 if args.no_graphs else 1)
    solve_s = time.perf_counter() - t0
    tts = np.array([r.solve_seconds for r in rs])
    plain_n, plain_s, worst = 0, 0.0, 0.0
    while plain_n < len(lps) and plain_s < args.c3_plain_seconds:
        t1 = time.perf_counter()
        p = solve_pdhg(ctx, lps[plain_n], tolerance=1e-4, use_graphs=0 if args.no_graphs else 1)
        plain_s += time.perf_counter() - t1
        worst = max(worst, abs(p.objective - rs[plain_n].objective) / max(1.0, abs(p.objective)))
        plain_n += 1
    cpu = None
    if not args.no_cpu_baseline:
        # The reference's solve_pdhg (oracle/_ref, CPU backend) on every 32nd
        # LP, one LP per host thread: the CPU seconds each LP takes give the
        # rate of one core, and the host's cores running one LP each give the
        # all-cores rate (a projection: the sample stands for the stack).
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        try:
            import _oracle
            ok = _oracle.ref is not None
        except Exception:
            ok = False
        if ok:
            from concurrent.futures import ThreadPoolExecutor
            picks = list(range(0, len(models), 32))
            threads = min(16, len(os.sched_getaffinity(0)))

            def one(i):
                r = _oracle.RefLp.from_natural(models[i])
                t1 = time.perf_counter()
                out = r.solve(tolerance=1e-4)
                return time.perf_counter() - t1, out.status == "optimal", out.iterations
            t1 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=threads) as pool:
                runs = list(pool.map(one, picks))
            wall = time.perf_counter() - t1
            cpu_s = sum(t for t, _, _ in runs)
            per_core = len(picks) / cpu_s
            cpu = {"value": round(per_core, 3), "unit": "LPs/s", "cores": 1, "kind": "reference",
                   "sample": f"every 32nd of the {len(models)} LPs ({len(picks)}: refinery plans "
                             f"and random LPs), reference solve_pdhg to 1e-4, one LP per thread on "
                             f"{threads} threads: {cpu_s:.1f} CPU s ({wall:.1f} s wall)",
                   "optimal": sum(ok for _, ok, _ in runs),
                   "all_cores_projected": round(per_core * threads, 2),
                   "all_cores_threads": threads}
    return {"workload": f"{len(lps)} independent LPs of uneven size (half refinery plans of 4-52 "
                        f"periods, half random LPs m~logU[1e3,5e4], n=2.5m), each its own K, "
                        f"rel KKT 1e-4",
            "cpu_reference": cpu,
            "rows": int(sum(lp.rows for lp in lps)), "cols": int(sum(lp.cols for lp in lps)),
            "nnz": int(sum(lp.nnz for lp in lps)), "build_s": round(build_s, 2),
            "lps_per_s": round(len(lps) / solve_s, 2), "seconds": round(solve_s, 3),
            "optimal": int(sum(r.status == "optimal" for r in rs)),
            "time_to_solve_s": {"p50": round(float(np.percentile(tts, 50)), 4),
                                "p99": round(float(np.percentile(tts, 99)), 4),
                                "max": round(float(tts.max()), 4)},
            "iterations": {"p50": int(np.percentile([r.iterations for r in rs], 50)),
                           "max": int(max(r.iterations for r in rs))},
            "phases_s": {k: round(v, 4) for k, v in rs[0].phases.items()},
            "autotune": [t for t in parse_tuning(ctx.tuning_report())
                         if t["kernel"].startswith("multi_")],
            "plain_sample": {"lps": plain_n, "seconds": round(plain_s, 3),
                             "lps_per_s": round(plain_n / plain_s, 2) if plain_s > 0 else None,
                             "max_rel_objective_diff": worst}}


def c4_section(ctx, args, base, cases):
    """Sensitivities on the case stack: the 61 cases batched to rel KKT
    1e-8, then the adjoint of the PDHG fixed point for L = w'x (w on the
    purchase and sale columns) of every case in one launch set, timed; the
    gradients against the committed one-sided finite differences and the
    reference simplex's basis sensitivities B^-1 at 200 coordinates per case
    (tests/golden/adjoint_reference.npz), on the coordinates where the two
    one-sided differences agree (elsewhere the derivative is set-valued).
    The reference has no adjoint, so no CPU figure stands beside it."""
    import numpy as np
    from anukulan_amd import generators as G, pdhg_adjoint, solve_pdhg_batch

    fx = np.load(os.path.join(ROOT, "tests", "golden", "adjoint_reference.npz"))
    lps = [G.case_model(base, c).to_lp() for c in cases]
    t0 = time.perf_counter()
    rs = solve_pdhg_batch(ctx, lps, tolerance=1e-8, max_iterations=2_000_000,
                          time_limit_seconds=1e9, use_graphs=0 if args.no_graphs else 1)
    solve_s = time.perf_counter() - t0
    ws = [fx[f"stack24_{i}_w_x"] for i in range(len(lps))]
    t0 = time.perf_counter()
    adj = pdhg_adjoint(ctx, lps, [r.x for r in rs], [r.y for r in rs], w_x=ws)
    adjoint_s = time.perf_counter() - t0
    unique = agree = 0
    worst_fd = worst_b = 0.0
    over = 0
    for i, a in enumerate(adj):
        g = (a.grad_c, a.grad_q, a.grad_lower, a.grad_upper)
        got = np.array([g[f][j] for f, j in fx[f"stack24_{i}_coords"]])
        p, m = fx[f"stack24_{i}_fd_plus"], fx[f"stack24_{i}_fd_minus"]
        scale = np.maximum(1.0, np.maximum(np.abs(p), np.abs(m)))
        u = np.isfinite(p) & np.isfinite(m) & (np.abs(p - m) <= 1e-6 * scale)
        fd = 0.5 * (p + m)
        basis = fx[f"stack24_{i}_basis"]
        e_fd = np.abs(got - fd) / np.maximum(1.0, np.abs(fd))
        v = u & (np.abs(basis - fd) <= 1e-6 * np.maximum(1.0, np.abs(fd)))
        e_b = np.abs(got - basis) / np.maximum(1.0, np.abs(basis))
        unique += int(u.sum())
        agree += int(v.sum())
        over += int((u & (e_fd > 1e-5)).sum())
        if u.any():
            worst_fd = max(worst_fd, float(e_fd[u].max()))
        if v.any():
            worst_b = max(worst_b, float(e_b[v].max()))
    total = 200 * len(adj)
    return {"workload": f"{len(lps)}-case {args.periods}-period stack: batched PDHG to rel KKT "
                        f"1e-8, then the adjoint of L = w'x on the purchase and sale columns",
            "cases": len(lps), "cases_optimal": int(sum(r.status == "optimal" for r in rs)),
            "solve_s": round(solve_s, 3), "adjoint_s": round(adjoint_s, 3),
            "adjoint_cases_per_s": round(len(lps) / adjoint_s, 2),
            "adjoint_converged": int(sum(a.converged for a in adj)),
            "adjoint_iterations_max": int(max(a.iterations for a in adj)),
            "coordinates": total, "coordinates_with_a_derivative": unique,
            "set_valued_share": round(1 - unique / total, 4),
            "max_rel_error_vs_differences": worst_fd,
            "coordinates_over_1e-5": over,
            "over_1e-5_note": "tests/test_gpu_adjoint.py shows each such coordinate to be one "
                              "along which the basis the adjoint completed is infeasible both "
                              "ways, and allows at most one in a thousand",
            "coordinates_where_basis_agrees": agree,
            "max_rel_error_vs_basis": worst_b}


def nodes_section(ctx, args):
    """Branch-and-bound node relaxations: the 64 nodes of the depth-6
    frontier of the 12-period refinery plan (crude purchases and product
    sales integer; branched on the six most fractional of them at the root
    optimum, the recipe in tests/golden/nodes_reference.json), each solved
    as the reference tree's first-order node solver solves it (default
    options),
    all 64 in one batch; beside it the plain GPU solver one node at a time
    and the reference's solve_pdhg (oracle/_ref, one thread), each on a
    bounded sample of the nodes."""
    import numpy as np
    from anukulan_amd import LP, generators as G, solve_pdhg
    from anukulan_amd.nodes import branch_frontier, solve_node_relaxations

    with open(os.path.join(ROOT, "tests", "golden", "nodes_reference.json")) as f:
        fx = json.load(f)["frontiers"]["refinery12_depth6"]
    lp = G.refinery_model(fx["periods"]).to_lp()
    e = lp.export()
    lo, hi = branch_frontier(e["lower"], e["upper"], fx["columns"],
                             [float.fromhex(v) for v in fx["values"]])
    solve_node_relaxations(ctx, lp, lo[:2], hi[:2])  # warm-up (tuner, graphs)
    t0 = time.perf_counter()
    got = solve_node_relaxations(ctx, lp, lo, hi)
    batched_s = time.perf_counter() - t0
    want = fx["nodes"]
    verdicts = sum((g.optimal, g.proved_infeasible) == (w["optimal"], w["proved_infeasible"])
                   for g, w in zip(got, want))
    worst = max(abs(g.bound - float.fromhex(w["bound"])) / max(1.0, abs(float.fromhex(w["bound"])))
                for g, w in zip(got, want) if w["optimal"])

    def node_lp(i):
        return LP.from_standard(lp.rows, lp.cols, e["num_equalities"], e["start"], e["index"],
                                e["value"], e["c"], e["q"], lo[i], hi[i], e["objective_scale"],
                                e["objective_offset"])

    plain_n, plain_s = 0, 0.0
    while plain_n < len(want) and plain_s < args.nodes_plain_seconds:
        nl = node_lp(plain_n)
        t1 = time.perf_counter()
        solve_pdhg(ctx, nl)
        plain_s += time.perf_counter() - t1
        plain_n += 1
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        try:
            import _oracle
            ok = _oracle.ref is not None
        except Exception:
            ok = False
        if ok:
            cpu_n, cpu_s = 0, 0.0
            while cpu_n < len(want) and cpu_s < args.nodes_cpu_seconds:
                R = _oracle.RefLp.from_export(dict(e, lower=lo[cpu_n], upper=hi[cpu_n]))
                t1 = time.perf_counter()
                R.solve(threads=1)
                cpu_s += time.perf_counter() - t1
                cpu_n += 1
            cpu = {"value": round(cpu_n / cpu_s, 3), "unit": "nodes/s", "cores": 1,
                   "kind": "reference", "sample": f"first {cpu_n} nodes of the frontier"}
    value = len(want) / batched_s
    return {"workload": f"depth-{fx['depth']} branch-and-bound frontier of the "
                        f"{fx['periods']}-period refinery plan ({len(want)} node relaxations "
                        f"sharing K, default options: rel KKT 1e-6)",
            "rows": lp.rows, "cols": lp.cols, "nnz": lp.nnz,
            "nodes_per_s": round(value, 2), "seconds": round(batched_s, 3),
            "iterations": int(sum(g.relaxation.iterations for g in got)),
            "verdicts_as_reference": f"{verdicts}/{len(want)}",
            "max_rel_bound_diff_vs_reference": worst,
            "plain_sample": {"nodes": plain_n, "seconds": round(plain_s, 3),
                             "nodes_per_s": round(plain_n / plain_s, 2) if plain_s > 0 else None},
            "cpu_baseline": cpu,
            "batched_vs_plain": round(value / (plain_n / plain_s), 2) if plain_s > 0 else None,
            "batched_vs_cpu": round(value / cpu["value"], 2) if cpu else None}


def finish_section(ctx, args, base, cases, lps, optima):
    """Exact finish of the case stack: every case solved by the batched
    device solver to a seed tolerance (1e-4, 1e-6 or 1e-8), then finished to
    an optimal basis on the host cores (crossover of the first-order point,
    then the bounded primal simplex from it), one case per thread; beside it
    the same finish from all logicals (cold) and the reference's simplex
    (oracle/_ref, one thread) on a bounded sample. Errors are against the
    reference simplex's FP64 optima."""
    from anukulan_amd import solve_pdhg_batch
    from anukulan_amd.finish import exact_finish

    seeds = {}
    for tol in (1e-4, 1e-6, 1e-8):
        t0 = time.perf_counter()
        pts = solve_pdhg_batch(ctx, lps, tolerance=tol)
        pdhg_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        warm = exact_finish(lps, [p.x for p in pts], [p.y for p in pts])
        fin_s = time.perf_counter() - t0
        err = ([abs(w.objective - o) / max(1.0, abs(o)) for w, o in zip(warm, optima)]
               if optima else [])
        seeds[f"{tol:g}"] = {"pdhg_s": round(pdhg_s, 3), "finish_s": round(fin_s, 3),
                             "cases_per_s": round(len(lps) / (pdhg_s + fin_s), 3),
                             "pivots": sum(w.iterations for w in warm),
                             "cases_optimal": sum(w.status == "optimal" for w in warm),
                             "max_rel_error_vs_simplex": max(err) if err else None}
    # The deciding cases only: the GPU bounds every case at 1e-4, and the
    # five with the best first-order objective are finished exactly, which
    # settles which case is best.
    t0 = time.perf_counter()
    pts = solve_pdhg_batch(ctx, lps, tolerance=1e-4)
    top = sorted(range(len(lps)), key=lambda i: pts[i].objective)[:5]
    fin = exact_finish([lps[i] for i in top], [pts[i].x for i in top], [pts[i].y for i in top])
    decide_s = time.perf_counter() - t0
    best = top[min(range(len(top)), key=lambda k: fin[k].objective)]
    deciding = {"seconds": round(decide_s, 3), "cases_finished": len(top),
                "best_case": best, "all_finished_optimal": all(f.status == "optimal" for f in fin)}
    if optima:
        deciding["best_case_is_the_simplex_best"] = best == min(range(len(optima)),
                                                                key=lambda i: optima[i])
        deciding["max_rel_error_vs_simplex"] = max(
            abs(f.objective - optima[i]) / max(1.0, abs(optima[i])) for f, i in zip(fin, top))
    # The reference's case-stack workflow with this library's pieces: the
    # base plan's first-order point (device, 1e-4) crossed over and finished,
    # then every what-if case started from the base's final basis on the
    # host cores (anuk_exact_finish_basis), not from its own point.
    from anukulan_amd import solve_pdhg
    t0 = time.perf_counter()
    p0 = solve_pdhg(ctx, lps[0], tolerance=1e-4)
    f0 = exact_finish(lps[:1], [p0.x], [p0.y])[0]
    base_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    rest = exact_finish(lps[1:], start=f0)
    rest_s = time.perf_counter() - t0
    wf = [f0] + rest
    wf_err = ([abs(w.objective - o) / max(1.0, abs(o)) for w, o in zip(wf, optima)]
              if optima else [])
    from_base = {"base_s": round(base_s, 3), "warm_s": round(rest_s, 3),
                 "cases_per_s": round(len(lps) / (base_s + rest_s), 3),
                 "base_pivots": f0.iterations, "warm_pivots": sum(w.iterations for w in rest),
                 "warm_phase_one_pivots": sum(w.info["phase_one_iterations"] for w in rest),
                 "cases_optimal": sum(w.status == "optimal" for w in wf),
                 "max_rel_error_vs_simplex": max(wf_err) if wf_err else None}
    t0 = time.perf_counter()
    cold = exact_finish(lps)
    cold_s = time.perf_counter() - t0
    cold_err = ([abs(c.objective - o) / max(1.0, abs(o)) for c, o in zip(cold, optima)]
                if optima else [])
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        try:
            import _oracle
            ok = _oracle.ref is not None
        except Exception:
            ok = False
        if ok:
            from anukulan_amd import generators as G
            n_done, spent = 0, 0.0
            while n_done < len(cases) and spent < args.finish_cpu_seconds:
                R = _oracle.RefLp.from_natural(G.case_model(base, cases[n_done]))
                t1 = time.perf_counter()
                R.simplex()
                spent += time.perf_counter() - t1
                n_done += 1
            cpu = {"value": round(n_done / spent, 3), "unit": "cases/s", "cores": 1,
                   "kind": "reference", "sample": f"reference simplex on the first {n_done} cases"}
            # The reference's own case-stack workflow (app/anukulan_cli.cpp:1065-1230,
            # `family --from-first`): the base plan solved cold, then each
            # case re-solved by the dual simplex warm-started from the base's
            # final basis. Its dual simplex stalls at the pivot limit on this
            # base plan from all logicals, so the base is solved by its
            # primal (the reference's default algorithm).
            R0 = _oracle.RefLp.from_natural(G.case_model(base, cases[0]))
            t1 = time.perf_counter()
            b0 = R0.simplex_warm(dual=False)
            base_s = time.perf_counter() - t1
            warm_n, warm_s, warm_piv, warm_ok = 0, 0.0, 0, 0
            while 1 + warm_n < len(cases) and warm_s < args.finish_cpu_seconds:
                R = _oracle.RefLp.from_natural(G.case_model(base, cases[1 + warm_n]))
                t1 = time.perf_counter()
                w = R.simplex_warm(dual=True, start=(b0["basic"], b0["var_status"]))
                warm_s += time.perf_counter() - t1
                warm_piv += w["pivots"]
                warm_ok += w["status"] == 0
                warm_n += 1
            per_case = warm_s / max(warm_n, 1)
            cpu["warm_stack_workflow"] = {
                "kind": "reference", "cores": 1, "base_cold_s": round(base_s, 3),
                "base_status": int(b0["status"]), "warm_cases": warm_n,
                "warm_cases_optimal": warm_ok, "warm_s": round(warm_s, 3),
                "warm_pivots": warm_piv,
                "warm_cases_per_s": round(warm_n / warm_s, 3) if warm_s > 0 else None,
                "stack_cases_per_s_projected": round(
                    len(cases) / (base_s + per_case * (len(cases) - 1)), 3),
                "sample": f"base plan cold (primal), then the first {warm_n} what-if cases "
                          "warm-started from its basis (dual simplex), one thread"}
    best = max(seeds, key=lambda k: seeds[k]["cases_per_s"])
    threads = min(16, len(os.sched_getaffinity(0)))
    top_rate = max(seeds[best]["cases_per_s"], from_base["cases_per_s"])
    return {"workload": f"{len(lps)}-case stack finished to optimal bases on {threads} host "
                        f"threads: each case from its own batched PDHG point (crossover, then "
                        f"primal simplex), or every case from the base plan's final basis",
            "cases_per_s": top_rate,
            "best_method": ("from the base plan's basis" if top_rate == from_base["cases_per_s"]
                            else f"own PDHG point at {best}"),
            "best_seed_tolerance": best,
            "by_seed_tolerance": seeds,
            "from_base_basis": from_base,
            "deciding_cases": deciding,
            "cold": {"seconds": round(cold_s, 3), "cases_per_s": round(len(lps) / cold_s, 3),
                     "pivots": sum(c.iterations for c in cold),
                     "max_rel_error_vs_simplex": max(cold_err) if cold_err else None},
            "cpu_baseline": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--periods", type=int, default=24)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--cases", type=int, default=61,
                    help="solve only the first N cases of the stack (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-cores reference baseline (default: the usable "
                         "cores, at most 16)")
    ap.add_argument("--no-profile", action="store_true",
                    help="no HIP-event timing pass (under an external profiler)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly (for external profilers)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the batched kernels' default launch shape instead of timing "
                         "the shapes (under an external profiler, whose overhead would skew "
                         "the timing)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 1M-column random LP")
    ap.add_argument("--no-c3", action="store_true", help="skip the uneven 512-LP stack")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the sensitivities of the 61-case 24-period stack")
    ap.add_argument("--c3-count", type=int, default=512)
    ap.add_argument("--c3-plain-seconds", type=float, default=10.0,
                    help="time spent on plain one-at-a-time solves of the uneven stack's LPs")
    ap.add_argument("--no-finish", action="store_true", help="skip the exact finish of the stack")
    ap.add_argument("--finish-cpu-seconds", type=float, default=8.0,
                    help="time for each reference simplex sample of the exact finish")
    ap.add_argument("--no-nodes", action="store_true",
                    help="skip the branch-and-bound node frontier")
    ap.add_argument("--nodes-plain-seconds", type=float, default=8.0)
    ap.add_argument("--nodes-cpu-seconds", type=float, default=8.0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB and exit: the "
                         "known-size reference for HBM byte counters")
    ap.add_argument("--no-c0", action="store_true", help="skip the 12-period plain solve")
    ap.add_argument("--c0-reps", type=int, default=5)
    ap.add_argument("--c1-host-reps", type=int, default=5,
                    help="reference host SpMV repetitions on the 1M-column LP")
    ap.add_argument("--c1-host-iterations", type=int, default=160,
                    help="iteration cap of the reference's host solve sample on the 1M-column LP")
    ap.add_argument("--c1-rows", type=int, default=400_000)
    ap.add_argument("--c1-cols", type=int, default=1_000_000)
    ap.add_argument("--c1-reps", type=int, default=50)
    ap.add_argument("--c1-seconds", type=float, default=60.0,
                    help="time limit of the random LP's plain solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(a case stack would shard by case across GPUs, one process each; not implemented)",
              file=sys.stderr)
        sys.exit(2)

    from anukulan_amd import Context, solve_pdhg, solve_pdhg_batch

    if args.copy_probe_mb:
        ctx = Context(0)
        n = args.copy_probe_mb * (1 << 20) // 8
        a, b = ctx.allocate(n), ctx.allocate(n)
        ctx.set_profiling(True)
        ctx.copy(a, b)
        row = parse_profile(ctx.profile_report())["copy"]
        print(json.dumps({"copy_probe_bytes": 8 * n, "mean_us": row["mean_us"],
                          "GB/s": row["gbs"]}))
        a.free()
        b.free()
        ctx.close()
        return

    base, cases, lps = build_stack(args.periods, args.cases)
    n, m, nnz = lps[0].cols, lps[0].rows, lps[0].nnz
    graphs = 0 if args.no_graphs else 1
    ctx = Context(0)

    tune = 0 if args.no_autotune else 1

    def batched():
        return solve_pdhg_batch(ctx, lps, tolerance=args.tol, use_graphs=graphs, autotune=tune)

    for _ in range(args.warmup):
        batched()
    ctx.synchronize()
    t0 = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results = batched()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    value = len(lps) * args.steps / elapsed
    solved = sum(1 for r in results if r.status == "optimal")
    lane_iters = sum(r.iterations for r in results)
    tuning = ctx.tuning_report()
    optima = simplex_optima(args.periods, len(lps))

    plain = None
    if not args.no_plain:
        t0 = time.perf_counter()
        pres = [solve_pdhg(ctx, lp, tolerance=args.tol, use_graphs=graphs) for lp in lps]
        pt = time.perf_counter() - t0
        worst = max(abs(a.objective - b.objective) / max(1.0, abs(b.objective))
                    for a, b in zip(results, pres))
        plain = {"value": round(len(lps) / pt, 3), "unit": "cases/s",
                 "seconds": round(pt, 3), "iterations": sum(r.iterations for r in pres),
                 "cases_optimal": sum(1 for r in pres if r.status == "optimal"),
                 "max_rel_objective_diff_vs_batched": worst,
                 "error_vs_simplex": error_vs_optimum(pres, optima)}

    roofline, kernels = None, {}
    if not args.no_profile:
        ctx.set_profiling(True)
        solve_pdhg_batch(ctx, lps, tolerance=args.tol, use_graphs=0)
        kernels = parse_profile(ctx.profile_report())
        ctx.set_profiling(False)
        ranked = sorted(((k, v) for k, v in kernels.items() if v["gbs"] > 0),
                        key=lambda kv: -kv[1]["total_ms"])
        if ranked:
            name, k = ranked[0]
            roofline = {"bound": "hbm", "kernel": name, "achieved": round(k["gbs"], 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(k["gbs"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(name),
                        "traffic_source": "profiles/pmc_traffic.txt (rocprofv3 --pmc: size-resolved L2-to-memory read requests plus WRITE_SIZE)",
                        "mean_launch_us": k["mean_us"], "launches_timed": k["timed"]}

    c0 = None if args.no_c0 else c0_section(ctx, args)
    c1 = None if args.no_c1 else c1_section(ctx, args)
    c3 = None if args.no_c3 else c3_section(ctx, args)
    c4 = None if args.no_c4 or args.periods != 24 or args.cases != 61 else c4_section(ctx, args, base, cases)
    nodes = None if args.no_nodes else nodes_section(ctx, args)
    finish = None if args.no_finish else finish_section(ctx, args, base, cases, lps, optima)

    cpu = cpu_par = cpu_thr = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(base, cases, args.tol, args.cpu_seconds)
        threads = args.cpu_threads or host_info()["pool_threads"]
        cpu_par = cpu_baseline_parallel(base, cases, args.tol, threads)
        cpu_thr = cpu_baseline_threaded(base, cases, args.tol, args.cpu_seconds / 3)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "cases/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"refinery case stack: {args.periods}-period base plan + "
                               f"{len(lps) - 1} what-if cases sharing K, each to rel KKT "
                               f"{args.tol:g}",
                   "solver": "batched PDHG (one wave64 lane per case, K read once per step)",
                   "rows": m, "cols": n, "nnz": nnz, "cases": len(lps),
                   "cases_optimal": solved, "lane_iterations_per_step": lane_iters,
                   "parallelism": "single GPU"},
        "error_vs_simplex": error_vs_optimum(results, optima),
        "roofline": roofline, "cpu_baseline": cpu, "cpu_baseline_all_cores": cpu_par,
        "cpu_baseline_threaded_backend": cpu_thr, "host": host_info(),
        "plain": plain, "c0": c0, "c1": c1, "c3": c3, "c4": c4, "nodes": nodes, "finish": finish,
        "batch_phases_s": {k: round(v, 4) for k, v in results[0].phases.items()},
        "autotune": parse_tuning(tuning),
        "kernels": {k: {"mean_us": v["mean_us"], "calls": v["calls"], "GB/s": v["gbs"]}
                    for k, v in kernels.items()},
    }
    if plain:
        out["batched_vs_plain"] = round(value / plain["value"], 2)
    if cpu and cpu.get("value"):
        out["batched_vs_cpu_baseline"] = round(value / cpu["value"], 2)
    if cpu_par and cpu_par.get("value"):
        out["batched_vs_cpu_all_cores"] = round(value / cpu_par["value"], 2)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
