#!/usr/bin/env python
"""Benchmark of the Caputo L1 history-sum solver on one MI355X.

Workload (default): 1D time-fractional moisture transport, N=10000 nodes,
M=20000 steps, tau=5e-4, alpha=0.7, seeded initial head (seed 1001); FP64
weights, FP64 partitioned-Thomas step, Picard iteration on the VGM
coefficients (at most 3 iterations, tol 1e-12; --picard 1 = lagged).  A bench
"step" is one complete solve of that workload (all M time steps) with the
initial field already in HBM; value = N*M / (wall time of one solve).

Variants (history-sum forms): plain FP64 streaming, FP64 MFMA time-block
GEMM (block 64), split bf16 hi+lo MFMA time-block GEMM (block 64, fp32
accumulate flushed to fp64 per 32-level group).  The headline value is the
split bf16 form; every variant reports its error against the plain FP64
result.  Prints ONE JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0
MI355X_PEAK_TFLOPS = {"f64": 78.6, "bf16": 2500.0, "f16": 2500.0}

VARIANT_DTYPE = {"plain": "f64", "mfma_f64": "f64", "split_bf16": "bf16x2->f32->f64",
                 "split_f16": "f16x2->f32->f64", "split_bf16x3": "bf16x3->f32->f64"}


def run_variant(F, workload, mode, block, steps, warmup, prof_mask_timed=None, nodes=None, msteps=None,
                picard=(1, 0.0), gamma=0.0):
    kw = {} if workload == "3d" else dict(picard_max=picard[0], picard_tol=picard[1])
    spec, u0 = F.workload(workload, hist_mode=mode, block=block, nodes=nodes, steps=msteps, gamma=gamma, **kw)
    tuned = None
    if block == 0 and mode != "plain":
        # untimed: every candidate block runs the whole solve once
        b, ms = F.autotune_block(spec, u0, trial_steps=spec.steps)
        spec.block = b
        tuned = {"block": b, "trial_ms_per_step": ms}
        print(f"[bench] {mode}: autotuned block {b} ({ms * 1e3:.1f} us/step)", file=sys.stderr, flush=True)
    s = F.Solver(spec)
    s.set_initial(u0)
    # profiled warmup: per-class device times with hipEvents on the solver stream
    s.set_profiling(3)
    s.advance(0)
    s.sync()
    u = s.solution()
    prof = s.stats()
    for _ in range(max(0, warmup - 1)):
        s.reset(u0)
        s.advance(0)
        s.sync()
    # timed solves; events only around the kernel class named by prof_mask_timed
    mask = prof_mask_timed if prof_mask_timed is not None else 0
    times = []
    timed_stats = []
    for _ in range(steps):
        s.reset(u0)
        s.set_profiling(mask)
        t0 = time.perf_counter()
        s.advance(0)
        s.sync()
        times.append(time.perf_counter() - t0)
        timed_stats.append(s.stats())
    u_last = s.solution()
    s.close()
    return dict(spec=spec, u=u_last, u_first=u, times=times, prof=prof, timed=timed_stats, tuned=tuned)


def cpu_baseline(N, M, seconds_hint=10.0, alpha=0.7, tau=5e-4, seed=1001, picard=(1, 0.0)):
    """FP64 oracle (oracle/, the reference loop order: node-outer, t-inner,
    OpenMP over nodes) timed on this host: the history sum at the run-average
    history length n = M/2 over all N nodes, repeated for about seconds_hint
    seconds of CPU work (median of the repetitions), plus one oracle step (VGM
    coefficients + Thomas, median over repeated 4-step solves);
    node*steps/s = N / (t_sum + t_step)."""
    import numpy as np
    from oracle import oracle as O
    O.lib()
    cores = min(16, os.cpu_count() or 1)
    n = M // 2
    rng = np.random.default_rng(0)
    old = np.cumsum(rng.standard_normal((n + 1, N)) * 1e-4, axis=0)
    Time = O.time_grid(n + 1, tau)
    _, kb = O.l1_weights(alpha, n, Time, weight_mode=1)
    taus = np.full(n + 2, tau)
    reps = []
    t_start = time.perf_counter()
    while len(reps) < 3 or time.perf_counter() - t_start < seconds_hint:
        t0 = time.perf_counter()
        O.time_summing(old, kb, taus, nthreads=cores)
        reps.append(time.perf_counter() - t0)
    t_sum = statistics.median(reps)
    t_sum_total = time.perf_counter() - t_start
    del old
    u0 = O.init_1d(N, seed)
    p = O.make_params(N, 4, tau, alpha, u_top=u0[0], nthreads=cores, picard_max=picard[0], picard_tol=picard[1])
    steps = []
    t_start = time.perf_counter()
    while len(steps) < 3 or (time.perf_counter() - t_start < 1.0 and len(steps) < 50):
        t0 = time.perf_counter()
        O.solve_1d(p, u0)
        steps.append((time.perf_counter() - t0) / 4)
    t_step = statistics.median(steps)
    return {
        "value": N / (t_sum + t_step),
        "unit": "node*steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle time_summing over n={n} levels x N={N} nodes (run-average history "
                   f"length for M={M}), {len(reps)} reps in {t_sum_total:.1f} s, median {t_sum * 1e3:.3f} ms; "
                   f"+ oracle step {t_step * 1e3:.3f} ms (median of {len(steps)} 4-step solves); "
                   f"extrapolated to the whole run"),
    }


def ensemble_run(F, count, nodes, msteps, hist_mode, picard, reps=3, uneven=False):
    """`count` independent 1D solves (alpha spread over [0.5, 0.9], own
    initial states; with `uneven`, N spread over [nodes/4, nodes] and M over
    [msteps/2, msteps] as well) timed three ways: one batch
    (fpde_solver_advance_batch: one cooperative launch per history block for
    all of them), one after another, and each on its own stream at once.
    Returns the medians and the largest difference of the batch results from
    the solo ones."""
    import numpy as np
    alphas = np.linspace(0.5, 0.9, count)
    if uneven:
        ns = [int(v) for v in np.linspace(nodes // 4, nodes, count)[::-1]]
        ms = [int(v) for v in np.linspace(msteps // 2, msteps, count)]
    else:
        ns, ms = [nodes] * count, [msteps] * count
    u0s = [F.init_1d(n, 700 + i) for i, n in enumerate(ns)]
    specs = [F.make_spec(n, m, 1e-3, alpha=float(a), hist_mode=hist_mode, block=64,
                         picard_max=picard[0], picard_tol=picard[1], u_top=u0[0])
             for n, m, a, u0 in zip(ns, ms, alphas, u0s)]
    solvers = [F.Solver(sp) for sp in specs]

    def reset_all():
        for s, u0 in zip(solvers, u0s):
            s.reset(u0)
        for s in solvers:
            s.sync()

    def batch():
        F.advance_batch(solvers, 0)

    def sequential():
        for s in solvers:
            s.advance(0)
            s.sync()

    def streams():
        for s in solvers:
            s.advance(0)
        for s in solvers:
            s.sync()

    times, sols = {}, {}
    for name, fn in (("batch", batch), ("sequential", sequential), ("streams", streams)):
        ts = []
        for _ in range(reps):
            reset_all()
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        times[name] = statistics.median(ts)
        sols[name] = [s.solution() for s in solvers]
        print(f"[ensemble] {name}: {times[name]:.3f} s for {count} solves", file=sys.stderr, flush=True)
    for s in solvers:
        s.close()
    diff = max(float(np.abs(b - a).max() / np.abs(a).max()) for a, b in zip(sols["sequential"], sols["batch"]))
    work = sum(n * m for n, m in zip(ns, ms))
    return {"count": count, "nodes": ns if uneven else nodes, "steps": ms if uneven else msteps,
            "alphas": [float(a) for a in alphas],
            "hist_mode": hist_mode, "seconds": times,
            "node_steps_per_s": {k: work / v for k, v in times.items()},
            "batch_speedup_vs_sequential": times["sequential"] / times["batch"],
            "batch_speedup_vs_streams": times["streams"] / times["batch"],
            "batch_max_rel_diff_vs_solo": diff}


def serve_mix(F, n_req, rate, seed=1004, quantum=256, budget_gb=200.0, max_active=16, hist_mode="split_f16"):
    """The serving mix: Poisson arrivals at `rate` req/s; 80% 1D (N~U[500,20000],
    M~U[1000,40000]), 20% 3D (32^3..96^3, M~U[200,1000]); alpha~U(0.3,0.95);
    25% distributed order (16 Gauss-Legendre nodes on [0.1,1]); seeded heads."""
    import numpy as np
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n_req):
        t += rng.exponential(1.0 / rate)
        is3d = rng.random() < 0.2
        do = rng.random() < 0.25
        alpha = float(rng.uniform(0.3, 0.95))
        if is3d:
            n = int(rng.integers(32, 97))
            M = int(rng.integers(200, 1001))
            u0 = F.init_3d(n, 100000 + i)
            kw = dict(dims=3, block=16)
        else:
            n = int(rng.integers(500, 20001))
            M = int(rng.integers(1000, 40001))
            u0 = F.init_1d(n, 100000 + i)
            kw = dict(dims=1, block=64)
        if do:
            kw.update(do_q=16, do_rule=0, do_a0=0.1, do_a1=1.0)
        tau = 1e-3 if is3d else 5e-4
        spec = F.make_spec(n, M, tau, alpha=alpha, hist_mode=hist_mode, u_top=u0[0], **kw)
        work = (n ** 3 if is3d else n) * M
        reqs.append((t, spec, u0, work))
    srv = F.Server(mem_budget_bytes=int(budget_gb * 2 ** 30), max_active=max_active)
    ids = []
    t0 = time.perf_counter()
    k = 0
    t_report = t0
    while k < len(reqs) or srv.poll(quantum) > 0:
        now = time.perf_counter() - t0
        while k < len(reqs) and reqs[k][0] <= now:
            ids.append((srv.submit(reqs[k][1], reqs[k][2]), reqs[k][3]))
            k += 1
        if k < len(reqs) and srv.poll(quantum) == 0:
            time.sleep(max(0.0, min(0.01, reqs[k][0] - (time.perf_counter() - t0))))
        if time.perf_counter() - t_report > 20.0:          # progress for long mixes
            t_report = time.perf_counter()
            done = sum(1 for rid, _ in ids if srv.info(rid)["status"] in (0, 1))
            print(f"[serve] {t_report - t0:.0f}s: {k} submitted, {done} finished", file=sys.stderr, flush=True)
    makespan = time.perf_counter() - t0
    lat, done_work, status = [], 0, []
    for rid, work in ids:
        inf = srv.info(rid)
        status.append(inf["status"])
        if inf["status"] == 0:
            lat.append(inf["finish_s"] - inf["submit_s"])
            done_work += work
    srv.close()
    lat = np.array(lat)
    return {
        "requests": n_req, "completed": int((np.array(status) == 0).sum()),
        "rejected": int((np.array(status) == 1).sum()), "rate_per_s": rate,
        "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
        "mean_latency_s": float(lat.mean()) if lat.size else None,
        "makespan_s": makespan, "node_steps_per_s": done_work / makespan,
        "hist_mode": hist_mode, "max_active": max_active, "quantum": quantum,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves of the headline variant")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1d_large")
    ap.add_argument("--headline", default="split_bf16")
    ap.add_argument("--compare", default="plain,mfma_f64")
    ap.add_argument("--block", type=int, default=64,
                    help="history block size 16/32/64/128; 0: autotune (timed trial of each, untimed)")
    ap.add_argument("--nodes", type=int, default=None)
    ap.add_argument("--msteps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--requests", type=int, default=512, help="serving mix size (--workload serve)")
    ap.add_argument("--rate", type=float, default=16.0, help="serving arrival rate, requests/s")
    ap.add_argument("--max-active", type=int, default=16)
    ap.add_argument("--count", type=int, default=6, help="solves in the batch (--workload ensemble)")
    ap.add_argument("--quantum", type=int, default=256, help="steps per scheduling round (--workload serve)")
    ap.add_argument("--uneven", action="store_true",
                    help="--workload ensemble: spread N over [nodes/4, nodes] and M over [M/2, M] too")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the six-solve batch measurement added to the headline line")
    ap.add_argument("--gamma", type=float, default=0.0,
                    help="space-fractional order in (0,1) (0: classical diffusion)")
    ap.add_argument("--picard", type=int, default=3,
                    help="Picard iterations per step on the nonlinear coefficient (1D; 3D LOD uses 1)")
    ap.add_argument("--picard-tol", type=float, default=1e-12,
                    help="Picard stops once max|U^k - U^(k-1)| <= tol")
    ap.add_argument("--calibrate", type=int, default=0,
                    help="first copy this many MiB device-to-device (known-size kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import fpde_hpc_amd as F
    if F.lib().fpde_device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if args.calibrate > 0:
        import ctypes
        L = F.lib()
        ctx, b0, b1 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nb = args.calibrate << 20
        F.check(L.fpde_cl_ctx_create(0, ctypes.byref(ctx)))
        F.check(L.fpde_cl_buf_alloc(ctx, nb, None, ctypes.byref(b0)))
        F.check(L.fpde_cl_buf_alloc(ctx, nb, None, ctypes.byref(b1)))
        for _ in range(3):
            F.check(L.fpde_cl_buf_copy(ctx, b1, 0, b0, 0, nb))
        L.fpde_cl_buf_free(b0)
        L.fpde_cl_buf_free(b1)
        L.fpde_cl_ctx_destroy(ctx)
        print(f"[bench] calibration copy: 3 x {args.calibrate} MiB device-to-device", file=sys.stderr, flush=True)
    if args.workload == "counters":
        # Short program for rocprofv3 --pmc passes: a known-size device copy
        # (counter calibration), the plain FP64 history stream (n=2000 levels,
        # N=10000), and a split-bf16 and an FP64-MFMA solve (N=10000, M=2000).
        import ctypes
        L = F.lib()
        ctx = ctypes.c_void_p()
        F.check(L.fpde_cl_ctx_create(0, ctypes.byref(ctx)))
        nb = 1 << 30
        bufs = [ctypes.c_void_p() for _ in range(5)]
        F.check(L.fpde_cl_buf_alloc(ctx, nb, None, ctypes.byref(bufs[0])))
        F.check(L.fpde_cl_buf_alloc(ctx, nb, None, ctypes.byref(bufs[1])))
        F.check(L.fpde_cl_buf_copy(ctx, bufs[1], 0, bufs[0], 0, nb))
        print("[counters] copy 1 GiB", file=sys.stderr, flush=True)
        n, nodes, ld = 2000, 10000, 10048
        D = np.random.default_rng(1).standard_normal((n, ld))
        w = np.random.default_rng(2).random(n + 8)
        F.check(L.fpde_cl_buf_alloc(ctx, D.nbytes, D.ctypes.data, ctypes.byref(bufs[2])))
        F.check(L.fpde_cl_buf_alloc(ctx, w.nbytes, w.ctypes.data, ctypes.byref(bufs[3])))
        F.check(L.fpde_cl_buf_alloc(ctx, ld * 8, None, ctypes.byref(bufs[4])))
        p = L.fpde_cl_buf_device_ptr
        F.check(L.fpde_history_sum_f64(p(bufs[4]), p(bufs[2]), n, nodes, ld, p(bufs[3])))
        print("[counters] plain history sum n=2000 N=10000", file=sys.stderr, flush=True)
        for mode in ("split_bf16", "mfma_f64"):
            spec, u0 = F.workload("1d_large", hist_mode=mode, steps=2000, picard_max=args.picard,
                                  picard_tol=args.picard_tol)
            F.solve(spec, u0)
            print(f"[counters] {mode} solve N=10000 M=2000", file=sys.stderr, flush=True)
        for bb in bufs:
            L.fpde_cl_buf_free(bb)
        L.fpde_cl_ctx_destroy(ctx)
        print(json.dumps({"counters_program": "done"}))
        return
    if args.workload == "ensemble":
        r = ensemble_run(F, args.count, args.nodes or 10000, args.msteps or 2000, args.headline,
                         (args.picard, args.picard_tol), reps=max(1, args.steps), uneven=args.uneven)
        print(json.dumps({
            "metric": "node*steps/s, ensemble of independent 1D solves in one batch",
            "value": r["node_steps_per_s"]["batch"], "unit": "node*steps/s", "n_gpus": 1, "steps": max(1, args.steps),
            "warmup": 0, "ms_per_step": r["seconds"]["batch"] * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": VARIANT_DTYPE.get(args.headline, "f64"), "data": "synthetic",
            "config": {"workload": f"{r['count']} x 1D N={r['nodes']}, M={r['steps']}, alpha 0.5..0.9",
                       "uneven": bool(args.uneven),
                       "parallelism": "single GPU, one cooperative launch per history block for the batch"},
            "ensemble": r}))
        return
    if args.workload == "serve":
        r = serve_mix(F, args.requests, args.rate, quantum=args.quantum, max_active=args.max_active,
                      hist_mode=args.headline)
        print(json.dumps({
            "metric": "p50/p99 solve latency and aggregate node*steps/s, serving mix (continuous batching)",
            "value": r["node_steps_per_s"], "unit": "node*steps/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": r["makespan_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": VARIANT_DTYPE.get(args.headline, "f64"), "data": "synthetic",
            "config": {"workload": "serving mix: Poisson arrivals, 80% 1D / 20% 3D, 25% distributed order",
                       "parallelism": "single GPU; 1D block-form requests share batched cooperative launches, "
                                      "the others one stream per request"},
            "serving": r}))
        return

    picard = (args.picard, args.picard_tol)
    # FP64 plain result first: the error reference for every other variant
    results = {}
    compare = [c for c in args.compare.split(",") if c]
    order = [m for m in ["plain"] if m in compare or args.headline == "plain"]
    order += [m for m in compare if m not in order and m != args.headline]
    order += [args.headline]
    for mode in order:
        K = args.steps if mode == args.headline else 1
        r = run_variant(F, args.workload, mode, args.block, K, args.warmup, prof_mask_timed=None,
                        nodes=args.nodes, msteps=args.msteps, picard=picard, gamma=args.gamma)
        # decide the dominant kernel class from the profiled warmup, then time
        # it live over the timed region with events on the solver stream
        prof = r["prof"]
        dom = 0 if prof["ms_history"] >= prof["ms_step"] else 1
        r2 = run_variant(F, args.workload, mode, r["spec"].block if args.block == 0 else args.block, K, 0,
                         prof_mask_timed=1 << dom,
                         nodes=args.nodes, msteps=args.msteps, picard=picard,
                         gamma=args.gamma) if mode == args.headline else None
        r["dominant"] = dom
        r["timed_live"] = r2
        results[mode] = r
        print(f"[bench] {mode}: solve {statistics.mean(r['times']):.3f}s "
              f"(hist {prof['ms_history']:.1f} ms / {prof['history_launches']} launches, "
              f"step {prof['ms_step']:.1f} ms / {prof['step_launches']} launches)", file=sys.stderr, flush=True)

    ref_u = results["plain"]["u"] if "plain" in results else None
    spec = results[args.headline]["spec"]
    N, M = spec.nodes, spec.steps
    unknowns = N if spec.dims == 1 else N ** 3
    variants = {}
    for mode, r in results.items():
        t = statistics.mean(r["times"])
        e = {}
        if ref_u is not None:
            d = np.abs(r["u"] - ref_u)
            e = {"rel_err_vs_fp64": float(d.max() / np.abs(ref_u).max()), "max_abs_err_vs_fp64": float(d.max())}
        p = r["prof"]
        variants[mode] = {
            "node_steps_per_s": unknowns * M / t, "ms_per_solve": t * 1e3, **e,
            "history_ms": p["ms_history"], "history_launches": p["history_launches"],
            "step_ms": p["ms_step"], "step_launches": p["step_launches"],
            "history_alg_bytes": p["history_alg_bytes"], "history_alg_flops": p["history_alg_flops"],
        }
        if p["ms_history"] > 0:
            if mode == "plain":
                variants[mode]["history_hbm_GBs"] = p["history_alg_bytes"] / (p["ms_history"] * 1e-3) / 1e9
            else:
                variants[mode]["history_TFLOPs"] = p["history_alg_flops"] / (p["ms_history"] * 1e-3) / 1e12

    head = results[args.headline]
    t_mean = statistics.mean(head["times"])
    # roofline of the dominant kernel class of the headline variant, from the
    # live-timed solves (events only around that class)
    live = head["timed_live"]
    ts = live["timed"]
    dom = head["dominant"]
    if dom == 0:
        ms = statistics.mean(x["ms_history"] for x in ts)
        nl = statistics.mean(x["history_launches"] for x in ts)
        if args.headline == "plain":
            alg = statistics.mean(x["history_alg_bytes"] for x in ts)
            achieved = alg / (ms * 1e-3) / 1e9
            roof = {"kernel": "k_hist_plain", "bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS,
                    "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                    "alg_bytes_per_launch": alg / max(nl, 1), "avg_launch_us": ms * 1e3 / max(nl, 1)}
        else:
            alg = statistics.mean(x["history_alg_flops"] for x in ts)
            pk = MI355X_PEAK_TFLOPS["f64" if args.headline == "mfma_f64" else "bf16"]
            achieved = alg / (ms * 1e-3) / 1e12
            roof = {"kernel": "k_gemm_f64" if args.headline == "mfma_f64" else "k_pack_split+k_gemm_split",
                    "bound": "mfma", "achieved": achieved, "peak": pk, "unit": "TFLOP/s",
                    "frac": achieved / pk, "traffic": None,
                    "alg_flops_per_launch": alg / max(nl, 1), "avg_launch_us": ms * 1e3 / max(nl, 1)}
    else:
        ms = statistics.mean(x["ms_step"] for x in ts)
        nl = statistics.mean(x["step_launches"] for x in ts)
        # step kernel algorithmic bytes per launch: U^n in, U^{n+1} and D out (24 B/node)
        # plus the near-field history terms it reads (8 B per term)
        # a launch covers M/launches steps (cooperative block launches: a
        # whole history block; otherwise one Picard iteration)
        steps_per_launch = M / max(nl, 1)
        bpl = 24.0 * unknowns * steps_per_launch
        achieved = bpl / (ms * 1e-3 / max(nl, 1)) / 1e9
        kname = ("k_step1d_block" if steps_per_launch > 1.5 else "k_step1d") if spec.dims == 1 else "lod3d sweeps"
        roof = {"kernel": kname, "bound": "hbm",
                "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                "alg_bytes_per_launch": bpl, "avg_launch_us": ms * 1e3 / max(nl, 1),
                "steps_per_launch": steps_per_launch,
                "note": "latency-bound sequential step chain (Picard iterations, grid barriers)"}

    # HBM traffic per launch of the dominant kernel, from the committed
    # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (calibrated on a copy)
    try:
        tr = json.load(open(os.path.join(ROOT, "profiles", "hbm_counters.json")))
        k = tr["kernels"].get(roof["kernel"])
        if k and spec.dims == 1 and N == tr["nodes"]:
            roof["traffic"] = k["fetch_bytes"] + k["write_bytes"]
            roof["traffic_source"] = "profiles/hbm_counters.json (rocprofv3 --pmc, copy-calibrated)"
    except (OSError, ValueError, KeyError):
        pass
    # the CPU baseline samples the classical scheme (the oracle's spatial sum
    # is a serial O(N^2) loop per step and is not part of the history path)
    cpu = None if (args.no_cpu or spec.dims != 1 or spec.do_q or spec.gamma) else cpu_baseline(
        N, M, alpha=spec.alpha, tau=spec.tau, picard=(spec.picard_max, spec.picard_tol))
    out = {
        "metric": "node*steps/s, mixed-MFMA history-sum solver (vs FP64 GPU at <=1e-8 rel err)",
        "value": unknowns * M / t_mean,
        "unit": "node*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mean * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": VARIANT_DTYPE.get(args.headline, "f64"),
        "data": "synthetic (seeded initial head, mt19937_64 seed 1001)",
        "config": {"workload": (f"{args.workload}: {'3D LOD ' + str(N) + '^3' if spec.dims == 3 else '1D N=' + str(N)} "
                                f"time-fractional moisture transport M={M} tau={spec.tau} "
                                + (f"distributed order q={spec.do_q} on [{spec.do_a0},{spec.do_a1}]" if spec.do_q
                                   else f"alpha={spec.alpha}") + f"; history {args.headline} block {spec.block}"
                                + (" (autotuned)" if args.block == 0 else "")
                                + (f"; Picard <= {spec.picard_max} iterations, tol {spec.picard_tol:g}"
                                   if spec.dims == 1 else "; LOD with lagged coefficients")
                                + (f"; space-fractional explicit scheme gamma={spec.gamma:g}" if spec.gamma else "")),
                   "nodes": N, "time_steps": M, "history_mode": args.headline, "block": spec.block,
                   "picard_max": spec.picard_max, "picard_tol": spec.picard_tol, "gamma": spec.gamma,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    if results[args.headline].get("tuned"):
        out["config"]["block_autotune"] = results[args.headline]["tuned"]
    if "plain" in variants:
        out["speedup_vs_plain_fp64"] = variants["plain"]["ms_per_solve"] / (t_mean * 1e3)
        if ref_u is not None:
            out["rel_err_vs_fp64"] = variants[args.headline].get("rel_err_vs_fp64")
    if not args.no_ensemble and args.workload == "1d_large" and args.nodes is None and spec.dims == 1:
        # independent solves of the same column (alpha 0.5..0.9, M=2000), one
        # batch against one after another: not part of the headline figure
        out["ensemble"] = ensemble_run(F, 6, N, 2000, args.headline, picard, reps=3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
