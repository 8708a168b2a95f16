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


def cpu_info():
    """Threads for the OpenMP oracle and the host CPU model.  The threads are
    the host share this process is given (OMP_NUM_THREADS; on the GPU box
    that is the CPU share of one GPU), else every core of the host."""
    ncpu = os.cpu_count() or 1
    try:
        cores = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu))))
    except ValueError:
        cores = ncpu
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model, ncpu


def _median_time(fn, min_reps, seconds):
    reps = []
    t_start = time.perf_counter()
    while len(reps) < min_reps or time.perf_counter() - t_start < seconds:
        t0 = time.perf_counter()
        fn()
        reps.append(time.perf_counter() - t0)
    return statistics.median(reps), len(reps), time.perf_counter() - t_start


def cpu_history_step(O, cores, nodes, n, tau, orders, node_frac=1.0, seconds=5.0):
    """Oracle time of the history term of one step at history length n (the
    run average n = M/2): per order the weights kb_3_cache of that step
    (precalc_kb_3) and time_summing over the nodes (reference loop order:
    node-outer, t-inner, OpenMP over nodes).  With node_frac < 1 the sum runs
    on that share of the nodes and is scaled up (the sum is independent per
    node).  Returns (seconds per step, repetitions, sampled seconds)."""
    import numpy as np
    ns = max(1, int(round(nodes * node_frac)))
    rng = np.random.default_rng(0)
    old = np.cumsum(rng.standard_normal((n + 1, ns)) * 1e-4, axis=0)
    Time = O.time_grid(n + 1, tau)
    taus = np.full(n + 2, tau)

    def one():
        for a in orders:
            _, kb = O.l1_weights(a, n, Time, weight_mode=1)
            O.time_summing(old, kb, taus, nthreads=cores)
    t, reps, total = _median_time(one, 3, seconds)
    return t / node_frac, reps, total


def cpu_baselines(F, picard):
    """The FP64 oracle (oracle/, restating the reference path) timed on this
    host: config 0 in full (median of 3 solves, and its result, against which
    the GPU results of config 0 are compared); configs 1-3 by sampling one
    step at the run-average history length n = M/2 (history term + one step
    of coefficients and line solves), extrapolated to node*steps/s.  The
    cpu_baseline object of the line is config 1's."""
    import numpy as np
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle as O
    O.lib()
    cores, model, ncpu = cpu_info()
    out = {"cores": cores, "host_cpus": ncpu, "cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}
    # config 0: full solves
    spec0, u0 = F.workload("1d_small", picard_max=picard[0], picard_tol=picard[1])
    p0 = O.make_params(spec0.nodes, spec0.steps, spec0.tau, spec0.alpha, ksat=spec0.ksat, picard_max=picard[0],
                       picard_tol=picard[1], nthreads=cores)
    res = {}

    def full0():
        res["u"] = O.solve_1d(p0, u0)
    t0, reps0, _ = _median_time(full0, 3, 0.0)
    out["config0"] = {"value": spec0.nodes * spec0.steps / t0, "unit": "node*steps/s", "seconds": t0,
                      "sample": f"full solve N={spec0.nodes} M={spec0.steps}, median of {reps0}"}
    print(f"[cpu] config 0 full solve {t0:.2f} s ({cores} threads)", file=sys.stderr, flush=True)
    # config 1: history at n = M/2 over all nodes + one step
    spec1, u1 = F.workload("1d_large", picard_max=picard[0], picard_tol=picard[1])
    N1, M1 = spec1.nodes, spec1.steps
    th, reps, tot = cpu_history_step(O, cores, N1, M1 // 2, spec1.tau, [spec1.alpha], seconds=10.0)
    p1 = O.make_params(N1, 4, spec1.tau, spec1.alpha, ksat=spec1.ksat, nthreads=cores, picard_max=picard[0],
                       picard_tol=picard[1])
    ts, sreps, _ = _median_time(lambda: O.solve_1d(p1, u1), 3, 1.0)
    ts /= 4
    out["config1"] = {"value": N1 / (th + ts), "unit": "node*steps/s", "seconds_per_step": th + ts,
                      "sample": (f"history term at n={M1 // 2} levels x N={N1} nodes (the run average for "
                                 f"M={M1}): {reps} reps in {tot:.1f} s, median {th * 1e3:.2f} ms; + one step "
                                 f"{ts * 1e3:.3f} ms (median of {sreps} 4-step solves); extrapolated to the run")}
    print(f"[cpu] config 1 sampled: {th * 1e3:.1f} ms history + {ts * 1e3:.2f} ms step", file=sys.stderr, flush=True)
    # config 2: distributed order, 32 Gauss-Legendre orders (the oracle sums per order)
    spec2, u2 = F.workload("do", picard_max=picard[0], picard_tol=picard[1])
    al, _ = O.gauss_legendre(spec2.do_q, spec2.do_a0, spec2.do_a1)
    th2, reps2, tot2 = cpu_history_step(O, cores, spec2.nodes, spec2.steps // 2, spec2.tau, list(al), seconds=5.0)
    p2 = O.make_params(spec2.nodes, 4, spec2.tau, spec2.alpha, ksat=spec2.ksat, nthreads=cores,
                       picard_max=picard[0], picard_tol=picard[1], do_q=spec2.do_q, do_rule=spec2.do_rule,
                       do_a0=spec2.do_a0, do_a1=spec2.do_a1)
    ts2, sreps2, _ = _median_time(lambda: O.solve_1d(p2, u2), 3, 1.0)
    ts2 /= 4
    out["config2"] = {"value": spec2.nodes / (th2 + ts2), "unit": "node*steps/s",
                      "seconds_per_step": th2 + ts2,
                      "sample": (f"history term of {spec2.do_q} orders at n={spec2.steps // 2} levels x "
                                 f"N={spec2.nodes}: {reps2} reps in {tot2:.1f} s, median {th2 * 1e3:.1f} ms; + one "
                                 f"step {ts2 * 1e3:.2f} ms (median of {sreps2} 4-step solves); extrapolated")}
    print(f"[cpu] config 2 sampled: {th2 * 1e3:.1f} ms history + {ts2 * 1e3:.2f} ms step", file=sys.stderr,
          flush=True)
    # config 3: 128^3, history on 1/8 of the nodes (scaled), one LOD step
    spec3, u3 = F.workload("3d")
    nn = spec3.nodes ** 3
    th3, reps3, tot3 = cpu_history_step(O, cores, nn, spec3.steps // 2, spec3.tau, [spec3.alpha], node_frac=0.125,
                                        seconds=5.0)
    p3 = O.make_params(spec3.nodes, 2, spec3.tau, spec3.alpha, ksat=spec3.ksat, nthreads=cores)
    ts3, sreps3, _ = _median_time(lambda: O.solve_3d(p3, u3), 1, 0.0)
    ts3 /= 2
    out["config3"] = {"value": nn / (th3 + ts3), "unit": "node*steps/s", "seconds_per_step": th3 + ts3,
                      "sample": (f"history term at n={spec3.steps // 2} levels on 1/8 of the {nn} nodes, scaled: "
                                 f"{reps3} reps in {tot3:.1f} s, median {th3 * 1e3:.0f} ms; + one LOD step "
                                 f"{ts3 * 1e3:.0f} ms (a 2-step solve); extrapolated")}
    print(f"[cpu] config 3 sampled: {th3 * 1e3:.0f} ms history + {ts3 * 1e3:.0f} ms step", file=sys.stderr,
          flush=True)
    return out, res["u"]


def config_runs(F, picard, u_cpu0):
    """One MI355X solve of each of configs 0, 2 and 3 in the plain FP64 form
    and the mixed form (after one untimed solve each), with the mixed result's
    error against the FP64 GPU result, and for config 0 both against the
    oracle's full CPU solve.  Outside the timed region of the headline."""
    import numpy as np
    out = {}
    # history blocks: 64 for the 1D forms; 64 for 3D as well, where the near
    # field is folded in 16-level groups (0.167 s against 0.171 s at 32)
    for key, name, mixed, block in (("config0", "1d_small", "split_bf16", 64), ("config2", "do", "split_f16", 64),
                                    ("config3", "3d", "split_bf16", 64)):
        rs = {}
        for mode in ("plain", mixed):
            r = run_variant(F, name, mode, block, 1, 1, picard=picard)
            rs[mode] = r
        spec = rs[mixed]["spec"]
        unknowns = spec.nodes if spec.dims == 1 else spec.nodes ** 3
        ref = rs["plain"]["u"]
        e = {}
        for mode, r in rs.items():
            t = r["times"][0]
            e[mode] = {"seconds": t, "node_steps_per_s": unknowns * spec.steps / t}
        d = np.abs(rs[mixed]["u"] - ref)
        e[mixed]["rel_err_vs_fp64_gpu"] = float(d.max() / np.abs(ref).max())
        e[mixed]["max_abs_err_vs_fp64_gpu"] = float(d.max())
        if key == "config0" and u_cpu0 is not None:
            for mode, r in rs.items():
                dc = np.abs(r["u"] - u_cpu0)
                e[mode]["rel_err_vs_cpu_oracle"] = float(dc.max() / np.abs(u_cpu0).max())
        out[key] = {"workload": name, "nodes": spec.nodes, "dims": spec.dims, "time_steps": spec.steps,
                    "mixed": mixed, "block": spec.block, "variants": e,
                    "speedup_mixed_vs_plain": e["plain"]["seconds"] / e[mixed]["seconds"]}
        print(f"[configs] {key}: plain {e['plain']['seconds']:.3f} s, {mixed} {e[mixed]['seconds']:.3f} s, "
              f"rel err {e[mixed]['rel_err_vs_fp64_gpu']:.2e}", file=sys.stderr, flush=True)
    return out


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
    # 256-node chunks: fewer workgroups per solve, the faster form in a batch
    specs = [F.make_spec(n, m, 1e-3, alpha=float(a), hist_mode=hist_mode, block=64,
                         picard_max=picard[0], picard_tol=picard[1], u_top=u0[0], chunk_rows=256)
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


def serve_mix(F, n_req, rate, seed=1004, quantum=256, budget_gb=200.0, max_active=16, hist_mode="split_bf16"):
    """The serving mix (fpde_hpc_amd.serving_mix): requests submitted at their
    Poisson arrival times while the server runs; latency = submit to
    retirement on the server clock."""
    import numpy as np
    reqs = F.serving_mix(n_req, rate, seed, hist_mode)
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
    ap.add_argument("--no-configs", action="store_true", help="skip the configs 0/2/3 solves of the default run")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving mix of the default run")
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
        # a launch covers M/launches steps (cooperative block launches: a
        # whole history block; otherwise one Picard iteration).  Algorithmic
        # bytes (counted by the solver per step): 100 B per node for the step
        # itself (coefficients, rows, line solve) + 8 B per near-field history
        # term the step kernel sums
        steps_per_launch = M / max(nl, 1)
        alg = statistics.mean(x["step_alg_bytes"] for x in ts)
        bpl = alg / max(nl, 1)
        achieved = alg / (ms * 1e-3) / 1e9
        kname = ("k_step1d_block" if steps_per_launch > 1.5 else "k_step1d") if spec.dims == 1 else "lod3d sweeps"
        roof = {"kernel": kname, "bound": "hbm",
                "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                "alg_bytes_per_launch": bpl, "avg_launch_us": ms * 1e3 / max(nl, 1),
                "steps_per_launch": steps_per_launch,
                "alg_bytes_per_node_step": bpl / (unknowns * max(steps_per_launch, 1e-9)),
                "alg_bytes_rule": "100 B per node and step (coefficients, rows, line solve) + 8 B per near-field "
                                  "history term summed inside the step kernel",
                "note": "latency-bound sequential step chain (Picard iterations, grid barriers); see latency"}

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
    # time per step of the headline solve and where it goes (profiled solve:
    # device time of the history kernels and of the step kernels per step)
    hp = head["prof"]
    latency = {"us_per_step": t_mean * 1e6 / M,
               "history_us_per_step": hp["ms_history"] * 1e3 / M, "step_us_per_step": hp["ms_step"] * 1e3 / M,
               "step_launch_us": roof.get("avg_launch_us") if dom == 1 else None,
               "picard_iterations_per_step": hp["picard_iters_total"] / max(M, 1)}
    default_run = (args.workload == "1d_large" and args.nodes is None and args.msteps is None and not args.gamma)
    # CPU baselines (the oracle on this host's cores, full config 0, sampled
    # configs 1-3) of the classical scheme
    cpu, cpu_all, u_cpu0 = None, None, None
    if default_run and not args.no_cpu:
        cpu_all, u_cpu0 = cpu_baselines(F, picard)
        c1 = cpu_all["config1"]
        cpu = {"value": c1["value"], "unit": c1["unit"], "cores": cpu_all["cores"], "kind": "port",
               "cpu_model": cpu_all["cpu_model"], "sample": c1["sample"]}
    out = {
        "metric": "node*steps/s mixed-MFMA vs FP64 GPU at <=1e-8 rel err; p50/p99 solve latency",
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
        "latency": latency,
        "variants": variants,
    }
    if spec.dims == 1:
        out["config"]["chunk_nodes"] = spec.chunk_rows if spec.chunk_rows else "automatic (192)"
    if results[args.headline].get("tuned"):
        out["config"]["block_autotune"] = results[args.headline]["tuned"]
    if "plain" in variants:
        out["speedup_vs_plain_fp64"] = variants["plain"]["ms_per_solve"] / (t_mean * 1e3)
        pv = variants["plain"]
        out["speedup_note"] = (
            f"plain runs its FP64 history stream between every two steps, so its steps launch per Picard "
            f"iteration ({pv['step_launches']} step launches, {pv['step_ms']:.0f} ms of step time); the "
            f"{args.headline} form runs one cooperative launch per history block "
            f"({variants[args.headline]['step_launches']} launches, {variants[args.headline]['step_ms']:.0f} ms)")
        if ref_u is not None:
            out["rel_err_vs_fp64"] = variants[args.headline].get("rel_err_vs_fp64")
    if cpu_all is not None:
        out["cpu_baselines"] = cpu_all
    if default_run and not args.no_configs:
        out["configs"] = config_runs(F, picard, u_cpu0)
    if default_run and not args.no_serve:
        out["serving"] = serve_mix(F, args.requests, args.rate, quantum=args.quantum, max_active=args.max_active,
                                   hist_mode=args.headline)
        sv = out["serving"]
        print(f"[serve] {sv['completed']}/{sv['requests']} done, p50 {sv['p50_latency_s']:.2f} s, "
              f"p99 {sv['p99_latency_s']:.2f} s, {sv['node_steps_per_s']:.3e} node*steps/s", file=sys.stderr,
              flush=True)
    if not args.no_ensemble and args.workload == "1d_large" and args.nodes is None and spec.dims == 1:
        # independent solves of the same column (alpha 0.5..0.9, M=2000), one
        # batch against one after another: not part of the headline figure
        out["ensemble"] = ensemble_run(F, 6, N, 2000, args.headline, picard, reps=3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
