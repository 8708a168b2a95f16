#!/usr/bin/env python
"""Benchmark of the Caputo L1 history-sum solver on one MI355X.

Workload (default): 1D time-fractional moisture transport, N=10000 nodes,
M=20000 steps, tau=5e-4, alpha=0.7, seeded initial head (seed 1001); FP64
weights, FP64 partitioned-Thomas step, lagged VGM coefficients.  A bench
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


def run_variant(F, workload, mode, block, steps, warmup, prof_mask_timed=None, nodes=None, msteps=None):
    spec, u0 = F.workload(workload, hist_mode=mode, block=block, nodes=nodes, steps=msteps)
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
    return dict(spec=spec, u=u_last, u_first=u, times=times, prof=prof, timed=timed_stats)


def cpu_baseline(N, M, seconds_hint=20.0):
    """FP64 oracle (oracle/, the reference loop order: node-outer, t-inner,
    OpenMP over nodes) timed on this host: the history sum at the run-average
    history length n = M/2 over all N nodes, plus one oracle step (VGM
    coefficients + Thomas), median of repetitions; node*steps/s =
    N / (t_sum + t_step)."""
    import numpy as np
    from oracle import oracle as O
    O.lib()
    cores = min(16, os.cpu_count() or 1)
    n = M // 2
    rng = np.random.default_rng(0)
    old = np.cumsum(rng.standard_normal((n + 1, N)) * 1e-4, axis=0)
    Time = O.time_grid(n + 1, 5e-4)
    _, kb = O.l1_weights(0.7, n, Time, weight_mode=1)
    tau = np.full(n + 2, 5e-4)
    reps = []
    t_start = time.perf_counter()
    while len(reps) < 3 or (time.perf_counter() - t_start < seconds_hint and len(reps) < 10):
        t0 = time.perf_counter()
        O.time_summing(old, kb, tau, nthreads=cores)
        reps.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > seconds_hint * 2:
            break
    t_sum = statistics.median(reps)
    del old
    u0 = O.init_1d(N, 1001)
    p = O.make_params(N, 4, 5e-4, 0.7, u_top=u0[0], nthreads=cores)
    t0 = time.perf_counter()
    O.solve_1d(p, u0)
    t_step = (time.perf_counter() - t0) / 4
    return {
        "value": N / (t_sum + t_step),
        "unit": "node*steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle time_summing over n={n} levels x N={N} nodes (run-average history "
                   f"length for M={M}), {len(reps)} reps median {t_sum:.3f}s, + oracle step "
                   f"{t_step * 1e3:.2f} ms; extrapolated to the whole run"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves of the headline variant")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1d_large")
    ap.add_argument("--headline", default="split_bf16")
    ap.add_argument("--compare", default="plain,mfma_f64")
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--nodes", type=int, default=None)
    ap.add_argument("--msteps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
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

    # FP64 plain result first: the error reference for every other variant
    results = {}
    compare = [c for c in args.compare.split(",") if c]
    order = [m for m in ["plain"] if m in compare or args.headline == "plain"]
    order += [m for m in compare if m not in order and m != args.headline]
    order += [args.headline]
    for mode in order:
        K = args.steps if mode == args.headline else 1
        r = run_variant(F, args.workload, mode, args.block, K, args.warmup, prof_mask_timed=None,
                        nodes=args.nodes, msteps=args.msteps)
        # decide the dominant kernel class from the profiled warmup, then time
        # it live over the timed region with events on the solver stream
        prof = r["prof"]
        dom = 0 if prof["ms_history"] >= prof["ms_step"] else 1
        r2 = run_variant(F, args.workload, mode, args.block, K, 0, prof_mask_timed=1 << dom,
                         nodes=args.nodes, msteps=args.msteps) if mode == args.headline else None
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
        bpl = 24.0 * unknowns
        achieved = bpl / (ms * 1e-3 / max(nl, 1)) / 1e9
        roof = {"kernel": "k_step1d" if spec.dims == 1 else "lod3d sweeps", "bound": "hbm",
                "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                "alg_bytes_per_launch": bpl, "avg_launch_us": ms * 1e3 / max(nl, 1),
                "note": "latency-bound sequential step chain"}

    cpu = None if args.no_cpu else cpu_baseline(N, M)
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
        "config": {"workload": f"{args.workload}: 1D time-fractional moisture transport N={N} M={M} "
                               f"tau={spec.tau} alpha={spec.alpha}; history {args.headline} block {args.block}",
                   "nodes": N, "time_steps": M, "history_mode": args.headline, "block": args.block,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    if "plain" in variants:
        out["speedup_vs_plain_fp64"] = variants["plain"]["ms_per_solve"] / (t_mean * 1e3)
        if ref_u is not None:
            out["rel_err_vs_fp64"] = variants[args.headline].get("rel_err_vs_fp64")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
