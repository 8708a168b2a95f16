"""The bench.py contract: one JSON line with the fields the driver reads, the
single-GPU refusal, and the CPU baseline object."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_cpu_baseline_pieces(monkeypatch):
    """The host share and CPU model, and the sampled history term of one step
    (per order; a node share scaled up)."""
    import bench
    from oracle import oracle as O
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    cores, model, ncpu = bench.cpu_info()
    assert cores == min(3, ncpu) and ncpu >= 1 and model
    t1, reps, total = bench.cpu_history_step(O, 1, 400, 200, 1e-3, [0.8], seconds=0.1)
    t2, _, _ = bench.cpu_history_step(O, 1, 400, 200, 1e-3, [0.8, 0.5], node_frac=0.5, seconds=0.1)
    assert t1 > 0 and t2 > 0 and reps >= 3 and total > 0


@pytest.mark.gpu
def test_bench_json_line():
    """A short run of the headline workload (200 steps): every field of the
    line, the roofline object of the dominant kernel, and the error of the
    mixed form against the FP64 stream."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--msteps", "200", "--steps", "1",
                        "--warmup", "0", "--no-cpu", "--compare", "plain"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.strip().splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in out, k
    assert out["n_gpus"] == 1 and out["steps"] == 1 and out["warmup"] == 0
    assert out["value"] > 0 and out["higher_is_better"] is True and out["vs_baseline"] is None
    assert out["config"]["time_steps"] == 200 and "workload" in out["config"]
    roof = out["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in roof, k
    assert roof["bound"] in ("hbm", "mfma") and 0 < roof["frac"] <= 1.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-9
    if roof["kernel"] == "k_step1d_block":         # 100 B/node*step + 8 B per near-field term
        assert 100.0 < roof["alg_bytes_per_node_step"] < 100.0 + 8.0 * 129
    lat = out["latency"]
    assert lat["us_per_step"] > 0 and lat["step_us_per_step"] > 0 and 1.0 <= lat["picard_iterations_per_step"] <= 3.0
    assert out["rel_err_vs_fp64"] <= 1e-8
    ens = out["ensemble"]                          # six solves as one batch vs one after another
    assert ens["count"] == 6 and ens["batch_max_rel_diff_vs_solo"] == 0.0
    assert ens["batch_speedup_vs_sequential"] > 1.0
