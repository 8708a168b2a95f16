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


def test_cpu_baseline_object():
    import bench
    cpu = bench.cpu_baseline(200, 400, seconds_hint=0.2, alpha=0.8, tau=1e-3, seed=1000, picard=(3, 1e-12))
    assert cpu["value"] > 0 and cpu["unit"] == "node*steps/s"
    assert cpu["kind"] == "port" and cpu["cores"] >= 1
    assert "n=200 levels x N=200 nodes" in cpu["sample"] and "reps" in cpu["sample"]


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
    assert out["rel_err_vs_fp64"] <= 1e-8
    ens = out["ensemble"]                          # six solves as one batch vs one after another
    assert ens["count"] == 6 and ens["batch_max_rel_diff_vs_solo"] == 0.0
    assert ens["batch_speedup_vs_sequential"] > 1.0
