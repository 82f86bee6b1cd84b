"""bench.py's output contract on a small lattice: one JSON line on stdout with the metric,
value, timing and configuration keys, the roofline object of the fused kernel and the CPU
baseline (a short sample). The other bench lines are switched off to keep it short."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line(gpu):
    print("bench.py on SU(3) 16^3, 4 steps", flush=True)
    cmd = [sys.executable, "bench.py", "--dims", "16", "16", "16", "--steps", "4", "--warmup", "1",
           "--kernel-reps", "3", "--cpu-seconds", "0.5", "--no-autotune", "--no-cpic", "--no-batch", "--no-adjoint",
           "--no-implicit", "--no-cgc"]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 4 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["unit"] == "link-updates/s"
    assert d["config"]["lattice"] == [16, 16, 16] and "workload" in d["config"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1
