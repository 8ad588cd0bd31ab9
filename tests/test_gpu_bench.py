"""bench.py keeps its output contract: one JSON line with the metric, the
whole-job value, the roofline and CPU-baseline objects, on the small 64^3
version of the benchmark operator; more than one GPU is refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_bench_json_contract():
    p = subprocess.run([sys.executable, "bench.py", "--workload", "cfg4-64", "--steps", "1", "--warmup", "0",
                        "--no-tune"], capture_output=True, text=True, timeout=110, cwd=ROOT)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 0
    assert d["value"] > 0 and d["unit"] == "histories/s" and d["higher_is_better"] is True
    assert d["config"]["workload"].startswith("cfg4-64") and d["vs_baseline"] is None
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert 0 < r["frac"] < 1 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-12
    c = d["cpu_baseline"]
    assert c["value"] > 0 and c["cores"] >= 1 and c["kind"] in ("port", "reference") and c["sample"]
    assert d["achieved_tol"] <= 1e-8
    assert d["learned_proposal"]["choice"] in ("plain", "diagonal", "network")
    assert d["plain"]["achieved_tol"] <= 1e-8 and d["speedup_vs_plain"] > 0


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True, timeout=60,
                       cwd=ROOT)
    assert p.returncode == 2 and "one GPU" in p.stderr
