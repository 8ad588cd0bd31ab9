"""The CPU baseline harness (oracle/cpu_baseline.py): reference-semantics
MCSA over several sets (threads), its fields, the quirk switch, and the
extrapolation arithmetic.  CPU only."""
import json
import os
import subprocess
import sys

import numpy as np

from mcls_amd import capi
from oracle import cpu_baseline as cb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _small():
    return capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 10, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2))


def test_full_solve_converges_over_sets():
    prob = _small()
    r = cb.run(prob, False, 16 * prob.n, cores=3, iters=60, full=True)
    assert r["converged"] and r["achieved_tol"] <= 1e-8
    assert r["sets"] == 3 and r["histories_per_set"] == 16 * prob.n // 3
    assert r["residual_history"][0] == 1.0 and r["residual_history"][-1] <= 1e-8
    assert r["histories_per_s"] > 0 and r["steps_per_history"] > 1


def test_extrapolation_and_quirk():
    prob = _small()
    r = cb.run(prob, False, 4000, cores=2, iters=2, workload_histories=1 << 20)
    assert not r["converged"] and r["iterations"] == 2
    ex = r["extrapolated"]
    assert np.isclose(ex["seconds_mc_per_iteration"], (1 << 20) / r["histories_per_s"])
    # the signed b[0] seed of the reference's source CDF: with a negative
    # source (b < 0 stays < 0 through the Richardson step, H >= 0 here) the
    # CDF starts below zero and the reference emits fewer histories
    prob.b[:] = -np.abs(prob.b) - 0.1
    a = cb.run(prob, False, 4000, cores=2, iters=1, quirk=False, workload_histories=4000)
    q = cb.run(prob, False, 4000, cores=2, iters=1, quirk=True, workload_histories=4000)
    assert q["reference_quirk"] and not a["reference_quirk"]
    assert q["histories"] < a["histories"]


def test_cli_prints_one_json_line():
    out = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--workload", "cfg4-64", "--iters", "1",
                          "--cores", "2", "--histories", "20000", "--time-to-tol-iterations", "7"],
                         capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert out.returncode == 0, out.stderr
    d = json.loads(out.stdout.strip().splitlines()[-1])
    assert d["workload"] == "cfg4-64" and d["cores"] == 2 and d["n"] == 64 ** 3
    assert np.isclose(d["extrapolated"]["time_to_tol_s"], 7 * d["extrapolated"]["seconds_mc_per_iteration"])
