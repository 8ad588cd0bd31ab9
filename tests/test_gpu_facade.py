"""The C++ facade (mcls_amd/host) on the GPU: runs tests/cpp/bin/test_facade,
which transcribes the reference's MCSASolverManager / MonteCarloSolverManager
/ SolverFactory tests."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_facade")


@pytest.mark.gpu
def test_facade_program():
    assert os.path.exists(BIN), "build first"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=110, cwd=ROOT)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL" in p.stdout
