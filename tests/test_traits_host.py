"""The facade's MatrixTraits / VectorTraits members and LinearProblem's
operator family against dense arithmetic: tests/cpp/bin/test_traits, a
host-only program (no GPU library)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_traits")


def test_traits_program():
    assert os.path.exists(BIN), "build first"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=60, cwd=ROOT)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL" in p.stdout
