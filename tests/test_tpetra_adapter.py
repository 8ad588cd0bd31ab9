"""The Tpetra binding (mcls_amd/host/MCLS_TpetraAdapter.hpp) against a double
of the Tpetra API it calls (tests/cpp/tpetra_double.hpp; Trilinos is not
installed here, so the binding is not checked against Trilinos itself):
tests/cpp/bin/test_tpetra_adapter.  On the CPU, every trait member against
the facade's own CSR traits and the reference's vector known answers; on the
GPU (--gpu), both managers solve through the binding and match the facade's
own types."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_tpetra_adapter")


def _run(args, timeout):
    assert os.path.exists(BIN), "build first"
    p = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL" in p.stdout
    return p.stdout


def test_tpetra_traits_on_host():
    out = _run([], 60)
    assert "ok Tpetra matrix traits" in out and "ok Tpetra vector traits" in out


@pytest.mark.gpu
def test_tpetra_binding_managers_on_gpu():
    out = _run(["--gpu"], 110)
    assert "ok Tpetra binding on the GPU" in out
