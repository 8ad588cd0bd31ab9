"""The facade's MatrixMarket reader (mcls_amd/host/MCLS_MatrixMarket.hpp)
against scipy.io.mmread on the same files, its error handling, and a
vector write/read round trip.  CPU only: the driver links no GPU library."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "bin", "test_mm")
DATA = os.path.join(ROOT, "tests", "data", "mm")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", ROOT, "tests/cpp/bin/test_mm"], check=True, capture_output=True)
    return DRIVER


def _run(driver, *args):
    return subprocess.run([driver, *args], capture_output=True, text=True)


def _read_matrix(driver, name):
    out = _run(driver, "matrix", os.path.join(DATA, name))
    assert out.returncode == 0, out.stdout
    lines = out.stdout.split("\n")
    n, nnz = map(int, lines[0].split())
    M = np.zeros((n, n))
    rows = []
    for ln in lines[1:1 + nnz]:
        i, j, v = ln.split()
        M[int(i), int(j)] += float(v)
        rows.append((int(i), int(j)))
    return M, rows


@pytest.mark.parametrize("name", ["general.mtx", "symmetric.mtx", "skew.mtx", "pattern.mtx"])
def test_matrix_matches_scipy(driver, name):
    M, rows = _read_matrix(driver, name)
    ref = scipy.io.mmread(os.path.join(DATA, name))
    ref = np.asarray(ref.todense() if hasattr(ref, "todense") else ref, dtype=float)
    assert np.array_equal(M, ref)
    # CSR rows ascending in column, one entry per (i, j)
    assert rows == sorted(rows) and len(set(rows)) == len(rows)


def test_repeated_entries_are_summed(driver):
    M, _ = _read_matrix(driver, "general.mtx")
    assert M[1, 2] == -0.75


def test_vector_and_round_trip(driver, tmp_path):
    out = _run(driver, "vector", os.path.join(DATA, "vector.mtx"))
    assert out.returncode == 0
    vals = [float(x) for x in out.stdout.split()[1:]]
    ref = np.asarray(scipy.io.mmread(os.path.join(DATA, "vector.mtx"))).ravel()
    assert np.array_equal(vals, ref)
    rt = _run(driver, "roundtrip", os.path.join(DATA, "vector.mtx"), str(tmp_path / "v.mtx"))
    assert rt.returncode == 0 and rt.stdout.strip() == "same"


@pytest.mark.parametrize("name,msg", [("bad_square.mtx", "square"), ("bad_range.mtx", "out of range"),
                                      ("bad_truncated.mtx", "ends after"), ("bad_banner.mtx", "banner"),
                                      ("bad_complex.mtx", "complex")])
def test_errors(driver, name, msg):
    out = _run(driver, "matrix", os.path.join(DATA, name))
    assert out.returncode == 2 and out.stdout.startswith("ERROR") and msg in out.stdout
