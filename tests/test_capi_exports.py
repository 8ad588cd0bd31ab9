"""The built libraries load and export every symbol the headers declare
(no compute calls: runs without a GPU)."""
import ctypes as C
import os
import re

from mcls_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mcls_[a-z_0-9]+)\(", text, re.M)))


def test_gpu_lib_exports_every_declared_symbol():
    decl = _declared("mcls_gpu.h")
    assert len(decl) >= 20
    assert sorted(capi.GPU_SYMBOLS) == decl
    lib = C.CDLL(capi.GPU_LIB)
    for name in decl:
        assert hasattr(lib, name), name


def test_mlp_exports_every_declared_symbol():
    decl = _declared("mcls_mlp.h")
    assert sorted(capi.MLP_SYMBOLS) == decl
    lib = C.CDLL(capi.GPU_LIB)
    for name in decl:
        assert hasattr(lib, name), name
    text = open(os.path.join(ROOT, "include", "mcls_mlp.h")).read()
    assert f"MCLS_MLP_PARAMS {capi.MLP_PARAMS}" in text


def test_problem_lib_exports_every_declared_symbol():
    decl = _declared("mcls_problems.h")
    assert sorted(capi.PROBLEM_SYMBOLS) == decl
    lib = C.CDLL(capi.PROBLEMS_LIB)
    for name in decl:
        assert hasattr(lib, name), name


def test_no_device_is_a_clean_error():
    import pytest

    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    lib = capi.gpu_lib()
    h = C.c_void_p()
    rc = lib.mcls_gpu_create(C.byref(h), 0, 0)
    assert rc == capi.MCLS_ERR_NO_DEVICE
    assert b"device" in lib.mcls_gpu_last_error()


def test_facade_binary_built():
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "bin", "test_facade"))
