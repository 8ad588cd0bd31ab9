"""mcls_amd: MI355X (gfx950) adjoint Neumann-Ulam / MCSA solver.

The product is the C ABI library mcls_amd/lib/libmcls_gpu.so (include/mcls_gpu.h)
and the C++ host facade under mcls_amd/host/ that mirrors the MCLS API
(LinearProblem, MCSASolverManager, MonteCarloSolverManager, SolverFactory,
MatrixTraits/VectorTraits).  This Python package only exposes ctypes bindings
for tests and benchmarks.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
