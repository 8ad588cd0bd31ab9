"""Edge cases of the path on the GPU: the homogeneous system (b = 0,
src/MCLS_MCSASolverManager_impl.hpp:300-318 replaces the zero source norm by
1), a 1 x 1 system whose domain row is empty, absorbing states inside a
larger system (columns of A with only their diagonal), and a source that
is zero except on one state."""
import numpy as np
import pytest
import scipy.sparse as sp

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FORMS = [
    {"Walk Precision": "FP64", "Source Sampling Type": "Random", "Transition Sampler": "CDF"},
    {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "Alias"},
]


def _problem(S, b):
    S = S.tocsr()
    S.sort_indices()
    return capi.Problem(None, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(np.float64),
                        None, np.asarray(b, np.float64))


@pytest.mark.parametrize("form", range(len(FORMS)))
def test_homogeneous_system(form):
    n = 50
    S = sp.diags([np.full(n - 1, -1.0), np.full(n, 3.0), np.full(n - 1, -1.0)], [-1, 0, 1])
    prob = _problem(S, np.zeros(n))
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(FORMS[form], **{"Number of Histories": 1000, "Convergence Tolerance": 1e-8,
                                            "Maximum Iterations": 5, "History Length": 50}))
        x = ctx.mcsa_solve(prob.b)
        st = ctx.stats()
        assert np.array_equal(x, np.zeros(n))
        assert st.converged == 1 and st.iterations == 1 and st.source_norm == 1.0
        assert np.array_equal(ctx.mc_solve(prob.b), np.zeros(n))


@pytest.mark.parametrize("form", range(len(FORMS)))
def test_one_by_one_system(form):
    """H = I - a^-1 a = 0: the domain row is empty, every history tallies its
    birth and stops, and x = b / a exactly."""
    prob = _problem(sp.csr_matrix(np.array([[2.0]])), [1.0])
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(FORMS[form], **{"Number of Histories": 4096, "Convergence Tolerance": 1e-12,
                                            "Maximum Iterations": 3, "History Length": 10}))
        assert ctx.mc_solve(prob.b)[0] == 0.5
        x = ctx.mcsa_solve(prob.b)
        st = ctx.stats()
        assert x[0] == 0.5 and st.converged == 1 and st.steps == 0


def _absorbing(n=120, seed=4):
    """Columns 0, 7, 14, ... of A hold only their diagonal: those domain rows
    (rows of H^T) are empty and walks that reach them stop there."""
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=0.05, random_state=rng, format="lil")
    for j in range(0, n, 7):
        S[:, j] = 0.0
    S = sp.csr_matrix(S)
    S.data = -S.data * 0.4
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = S + sp.diags(np.abs(S).sum(axis=1).A1 * 1.5 + 1.0)
    return _problem(S, rng.random(n) - 0.3)


@pytest.mark.parametrize("form", range(len(FORMS)))
def test_absorbing_states_paths_and_tally(form):
    prob = _absorbing()
    A = po.Csr(prob.rp, prob.col, prob.values())
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    lens = np.diff(dom.rp)
    assert np.all(lens[::7] == 0) and lens.max() > 0
    f = FORMS[form]
    L, nh = 60, 6000
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(f, **{"Number of Histories": nh, "History Length": L, "Weight Cutoff": 1e-6,
                                  "Random Seed": 5}))
        rhs = prob.b * dinv
        gp, gw = ctx.record_paths(rhs, 2, 0, 3000, L)
        xg = ctx.mc_solve_raw(rhs, 2)
    mp = po.mc_params(nh, L, 1e-6, 64 if f["Walk Precision"] == "FP64" else 32,
                      source=po.SRC_RANDOM if f["Source Sampling Type"] == "Random" else po.SRC_STRATIFIED, seed=5,
                      sampler=po.SAMPLER_CDF if f["Transition Sampler"] == "CDF" else po.SAMPLER_ALIAS)
    op, ow = po.record_paths(dom, rhs, mp, 0, 3000, L, solve_index=2)
    assert np.array_equal(gp, op) and np.array_equal(gw, ow)
    # some walks end in an absorbing state before the length or weight limits
    last = np.array([p[p >= 0][-1] for p in gp])
    steps = (gp >= 0).sum(axis=1) - 1
    assert np.any((last % 7 == 0) & (steps < L))
    xo, _ = po.mc_solve(dom, rhs, mp, solve_index=2)
    tol = 1e-10 if f["Walk Precision"] == "FP64" else 2e-5
    assert np.linalg.norm(xg - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("form", range(len(FORMS)))
def test_single_state_source(form):
    """Every history is born in the one state with b != 0."""
    prob = _absorbing(60, 9)
    b = np.zeros(prob.n)
    b[17] = -2.5
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(FORMS[form], **{"Number of Histories": 2000, "History Length": 20, "Random Seed": 8}))
        gp, _ = ctx.record_paths(b, 0, 0, 2000, 20)
    assert np.all(gp[:, 0] == 17)
