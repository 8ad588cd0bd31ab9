"""The benchmark's own operator at full size (3-D 27-point stencil, 256^3 =
16.8M unknowns, 449M nonzeros, FP32 CSR): properties that do not depend on
the size.  Adjoint domain rows recomputed from A for a random sample of rows
(bit-exact), row weights W = sum |h| and CDFs ending at exactly 1 everywhere,
the forward domain's row sums inside the generator's [0.6, 0.9], and MCSA to
1e-8 for the chosen adjoint form and the forward method, checked with an
independent host SpMV of the true residual."""
import numpy as np
import pytest
import scipy.sparse as sp

from mcls_amd import capi

pytestmark = pytest.mark.gpu

N1D = 256


@pytest.fixture(scope="module")
def problem():
    spec = capi.make_spec(capi.STENCIL3D_27PT, N1D, seed=4, p0=0.6, p1=0.9, p2=0.8, p3=1.2)
    prob = capi.build_problem(spec, want64=False, want32=True, nthreads=16)
    A = sp.csr_matrix((prob.val32.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    return prob, A


@pytest.fixture(scope="module")
def ctx(problem):
    prob, _ = problem
    c = capi.Context(0)
    c.set_problem(prob, "f32")
    c.set_point_jacobi()
    yield c
    c.close()


def test_adjoint_domain_full_size(problem, ctx):
    prob, A = problem
    ctx.set_params({"Monte Carlo Type": "Adjoint"})
    ctx.build_domain()
    rp, col, h, cdf, W = ctx.domain()
    n = prob.n
    lens = np.diff(rp)
    assert lens.min() >= 1 and rp[-1] == len(h)
    last = rp[1:] - 1
    assert np.all(cdf[last] == 1.0)
    rows_sum = np.add.reduceat(np.abs(h), rp[:-1])
    assert np.allclose(rows_sum, W, rtol=4e-16 * 27, atol=0.0)
    # domain row i = column i of PL A: h_ij = delta_ij - a_ji / a_jj
    d = A.diagonal()
    dinv = ctx.get_left_prec_diag()
    assert np.array_equal(dinv, 1.0 / d)
    rng = np.random.default_rng(7)
    for i in rng.choice(n, 2000, replace=False):
        cols = col[rp[i]:rp[i + 1]]
        got = h[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0)
        want = []
        for j in cols:
            row = A.indices[A.indptr[j]:A.indptr[j + 1]]
            k = A.indptr[j] + np.searchsorted(row, i)
            m = A.data[k] * dinv[j]
            want.append(1.0 - m if j == i else -m)
        assert np.array_equal(got, np.array(want)), i


def test_forward_domain_row_sums(problem, ctx):
    prob, _ = problem
    ctx.set_params({"Monte Carlo Type": "Forward"})
    ctx.build_domain()
    rp, col, h, cdf, W = ctx.domain()
    ctx.set_params({"Monte Carlo Type": "Adjoint"})
    # rows of H: sum_j |h_ij| = sum_{j != i} |a_ij| / a_ii = s_i in [0.6, 0.9]
    # up to the FP32 rounding of the diagonal
    assert W.min() >= 0.6 * (1 - 1e-6) and W.max() <= 0.9 * (1 + 1e-6)
    assert np.all(cdf[rp[1:] - 1] == 1.0)


def _true_residual(A, b, x, dinv):
    r = b - A @ x
    return np.linalg.norm(dinv * r) / np.linalg.norm(dinv * b)


@pytest.mark.parametrize("method", ["Adjoint", "Forward"])
def test_mcsa_full_size(problem, ctx, method):
    prob, A = problem
    ctx.set_params({"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Estimator Type": "Collision",
                    "Weight Cutoff": 1e-2, "History Length": 1000, "Transition Sampler": "Alias",
                    "Number of Histories": 1 << 28, "Random Seed": 2024, "Convergence Tolerance": 1e-8,
                    "Maximum Iterations": 20, "Monte Carlo Type": method})
    x = ctx.mcsa_solve(prob.b)
    st = ctx.stats()
    ctx.set_params({"Monte Carlo Type": "Adjoint"})
    print(method, st.iterations, st.achieved_tol, st.ms_total)
    assert st.converged == 1 and st.achieved_tol <= 1e-8 and st.iterations <= 8
    # an independent host SpMV of the preconditioned residual
    assert _true_residual(A, prob.b, x, ctx.get_left_prec_diag()) <= 1.5e-8
