"""General (sparse) left / right preconditioners on the GPU against the
oracle: the composite operator's domain (src/MCLS_LinearProblem_impl.hpp:
139-221) bit-exactly, the preconditioned residual, the MC solve with its
x = PR x recovery, MCSA against the oracle and a direct solve, the
reference's prec_adjoint / prec_forward manager tests (PL = PR = I), graph
replay and the error paths."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rand_system(n=150, seed=5, dens=0.04):
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=dens, random_state=rng, format="csr")
    S.data = -S.data * 0.5
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.abs(S).sum(axis=1).A1 * 1.6 + 0.5)).tocsr()
    S.sort_indices()
    return S, rng.random(n)


def _block_inverse(S, bs=3):
    """PL: inverse of the diagonal bs x bs blocks of S (sparse, FP64)."""
    A = S.toarray()
    n = A.shape[0]
    P = np.zeros_like(A)
    for k in range(0, n, bs):
        e = min(n, k + bs)
        P[k:e, k:e] = np.linalg.inv(A[k:e, k:e])
    M = sp.csr_matrix(P)
    M.sort_indices()
    return M


def _scaling(n, seed=1):
    rng = np.random.default_rng(seed)
    M = sp.diags(rng.uniform(0.7, 1.3, n)).tocsr()
    # plus a sparse off-diagonal band so PR is not diagonal
    M = (M + sp.diags(rng.uniform(-0.05, 0.05, n - 1), 1)).tocsr()
    M.sort_indices()
    return M


def _problem(S, b):
    return capi.Problem(None, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(np.float64),
                        None, b)


def _set_prec(ctx, PL, PR, diag_left=False):
    if PL is not None:
        if diag_left:
            ctx.set_left_prec_diag(PL.diagonal())
        else:
            ctx.set_left_prec_matrix(PL.indptr, PL.indices, PL.data)
    else:
        ctx.clear_left_prec()
    if PR is not None:
        ctx.set_right_prec_matrix(PR.indptr, PR.indices, PR.data)
    else:
        ctx.clear_right_prec()


BASE = {"Neumann Relaxation": 1.0, "Composite Operator Threshold": 0.0, "Walk Precision": "FP64",
        "Source Sampling Type": "Random", "Estimator Type": "Collision", "Weight Cutoff": 1e-4,
        "History Length": 100, "Number of Histories": 4000, "Random Seed": 7, "Transition Sampler": "CDF",
        "Monte Carlo Type": "Adjoint"}


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("method", ["Adjoint", "Forward"])
@pytest.mark.parametrize("which", ["left", "right", "both", "diag+right"])
@pytest.mark.parametrize("thr", [0.0, 0.05])
def test_composite_domain_bitexact(ctx, method, which, thr):
    S, b = _rand_system()
    PL = _block_inverse(S)
    PR = _scaling(S.shape[0])
    if which == "diag+right":
        PL = sp.diags(1.0 / S.diagonal()).tocsr()
    use_l = which in ("left", "both", "diag+right")
    use_r = which in ("right", "both", "diag+right")
    ctx.set_problem(_problem(S, b), "f64")
    _set_prec(ctx, PL if use_l else None, PR if use_r else None, diag_left=which == "diag+right")
    ctx.set_params(dict(BASE, **{"Composite Operator Threshold": thr, "Neumann Relaxation": 0.9,
                                 "Monte Carlo Type": method}))
    ctx.build_domain()
    rp, col, h, cdf, W = ctx.domain()
    A = po.Csr.from_scipy(S)
    dom = po.build_domain_general(A, po.Csr.from_scipy(PL) if use_l else None, po.Csr.from_scipy(PR) if use_r else None,
                                  0.9, thr, forward=method == "Forward")
    assert np.array_equal(rp, dom.rp) and np.array_equal(col, dom.col)
    assert np.array_equal(h, dom.h) and np.array_equal(cdf, dom.cdf) and np.array_equal(W, dom.W)
    if thr > 0:  # the filter removed entries
        full = po.build_domain_general(A, po.Csr.from_scipy(PL) if use_l else None,
                                       po.Csr.from_scipy(PR) if use_r else None, 0.9, 0.0,
                                       forward=method == "Forward")
        assert full.nnz > dom.nnz


def test_composite_domain_fp32_operator(ctx):
    """An FP32 operator: its values widen to FP64 before the products."""
    prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 8, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)
    S = sp.csr_matrix((prob.val32.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    PL = _block_inverse(S, 4)
    PR = _scaling(prob.n, 3)
    ctx.set_problem(prob, "f32")
    _set_prec(ctx, PL, PR)
    ctx.set_params(dict(BASE))
    ctx.build_domain()
    rp, col, h, cdf, W = ctx.domain()
    dom = po.build_domain_general(po.Csr.from_scipy(S), po.Csr.from_scipy(PL), po.Csr.from_scipy(PR))
    assert np.array_equal(rp, dom.rp) and np.array_equal(col, dom.col) and np.array_equal(h, dom.h)


def test_prec_residual_matches_oracle(ctx):
    S, b = _rand_system()
    PL, PR = _block_inverse(S), _scaling(S.shape[0])
    ctx.set_problem(_problem(S, b), "f64")
    _set_prec(ctx, PL, PR)
    x = np.random.default_rng(4).random(S.shape[0])
    r, nrm = ctx.prec_residual(b, x)
    sys_ = po.PrecSystem(po.Csr.from_scipy(S), po.Csr.from_scipy(PL), po.Csr.from_scipy(PR))
    ro = sys_.prec_residual(b, x)
    assert np.allclose(r, ro, rtol=1e-12, atol=1e-14 * np.abs(ro).max())
    assert abs(nrm - np.linalg.norm(ro)) <= 1e-12 * np.linalg.norm(ro)


@pytest.mark.parametrize("method", ["Adjoint", "Forward"])
def test_mc_solve_recovers_right_prec(ctx, method):
    """MonteCarloSolverManager: source omega PL b, the composite's domain,
    then x = PR x; same streams as the oracle, so only the tally summation
    order differs."""
    S, b = _rand_system(seed=8)
    PL, PR = _block_inverse(S), _scaling(S.shape[0], 5)
    ctx.set_problem(_problem(S, b), "f64")
    _set_prec(ctx, PL, PR)
    nh = 20000
    ctx.set_params(dict(BASE, **{"Number of Histories": nh, "Monte Carlo Type": method, "Neumann Relaxation": 0.95}))
    xg = ctx.mc_solve(b)
    st = ctx.stats()
    sys_ = po.PrecSystem(po.Csr.from_scipy(S), po.Csr.from_scipy(PL), po.Csr.from_scipy(PR))
    mp = po.mc_params(nh, 100, 1e-4, 64, po.EST_FORWARD if method == "Forward" else po.EST_COLLISION, seed=7)
    xo, steps = po.mc_solve_general(sys_, b, mp, omega=0.95, forward=method == "Forward",
                                    solve_index=st.solve_index - 1)
    assert st.steps == steps
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", ["Adjoint", "Forward"])
def test_mcsa_with_left_and_right_prec(ctx, method):
    S, b = _rand_system(seed=9)
    PL, PR = _block_inverse(S), _scaling(S.shape[0], 6)
    ctx.set_problem(_problem(S, b), "f64")
    _set_prec(ctx, PL, PR)
    ctx.set_params(dict(BASE, **{"Number of Histories": 6000, "Convergence Tolerance": 1e-8,
                                 "Maximum Iterations": 40, "Monte Carlo Type": method}))
    x = ctx.mcsa_solve(b)
    st = ctx.stats()
    assert st.converged == 1 and st.achieved_tol <= 1e-8
    xd = sla.spsolve(S.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    sys_ = po.PrecSystem(po.Csr.from_scipy(S), po.Csr.from_scipy(PL), po.Csr.from_scipy(PR))
    mp = po.mc_params(6000, 100, 1e-4, 64, po.EST_FORWARD if method == "Forward" else po.EST_COLLISION, seed=7)
    xo, res, _ = po.mcsa_solve_general(sys_, b, params=po.solver_params(tol=1e-8, max_iters=40, mc=mp),
                                       solve_index0=st.solve_index - st.iterations, forward=method == "Forward")
    assert res["converged"] and abs(res["iters"] - st.iterations) <= 2
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def _reference_tridiag(n=10):
    # tstTpetraMCSASolverManager.cpp:207-262: 1 on the diagonal, -0.14 off it,
    # explicit zeros at (0, 2) and (n-1, n-3)
    rows, cols, vals = [], [], []
    rows += [0, 0, 0]
    cols += [0, 1, 2]
    vals += [1.0, -0.14, 0.0]
    for i in range(1, n - 1):
        rows += [i, i, i]
        cols += [i - 1, i, i + 1]
        vals += [-0.14, 1.0, -0.14]
    rows += [n - 1, n - 1, n - 1]
    cols += [n - 3, n - 2, n - 1]
    vals += [0.0, -0.14, 1.0]
    rp = np.zeros(n + 1, np.int64)
    for r in rows:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    return capi.Problem(None, rp, np.array(cols, np.int32), np.array(vals), None, np.zeros(n))


@pytest.mark.parametrize("method", ["Adjoint", "Forward"])
def test_reference_prec_managers_identity(ctx, method):
    """prec_adjoint / prec_forward (tstTpetraMCSASolverManager.cpp:207-356,
    :494-...) and adjoint_prec / forward_prec of the MC manager: PL = PR = I,
    b = -1 gives x < 0, b = 2 gives x > 0, in fewer than 20 iterations."""
    prob = _reference_tridiag()
    n = prob.n
    ctx.set_problem(prob, "f64")
    eye_rp, eye_col, eye_val = np.arange(n + 1), np.arange(n), np.ones(n)
    ctx.set_left_prec_matrix(eye_rp, eye_col, eye_val)
    ctx.set_right_prec_matrix(eye_rp, eye_col, eye_val)
    ctx.set_params(dict(BASE, **{"Convergence Tolerance": 1e-8, "Maximum Iterations": 20, "Weight Cutoff": 0.0,
                                 "History Length": 10, "Number of Histories": 10, "Monte Carlo Type": method}))
    for bval, sign in ((-1.0, -1), (2.0, 1), (-2.0, -1)):
        x = ctx.mcsa_solve(np.full(n, bval))
        st = ctx.stats()
        assert st.converged == 1 and st.iterations < 20 and st.achieved_tol > 0.0
        assert np.all(np.sign(x) == sign)
        xm = ctx.mc_solve(np.full(n, bval))
        assert np.all(np.sign(xm) == sign)


def test_graph_replay_matches_eager_with_prec():
    """One MCSA iteration captured in a hipGraph (SpMVs of PL and PR
    included) against eager launches, each in a fresh context so that both
    use the same stream indices: same paths, so only the FP64 tally
    summation order differs."""
    S, b = _rand_system(seed=12)
    PL, PR = _block_inverse(S), _scaling(S.shape[0], 7)
    out = []
    for graph in (0, 1):
        c = capi.Context(0)
        c.set_problem(_problem(S, b), "f64")
        _set_prec(c, PL, PR)
        c.set_params(dict(BASE, **{"Number of Histories": 6000, "Convergence Tolerance": 1e-8,
                                   "Maximum Iterations": 40, "Use Graph": graph, "Random Seed": 21}))
        x = c.mcsa_solve(b)
        st = c.stats()
        out.append((x, st.iterations, st.converged, st.achieved_tol))
        c.close()
    (xe, ie, ce, te), (xg, ig, cg, tg) = out
    assert ce == 1 and cg == 1 and ie == ig
    assert abs(te - tg) <= 1e-6 * te
    assert np.linalg.norm(xe - xg) <= 1e-9 * np.linalg.norm(xe)


def test_prec_errors(ctx):
    S, b = _rand_system(60)
    ctx.set_problem(_problem(S, b), "f64")
    n = S.shape[0]
    # columns not ascending
    rp = np.arange(n + 1) * 2
    col = np.tile([1, 0], n)
    with pytest.raises(capi.GpuError):
        ctx.set_left_prec_matrix(rp, col, np.ones(2 * n))
    # row pointers that do not start at 0
    with pytest.raises(capi.GpuError):
        ctx.set_right_prec_matrix(np.arange(n + 1) + 1, np.arange(n), np.ones(n))
    # column out of range
    with pytest.raises(capi.GpuError):
        ctx.set_right_prec_matrix(np.arange(n + 1), np.full(n, n + 3), np.ones(n))
    # a sparse PL has no diagonal to return
    ctx.set_left_prec_matrix(np.arange(n + 1), np.arange(n), np.ones(n))
    with pytest.raises(capi.GpuError):
        ctx.get_left_prec_diag()
    ctx.clear_left_prec()
    ctx.clear_right_prec()
