"""The oracle's composite operator with general preconditioners
(src/MCLS_LinearProblem_impl.hpp:139-221): transpose, row-by-row products
with the row filter of MT::multiply, and the preconditioned MCSA built on
them.  Pinned by the reference's own LinearProblem tests, dense products,
a summation-order case, the threshold rule recomputed from dense rows, the
diagonal path already pinned elsewhere, and direct solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import pyoracle as po


def _bidiag(n=10):
    # tstTpetraLinearProblem.cpp:354-408: rows 0..n-2 hold (i: 1, i+1: 2)
    rp, col, val = [0], [], []
    for i in range(n):
        if i < n - 1:
            col += [i, i + 1]
            val += [1.0, 2.0]
        rp.append(len(col))
    return po.Csr(rp, col, val, n)


def _rand(n, density, seed, diag=0.0, scale=1.0):
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=density, random_state=rng, format="csr") * scale
    S.data -= 0.5 * scale
    if diag:
        S = S + sp.eye(n) * diag
    S = S.tocsr()
    S.sort_indices()
    return po.Csr.from_scipy(S), S


def test_reference_composite_operator():
    """LinearProblem.CompositeOperator: PL = A^T on the bidiagonal A gives
    interior rows (i-1, i, i+1) = (2, 5, 2)."""
    A = _bidiag()
    D = po.composite(A, PL=po.transpose(A), forward=True)
    for i in range(1, 9):
        cs = D.col[D.rp[i]:D.rp[i + 1]]
        vs = D.val[D.rp[i]:D.rp[i + 1]]
        assert list(cs) == [i - 1, i, i + 1]
        assert list(vs) == [2.0, 5.0, 2.0]


def test_reference_apply_identity_preconditioners():
    """LinearProblem.Apply: with A = PL = PR = I the composite is I."""
    n = 10
    I = po.diag_csr(np.ones(n))
    for fwd in (True, False):
        C = po.composite(I, PL=I, PR=I, forward=fwd)
        assert np.array_equal(C.rp, np.arange(n + 1)) and np.array_equal(C.col, np.arange(n))
        assert np.array_equal(C.val, np.ones(n))


def test_transpose_matches_scipy():
    A, S = _rand(60, 0.1, 1)
    T = po.transpose(A)
    St = S.T.tocsr()
    St.sort_indices()
    assert np.array_equal(T.rp, St.indptr) and np.array_equal(T.col, St.indices)
    assert np.array_equal(T.val, St.data)


@pytest.mark.parametrize("seed", [2, 3])
def test_product_matches_dense(seed):
    X, SX = _rand(50, 0.12, seed)
    Y, SY = _rand(50, 0.12, seed + 10)
    C = po.spgemm(X, Y)
    dense = SX.toarray() @ SY.toarray()
    got = C.to_dense()
    assert np.allclose(got, dense, rtol=1e-13, atol=1e-15)
    # the pattern is the product's structural pattern less exact zeros
    struct = (np.abs(SX.toarray()) > 0).astype(int) @ (np.abs(SY.toarray()) > 0).astype(int)
    assert np.array_equal(got != 0, (struct > 0) & (got != 0))
    assert np.all(got[struct > 0] != 0)  # random values: no cancellation to zero
    for i in range(C.n):  # ascending columns
        assert np.all(np.diff(C.col[C.rp[i]:C.rp[i + 1]]) > 0)


def test_product_sums_in_row_order_and_drops_zero():
    """C_00 = 1e16*1 + 1*1 + (-1e16)*1 summed in X's row order is
    (1e16 + 1) - 1e16 = 0, which the filter removes; any order that adds the
    two large terms first would give 1.  C_01 = 1 + 2 survives."""
    X = po.Csr([0, 3, 3, 3], [0, 1, 2], [1e16, 1.0, -1e16], 3)
    Y = po.Csr([0, 2, 4, 5], [0, 1, 0, 1, 0], [1.0, 1.0, 1.0, 2.0, 1.0], 3)
    C = po.spgemm(X, Y)
    assert list(C.rp) == [0, 1, 1, 1]
    assert list(C.col) == [1] and list(C.val) == [1e16 + 2.0]
    # sanity: the three products of C_00 are what the test claims
    assert (1e16 + 1.0) - 1e16 == 0.0 and (1e16 - 1e16) + 1.0 == 1.0


@pytest.mark.parametrize("thr", [0.05, 0.3])
def test_filter_threshold_rule(thr):
    """Per row: drop |c| <= (signed row max) * threshold, recomputed from the
    unfiltered dense product; a row whose maximum is negative keeps all."""
    X, SX = _rand(40, 0.15, 7)
    Y, SY = _rand(40, 0.15, 8)
    # rows 0..9 of X made all-negative and Y all-positive: those product rows
    # have a negative maximum
    Y = po.Csr(Y.rp, Y.col, np.abs(Y.val), Y.n)
    xv = X.val.copy()
    xv[:X.rp[10]] = -np.abs(xv[:X.rp[10]])
    X = po.Csr(X.rp, X.col, xv, X.n)
    full = po.spgemm(X, Y, 0.0)
    C = po.spgemm(X, Y, thr)
    negmax_rows = 0
    for i in range(full.n):
        vals = full.val[full.rp[i]:full.rp[i + 1]]
        cols = full.col[full.rp[i]:full.rp[i + 1]]
        if len(vals) == 0:
            assert C.rp[i + 1] == C.rp[i]
            continue
        keep = np.abs(vals) > vals.max() * thr
        negmax_rows += vals.max() < 0
        assert np.array_equal(C.col[C.rp[i]:C.rp[i + 1]], cols[keep])
        assert np.array_equal(C.val[C.rp[i]:C.rp[i + 1]], vals[keep])
    assert negmax_rows > 0  # the signed-maximum rule is exercised


@pytest.mark.parametrize("forward", [False, True])
@pytest.mark.parametrize("thr", [0.0, 0.05])
def test_diagonal_left_prec_matches_diagonal_path(forward, thr):
    """A sparse PL that is diagonal builds the same domain, bit for bit, as
    the diagonal path (whose filter and products are pinned elsewhere)."""
    A, _ = _rand(80, 0.08, 11, diag=3.0)
    dinv = po.jacobi_dinv(A)
    d1 = po.build_domain(A, dinv, 0.9, thr, forward)
    d2 = po.build_domain_general(A, PL=po.diag_csr(dinv), omega=0.9, threshold=thr, forward=forward)
    for a in ("rp", "col", "h", "cdf", "W"):
        assert np.array_equal(getattr(d1, a), getattr(d2, a)), a


def test_adjoint_composite_is_transpose_of_forward_without_filter():
    A, SA = _rand(50, 0.1, 21, diag=2.0)
    PL, SL = _rand(50, 0.06, 22, diag=1.0)
    PR, SR = _rand(50, 0.06, 23, diag=1.0)
    fwd = po.composite(A, PL, PR, 0.0, forward=True).to_dense()
    adj = po.composite(A, PL, PR, 0.0, forward=False).to_dense()
    ref = SL.toarray() @ SA.toarray() @ SR.toarray()
    assert np.allclose(fwd, ref, rtol=1e-12, atol=1e-14)
    assert np.allclose(adj, ref.T, rtol=1e-12, atol=1e-14)


def _system(n=120, seed=5):
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=0.04, random_state=rng, format="csr")
    S.data = -S.data * 0.5
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.abs(S).sum(axis=1).A1 * 1.6 + 0.5)).tocsr()
    S.sort_indices()
    return po.Csr.from_scipy(S), S, rng.random(n)


def test_mcsa_general_with_diagonal_left_prec_equals_c_mcsa():
    """The preconditioned MCSA loop with PL = diag(dinv) as a sparse matrix
    reproduces the C oracle's MCSA with the diagonal PL exactly."""
    A, S, b = _system()
    dinv = po.jacobi_dinv(A)
    mp = po.mc_params(4000, 60, 1e-4, 64, seed=17)
    sp_ = po.solver_params(tol=1e-8, max_iters=30, mc=mp)
    x1, r1, h1 = po.mcsa_solve(A, b, dinv=dinv, params=sp_, solve_index0=3)
    x2, r2, h2 = po.mcsa_solve_general(po.PrecSystem(A, PL=po.diag_csr(dinv)), b, params=sp_, solve_index0=3)
    assert r1.iters == r2["iters"]
    assert np.array_equal(x1, x2)
    assert np.array_equal(h1, np.array(h2))


def test_mcsa_general_right_and_left_prec_solves():
    """Sparse PL (block-diagonal inverse) and PR (a diagonal scaling): the
    recovered x = PR y solves A x = b."""
    A, S, b = _system()
    n = A.n
    # PL: inverse of the 2x2 diagonal blocks of A
    Ad = S.toarray()
    PLd = np.zeros((n, n))
    for k in range(0, n, 2):
        PLd[k:k + 2, k:k + 2] = np.linalg.inv(Ad[k:k + 2, k:k + 2])
    PL = po.Csr.from_dense(PLd)
    PR = po.diag_csr(np.linspace(0.8, 1.2, n))
    sys_ = po.PrecSystem(A, PL=PL, PR=PR)
    mp = po.mc_params(8000, 80, 1e-4, 64, seed=5)
    x, res, hist = po.mcsa_solve_general(sys_, b, params=po.solver_params(tol=1e-8, max_iters=40, mc=mp))
    assert res["converged"]
    assert np.allclose(x, sla.spsolve(S.tocsc(), b), rtol=1e-6, atol=1e-9)
    # the residual is measured in y: PL (b - A PR y)
    y = x / np.linspace(0.8, 1.2, n)
    r = sys_.prec_residual(b, y)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(sys_.left(b)) * 1.0000001


def test_mc_solve_general_identity_preconditioners_match_plain():
    """MonteCarloSolverManager with PL = PR = I gives the unpreconditioned
    answer: the composite is A itself (the filter only drops zeros)."""
    A, S, b = _system(60, 9)
    n = A.n
    I = po.diag_csr(np.ones(n))
    mp = po.mc_params(2000, 40, 1e-3, 64, seed=2)
    x1, _ = po.mc_solve_general(po.PrecSystem(A, PL=I, PR=I), b, mp)
    x2, _ = po.mc_solve(po.build_domain(A), b, mp)
    assert np.array_equal(x1, x2)
