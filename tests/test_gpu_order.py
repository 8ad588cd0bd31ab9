"""State order (mcls_gpu_set_state_order): a context that numbers the
unknowns its own way solves P A P^T and hands every vector back in the
caller's numbering.  Against a context given P A P^T and P b directly: the
domain and the sampled paths are bit-exact (the paths mapped through the
order), MC tallies agree to the summation-order rounding, residuals are
bit-exact, and the MCSA solution matches a direct solve.  Against the
oracle on P A P^T: domain and paths bit-exact.  Plus the error paths."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

BASE = {"Neumann Relaxation": 1.0, "Composite Operator Threshold": 0.0, "Walk Precision": "FP64",
        "Source Sampling Type": "Random", "Estimator Type": "Collision", "Weight Cutoff": 1e-3,
        "History Length": 200, "Number of Histories": 8000, "Random Seed": 77, "Transition Sampler": "CDF",
        "Monte Carlo Type": "Adjoint"}
FORMS = {
    "plain": {},
    "chosen": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "Alias"},
    "forward": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "Alias",
                "Monte Carlo Type": "Forward"},
}


def permute_csr(rp, col, val, order):
    """P A P^T the way the library builds it: row k is row order[k], its
    columns renumbered by the inverse order, entries in their order."""
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    lens = np.diff(rp)[order]
    rpp = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=rpp[1:])
    idx = np.repeat(rp[:-1][order] - rpp[:-1], lens) + np.arange(rpp[-1])
    return rpp, inv[col[idx]].astype(np.int32), val[idx]


def _system(nx=12):
    prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, nx, seed=6, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)
    order = capi.brick_order(nx, nx, nx, 4, 4, 4)
    rpp, colp, valp = permute_csr(prob.rp, prob.col, prob.val32, order)
    pprob = capi.Problem(None, rpp, colp, None, valp, prob.b[order])
    return prob, pprob, order


def _ctx(prob, form, order=None, params=None):
    c = capi.Context(0)
    if order is not None:
        c.set_state_order(order)
    c.set_problem(prob, "f32")
    c.set_point_jacobi()
    c.set_params(dict(BASE, **FORMS[form], **(params or {})))
    return c


@pytest.mark.parametrize("form", sorted(FORMS))
def test_domain_and_paths_match_permuted_problem(form):
    prob, pprob, order = _system()
    rhs = np.random.default_rng(2).random(prob.n) - 0.3
    with _ctx(prob, form, order) as co, _ctx(pprob, form) as cp:
        co.build_domain()
        cp.build_domain()
        for a, b in zip(co.domain(), cp.domain()):
            assert np.array_equal(a, b)
        po_paths, pw = co.record_paths(rhs, 3, 0, 2000, 200)
        pp_paths, ppw = cp.record_paths(rhs[order], 3, 0, 2000, 200)
    mapped = np.where(pp_paths >= 0, order[np.maximum(pp_paths, 0)], -1)
    assert np.array_equal(po_paths, mapped) and np.array_equal(pw, ppw)
    assert (po_paths[:, 1] >= 0).mean() > 0.9


def test_domain_and_paths_match_oracle():
    prob, pprob, order = _system(10)
    A = po.Csr(pprob.rp, pprob.col, pprob.val32.astype(np.float64))
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    rhs = pprob.b * dinv
    L, nh = 150, 6000
    with _ctx(prob, "chosen", order, {"History Length": L, "Number of Histories": nh, "Random Seed": 9}) as co:
        co.build_domain()
        rp, col, h, cdf, W = co.domain()
        assert np.array_equal(rp, dom.rp) and np.array_equal(col, dom.col)
        assert np.array_equal(h, dom.h) and np.array_equal(cdf, dom.cdf) and np.array_equal(W, dom.W)
        user_rhs = np.empty_like(rhs)
        user_rhs[order] = rhs
        gp, gw = co.record_paths(user_rhs, 4, 100, 2500, L)
    mp = po.mc_params(nh, L, BASE["Weight Cutoff"], 32, source=po.SRC_STRATIFIED, seed=9, sampler=po.SAMPLER_ALIAS)
    op, ow = po.record_paths(dom, rhs, mp, 100, 2500, L, solve_index=4)
    assert np.array_equal(gp, np.where(op >= 0, order[np.maximum(op, 0)], -1))
    assert np.array_equal(gw, ow)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_mc_tally_and_residual_match_permuted_problem(form):
    prob, pprob, order = _system()
    rhs = np.random.default_rng(3).random(prob.n)
    x0 = np.random.default_rng(4).random(prob.n)
    with _ctx(prob, form, order) as co, _ctx(pprob, form) as cp:
        xo = co.mc_solve_raw(rhs, 8)
        xp = cp.mc_solve_raw(rhs[order], 8)
        assert co.stats().steps == cp.stats().steps
        tol = 1e-10 if form == "plain" else 2e-5
        assert np.linalg.norm(xo[order] - xp) <= tol * np.linalg.norm(xp)
        ro, no = co.prec_residual(prob.b, x0)
        rq, nq = cp.prec_residual(pprob.b, x0[order])
        assert np.array_equal(ro[order], rq)  # same rows, same entry order
        assert abs(no - nq) <= 1e-13 * nq
        assert np.array_equal(co.get_left_prec_diag()[order], cp.get_left_prec_diag())


@pytest.mark.parametrize("tile_prec", ["Walk", "Compensated", "FP64"])
@pytest.mark.parametrize("walks", [1, 2])
@pytest.mark.parametrize("tile", [256, 1024])
def test_ordered_tile_tally_matches_oracle(tile, walks, tile_prec):
    """The strided FP32 walk with a state order, whose LDS tile maps a state
    to set (state mod sets) instead of a hash, against the oracle on
    P A P^T: same step count, tallies to the FP32 partial sums' rounding
    ("Walk") or to the FP64 summation order ("Compensated", and "FP64", the
    256^3 workload's chosen form)."""
    prob, pprob, order = _system(10)
    A = po.Csr(pprob.rp, pprob.col, pprob.val32.astype(np.float64))
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    rhs = pprob.b * dinv
    nh, L = 40000, 100
    params = {"Number of Histories": nh, "History Length": L, "Weight Cutoff": 1e-2, "Random Seed": 5,
              "Tally Tile": tile, "Tally Tile Precision": tile_prec, "Walks Per Lane": walks,
              "Walk Block Size": 64, "Walk Blocks Per CU": 0}
    user_rhs = np.empty_like(rhs)
    user_rhs[order] = rhs
    with _ctx(prob, "chosen", order, params) as co:
        co.set_params({"Transition Sampler": "Alias Strided"})
        xg = co.mc_solve_raw(user_rhs, 3)
        steps = co.stats().steps
    mp = po.mc_params(nh, L, 1e-2, 32, po.EST_COLLISION, po.SRC_STRATIFIED, seed=5,
                      sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps_o = po.mc_solve(dom, rhs, mp, solve_index=3)
    assert steps == steps_o
    tol = 2e-5 if tile_prec == "Walk" else 1e-12
    assert np.linalg.norm(xg[order] - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("form", ["chosen", "plain"])
def test_mcsa_ordered_converges(form):
    prob, pprob, order = _system()
    params = {"Convergence Tolerance": 1e-8, "Maximum Iterations": 100, "Number of Histories": 40000,
              "Weight Cutoff": 1e-2}
    with _ctx(prob, form, order, params) as co, _ctx(pprob, form, None, params) as cp:
        x = co.mcsa_solve(prob.b)
        so = co.stats()
        xp = cp.mcsa_solve(pprob.b)
        sq = cp.stats()
    assert so.converged == 1 and so.achieved_tol <= 1e-8
    assert abs(so.iterations - sq.iterations) <= 1
    S = sp.csr_matrix((prob.val32.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    xd = sla.spsolve(S.tocsc(), prob.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    assert np.linalg.norm(x[order] - xp) <= 1e-6 * np.linalg.norm(xp)


def test_device_pointers_ordered():
    prob, _, order = _system()
    params = {"Convergence Tolerance": 1e-8, "Maximum Iterations": 100, "Number of Histories": 40000}
    with _ctx(prob, "chosen", order, params) as c:
        xh = c.mcsa_solve(prob.b)
        ih = c.stats().iterations
    with _ctx(prob, "chosen", order, params) as c:
        b, x = capi.DeviceArray(prob.n), capi.DeviceArray(prob.n)
        b.upload(prob.b)
        x.zero()
        c.mcsa_solve_device(b.ptr.value, x.ptr.value)
        xd = x.download()
        idv = c.stats().iterations
        b.close()
        x.close()
    assert ih == idv
    assert np.linalg.norm(xh - xd) <= 1e-6 * np.linalg.norm(xh)


def _block_inverse(S, bs=3):
    A = S.toarray()
    P = np.zeros_like(A)
    for k in range(0, A.shape[0], bs):
        e = min(A.shape[0], k + bs)
        P[k:e, k:e] = np.linalg.inv(A[k:e, k:e])
    M = sp.csr_matrix(P)
    M.sort_indices()
    return M


def test_sparse_preconditioners_ordered():
    """General PL and PR under an order: the context's composite operator is
    the one of P PL P^T, P A P^T, P PR P^T (rows re-sorted)."""
    rng = np.random.default_rng(5)
    n = 160
    S = sp.random(n, n, density=0.04, random_state=rng, format="csr")
    S.data = -S.data * 0.5
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.abs(S).sum(axis=1).A1 * 1.6 + 0.5)).tocsr()
    S.sort_indices()
    b = rng.random(n)
    PL = _block_inverse(S)
    PR = (sp.diags(rng.uniform(0.7, 1.3, n)) + sp.diags(rng.uniform(-0.05, 0.05, n - 1), 1)).tocsr()
    PR.sort_indices()
    order = rng.permutation(n).astype(np.int32)

    def pmat(M):
        M = M.tocsr()[order][:, order].tocsr()
        M.sort_indices()
        return M

    def make(perm):
        c = capi.Context(0)
        A, L, R, bb = (pmat(S), pmat(PL), pmat(PR), b[order]) if perm else (S, PL, PR, b)
        if not perm:
            c.set_state_order(order)
        c.set_problem(capi.Problem(None, A.indptr.astype(np.int64), A.indices.astype(np.int32),
                                   A.data.astype(np.float64), None, bb))
        c.set_left_prec_matrix(L.indptr, L.indices, L.data)
        c.set_right_prec_matrix(R.indptr, R.indices, R.data)
        c.set_params(dict(BASE, **{"Number of Histories": 20000, "History Length": 100, "Convergence Tolerance": 1e-8,
                                   "Maximum Iterations": 100}))
        return c, bb

    (co, bo), (cp, bp) = make(False), make(True)
    with co, cp:
        co.build_domain()
        cp.build_domain()
        for a, q in zip(co.domain(), cp.domain()):
            assert np.array_equal(a, q)
        xo = co.mc_solve(bo)
        xp = cp.mc_solve(bp)
        assert np.linalg.norm(xo[order] - xp) <= 1e-10 * np.linalg.norm(xp)
        x = co.mcsa_solve(bo)
        assert co.stats().converged == 1
    xd = sla.spsolve(S.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def test_importance_pieces_ordered():
    prob, pprob, order = _system(10)
    g = np.random.default_rng(8).uniform(0.5, 2.0, prob.n).astype(np.float32)
    v = np.random.default_rng(9).random(prob.n)
    with _ctx(prob, "chosen", order) as co, _ctx(pprob, "chosen") as cp:
        co.set_importance(g)
        assert np.array_equal(co.importance(), g)
        cp.set_importance(g[order])
        co.set_params({"Transition Sampler": "Learned Alias"})
        cp.set_params({"Transition Sampler": "Learned Alias"})
        xo = co.mc_solve_raw(v, 2)
        xp = cp.mc_solve_raw(v[order], 2)
        assert np.linalg.norm(xo[order] - xp) <= 2e-5 * np.linalg.norm(xp)
        Xo = co.importance_features(v)
        Xp = cp.importance_features(v[order])
        assert np.allclose(Xo[order], Xp, rtol=1e-6, atol=1e-7)
        mask = (np.arange(prob.n) % 3 != 0).astype(np.int8)
        co.set_importance_mask(mask)
        cp.set_importance_mask(mask[order])
        xo = co.mc_solve_raw(v, 3)
        xp = cp.mc_solve_raw(v[order], 3)
        assert np.linalg.norm(xo[order] - xp) <= 2e-5 * np.linalg.norm(xp)


def test_order_errors_and_reset():
    prob, _, order = _system(6)
    with capi.Context(0) as c:
        bad = order.copy()
        bad[3] = bad[4]
        with pytest.raises(capi.GpuError) as e:
            c.set_state_order(bad)
        assert e.value.code == capi.MCLS_ERR_INVALID_ARG
        bad = order.copy()
        bad[0] = prob.n
        with pytest.raises(capi.GpuError):
            c.set_state_order(bad)
        c.set_state_order(np.arange(prob.n - 1, dtype=np.int32))
        with pytest.raises(capi.GpuError) as e:
            c.set_problem(prob, "f32")  # n differs from the order's length
        assert e.value.code == capi.MCLS_ERR_INVALID_ARG
        c.set_state_order(order)
        c.set_problem(prob, "f32")
        c.set_point_jacobi()
        c.set_params(dict(BASE, **{"Number of Histories": 2000}))
        with pytest.raises(capi.GpuError) as e:
            c.set_systems(np.array([0, prob.n // 2, prob.n]))
        assert e.value.code == capi.MCLS_ERR_UNSUPPORTED
        xo = c.mc_solve_raw(prob.b, 1)
        c.set_state_order(order)  # drops the operator
        with pytest.raises(capi.GpuError) as e:
            c.mc_solve(prob.b)
        assert e.value.code == capi.MCLS_ERR_STATE
        c.set_state_order(None)
        c.set_problem(prob, "f32")
        c.set_point_jacobi()
        xn = c.mc_solve_raw(prob.b, 1)  # the caller's numbering again
    with _ctx(prob, "plain", None, {"Number of Histories": 2000}) as c:
        xr = c.mc_solve_raw(prob.b, 1)
    assert np.all(np.isfinite(xo))
    assert np.linalg.norm(xn - xr) <= 1e-10 * np.linalg.norm(xr)
