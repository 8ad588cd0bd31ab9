"""Strided alias tables ("Transition Sampler" = "Alias Strided", FP32
walks) on the GPU against the oracle: sampled paths and final weights
bit-exact (adjoint and forward, random and stratified births, operators with
absorbing rows and rows shorter than the stride), MC tallies to the FP32
summation order, MCSA to a direct solve, the LDS tile and the state order on
top, and the error paths - for the 16-byte records and their 8-byte form
("Strided Record Bytes" = 8: the same draws, the alias column read from the
alias slot's own record and W from the row array)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

BASE = {"Neumann Relaxation": 1.0, "Composite Operator Threshold": 0.0, "Walk Precision": "FP32",
        "Estimator Type": "Collision", "Weight Cutoff": 1e-3, "History Length": 300,
        "Transition Sampler": "Alias Strided", "Monte Carlo Type": "Adjoint"}


def _problem(kind):
    if kind == "s27":
        p = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 10, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                               want64=False, want32=True)
        return p, po.Csr(p.rp, p.col, p.val32.astype(np.float64)), "f32"
    if kind == "poisson":
        p = capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, 30, 30, seed=0))
        return p, po.Csr(p.rp, p.col, p.values()), "f64"
    # random nonsymmetric operator with absorbing states (columns of A with
    # only their diagonal: empty rows of H^T) and rows of 0..~12 entries
    rng = np.random.default_rng(4)
    n = 150
    S = sp.random(n, n, density=0.05, random_state=rng, format="lil")
    for j in range(0, n, 7):
        S[:, j] = 0.0
    S = sp.csr_matrix(S)
    S.data = -S.data * 0.4
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.abs(S).sum(axis=1).A1 * 1.5 + 1.0)).tocsr()
    S.sort_indices()
    p = capi.Problem(None, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(np.float64), None,
                     rng.random(n) - 0.3)
    return p, po.Csr(p.rp, p.col, p.values()), "f64"


@pytest.mark.parametrize("rec", [16, 8])
@pytest.mark.parametrize("forward", [False, True])
@pytest.mark.parametrize("source", ["Random", "Stratified"])
@pytest.mark.parametrize("kind", ["s27", "poisson", "absorbing"])
def test_paths_bitexact(kind, source, forward, rec):
    prob, A, dt = _problem(kind)
    L, nh = 300, 6000
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv, forward=forward)
    rhs = prob.b * dinv
    rhs[::5] *= -1.0
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Source Sampling Type": source, "Number of Histories": nh, "Random Seed": 41,
                                     "Monte Carlo Type": "Forward" if forward else "Adjoint",
                                     "Strided Record Bytes": rec}))
        count = 2000
        gp, gw = ctx.record_paths(rhs, 7, 50, count, L)
    est = po.EST_FORWARD if forward else po.EST_COLLISION
    mp = po.mc_params(nh, L, 1e-3, 32, est, po.SRC_STRATIFIED if source == "Stratified" else po.SRC_RANDOM,
                      seed=41, sampler=po.SAMPLER_ALIAS_STRIDED)
    op, ow = po.record_paths(dom, rhs, mp, 50, count, L, solve_index=7)
    assert np.array_equal(gp, op)
    assert np.array_equal(gw, ow)
    assert (gp[:, 1] >= 0).mean() > 0.5


@pytest.mark.parametrize("walks,rec", [(1, 16), (2, 16), (1, 8)])
@pytest.mark.parametrize("source", ["Random", "Stratified"])
@pytest.mark.parametrize("tile", [0, 256])
def test_tally_matches_oracle(source, tile, walks, rec):
    prob, A, dt = _problem("s27")
    nh, L = 40000, 100
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    rhs = prob.b * dinv
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Source Sampling Type": source, "Number of Histories": nh, "History Length": L,
                                     "Weight Cutoff": 1e-2, "Random Seed": 5, "Tally Tile": tile,
                                     "Tally Tile Precision": "Walk", "Walks Per Lane": walks,
                                     "Strided Record Bytes": rec}))
        xg = ctx.mc_solve_raw(rhs, 3)
        st = ctx.stats()
    mp = po.mc_params(nh, L, 1e-2, 32, po.EST_COLLISION,
                      po.SRC_STRATIFIED if source == "Stratified" else po.SRC_RANDOM, seed=5,
                      sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps = po.mc_solve(dom, rhs, mp, solve_index=3)
    assert st.steps == steps
    assert np.linalg.norm(xg - xo) <= 2e-5 * np.linalg.norm(xo)


@pytest.mark.parametrize("walks", [1, 2])
@pytest.mark.parametrize("batch", [64, 192, 4096])
def test_tally_over_grab_sizes_matches_oracle(batch, walks):
    """The births of a grab are computed 64 positions at a time and kept for
    the current and the next block: grabs of one block, of three and of 64,
    with one and two histories per lane, all tally what the oracle does."""
    prob, A, dt = _problem("s27")
    nh, L = 30000, 100
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    rhs = prob.b * dinv
    rhs[::7] *= -3.0
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Source Sampling Type": "Stratified", "Number of Histories": nh,
                                     "History Length": L, "Weight Cutoff": 1e-2, "Random Seed": 12,
                                     "Tally Tile": 512, "Tally Tile Precision": "FP64", "Walk Batch": batch,
                                     "Walks Per Lane": walks}))
        xg = ctx.mc_solve_raw(rhs, 4)
        st = ctx.stats()
    mp = po.mc_params(nh, L, 1e-2, 32, po.EST_COLLISION, po.SRC_STRATIFIED, seed=12,
                      sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps = po.mc_solve(dom, rhs, mp, solve_index=4)
    assert st.steps == steps
    assert np.linalg.norm(xg - xo) <= 1e-12 * np.linalg.norm(xo)


def test_record_forms_switch_and_rebuild():
    """8-byte records alone (no 16-byte tables), then two histories per lane
    (which builds the 16-byte ones), then back to 16-byte records with one
    history per lane: every switch walks the same histories, whose tallies
    match the oracle's."""
    prob, A, dt = _problem("s27")
    nh, L = 20000, 100
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv)
    rhs = prob.b * dinv
    mp = po.mc_params(nh, L, 1e-2, 32, po.EST_COLLISION, po.SRC_STRATIFIED, seed=8,
                      sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps = po.mc_solve(dom, rhs, mp, solve_index=6)
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Source Sampling Type": "Stratified", "Number of Histories": nh,
                                     "History Length": L, "Weight Cutoff": 1e-2, "Random Seed": 8,
                                     "Tally Tile": 256, "Tally Tile Precision": "FP64"}))
        for rec, walks in ((8, 1), (8, 2), (16, 1), (8, 1)):
            ctx.set_params({"Strided Record Bytes": rec, "Walks Per Lane": walks})
            xg = ctx.mc_solve_raw(rhs, 6)
            assert ctx.stats().steps == steps, (rec, walks)
            assert np.linalg.norm(xg - xo) <= 1e-12 * np.linalg.norm(xo), (rec, walks)


def test_forward_two_walks_per_lane_matches_oracle():
    """Two histories per lane on the forward method: the same per-state sums
    as the oracle (and as one history per lane) to the FP32 summation order."""
    prob, A, dt = _problem("s27")
    nh, L = 30000, 100
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv, forward=True)
    rhs = prob.b * dinv
    xs = []
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        for walks in (1, 2):
            ctx.set_params(dict(BASE, **{"Monte Carlo Type": "Forward", "Number of Histories": nh,
                                         "History Length": L, "Weight Cutoff": 1e-2, "Random Seed": 6,
                                         "Walks Per Lane": walks}))
            xs.append(ctx.mc_solve_raw(rhs, 2))
            st = ctx.stats()
    mp = po.mc_params(nh, L, 1e-2, 32, po.EST_FORWARD, po.SRC_RANDOM, seed=6, sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps = po.mc_solve(dom, rhs, mp, solve_index=2)
    assert st.steps == steps
    for x in xs:
        assert np.linalg.norm(x - xo) <= 2e-5 * np.linalg.norm(xo)


@pytest.mark.parametrize("walks,rec", [(1, 16), (2, 16), (1, 8)])
@pytest.mark.parametrize("ordered", [False, True])
def test_mcsa_converges(ordered, walks, rec):
    prob, A, dt = _problem("s27")
    with capi.Context(0) as ctx:
        if ordered:
            ctx.set_state_order(capi.brick_order(10, 10, 10, 4, 4, 4))
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Source Sampling Type": "Stratified", "Number of Histories": 16 * prob.n,
                                     "Weight Cutoff": 1e-1, "History Length": 1000, "Random Seed": 9,
                                     "Convergence Tolerance": 1e-8, "Maximum Iterations": 60,
                                     "Tally Tile": 512, "Tally Tile Precision": "Walk",
                                     "Walks Per Lane": walks, "Strided Record Bytes": rec}))
        x = ctx.mcsa_solve(prob.b)
        st = ctx.stats()
    assert st.converged == 1 and st.achieved_tol <= 1e-8
    S = sp.csr_matrix((prob.val32.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    xd = sla.spsolve(S.tocsc(), prob.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def test_errors():
    prob, A, dt = _problem("s27")
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, dt)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Number of Histories": 1000, "Walk Precision": "FP64"}))
        with pytest.raises(capi.GpuError) as e:
            ctx.mc_solve(prob.b)
        assert e.value.code == capi.MCLS_ERR_UNSUPPORTED
        ctx.set_params({"Walk Precision": "FP32", "Estimator Type": "Expected Value"})
        with pytest.raises(capi.GpuError) as e:
            ctx.mc_solve(prob.b)
        assert e.value.code == capi.MCLS_ERR_UNSUPPORTED
        ctx.set_params({"Estimator Type": "Collision"})
        assert np.all(np.isfinite(ctx.mc_solve(prob.b)))
        with pytest.raises(capi.GpuError) as e:
            ctx.set_params({"Strided Record Bytes": 12})
        assert e.value.code == capi.MCLS_ERR_INVALID_ARG
    # 8-byte records hold an alias slot of 5 bits: rows of more than 32 entries refuse them
    n = 40
    S = sp.csr_matrix(-0.01 * np.ones((n, n)) + np.diag(np.full(n, 2.0)))
    p = capi.Problem(None, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(np.float64), None,
                     np.ones(n))
    with capi.Context(0) as ctx:
        ctx.set_problem(p, "f64")
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Number of Histories": 1000, "Strided Record Bytes": 8}))
        with pytest.raises(capi.GpuError) as e:
            ctx.mc_solve(p.b)
        assert e.value.code == capi.MCLS_ERR_UNSUPPORTED
        ctx.set_params({"Strided Record Bytes": 16})
        assert np.all(np.isfinite(ctx.mc_solve(p.b)))
        with pytest.raises(capi.GpuError):
            ctx.set_params({"Walks Per Lane": 3})
    # a domain row of more than 64 entries
    n = 80
    S = sp.lil_matrix((n, n))
    S.setdiag(100.0)
    S[0, 1:] = -0.5
    S = S.tocsr()
    prob = capi.Problem(None, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data, None, np.ones(n))
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(BASE, **{"Number of Histories": 1000, "Monte Carlo Type": "Forward"}))
        with pytest.raises(capi.GpuError) as e:
            ctx.mc_solve(prob.b)
        assert e.value.code == capi.MCLS_ERR_UNSUPPORTED
