"""Random births walked in birth-state order ("Birth Order" = "State"): the
births are drawn ahead of the walk from each history's own stream and
sorted by state, so only the order in which histories are walked changes.
Paths and final weights are bit-exact against the oracle (which walks in
history-id order), tallies match it to the atomic summation order, windows
of history ids keep their meaning, the autotuner's pilots and MCSA run with
it, and a captured MCSA iteration refuses it."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

BASE = {"Neumann Relaxation": 1.0, "Composite Operator Threshold": 0.0, "Walk Precision": "FP64",
        "Source Sampling Type": "Random", "Estimator Type": "Collision", "Weight Cutoff": 0.0,
        "History Length": 10, "Number of Histories": 0, "Random Seed": 0, "Transition Sampler": "CDF",
        "Monte Carlo Type": "Adjoint", "Residual Format": "Sliced ELL", "Tally Tile": 0,
        "Tally Tile Precision": "FP64", "Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64,
        "Use Graph": 0}


def _s27(n=12):
    return capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, n, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)


def _setup(ctx, prob, f32, params):
    ctx.set_problem(prob, "f32" if f32 else "f64")
    ctx.set_point_jacobi()
    ctx.set_params(dict(BASE, **params))


def _oracle(prob, f32):
    val = prob.val32.astype(np.float64) if f32 else prob.values()
    A = po.Csr(prob.rp, prob.col, val)
    dinv = po.jacobi_dinv(A)
    return po.build_domain(A, dinv), dinv


@pytest.fixture
def octx(ctx):
    yield ctx
    ctx.set_params({"Birth Order": "Id", "Use Graph": 0, "Tally Tile": 0, "Walk Block Size": 256,
                    "Walk Blocks Per CU": 0, "Walk Batch": 64})


@pytest.mark.parametrize("sampler,prec", [("CDF", 64), ("Alias Strided", 32), ("Alias", 32)])
def test_ordered_births_paths_bitexact(octx, sampler, prec):
    ctx = octx
    prob = _s27()
    L, nh = 60, 30000
    _setup(ctx, prob, True, {"Walk Precision": f"FP{prec}", "Transition Sampler": sampler, "History Length": L,
                             "Weight Cutoff": 1e-3, "Number of Histories": nh, "Random Seed": 77})
    dom, dinv = _oracle(prob, True)
    rhs = prob.b * dinv
    rhs[::5] *= -1.0
    first, count = 1000, 8000
    ctx.set_params({"Birth Order": "Id"})
    ip, iw = ctx.record_paths(rhs, 9, first, count, L)
    ctx.set_params({"Birth Order": "State"})
    gp, gw = ctx.record_paths(rhs, 9, first, count, L)
    samp = {"CDF": po.SAMPLER_CDF, "Alias": po.SAMPLER_ALIAS, "Alias Strided": po.SAMPLER_ALIAS_STRIDED}[sampler]
    mp = po.mc_params(nh, L, 1e-3, prec, po.EST_COLLISION, po.SRC_RANDOM, seed=77, sampler=samp)
    op, ow = po.record_paths(dom, rhs, mp, first, count, L, solve_index=9)
    assert np.array_equal(gp, op) and np.array_equal(gw, ow)
    assert np.array_equal(gp, ip) and np.array_equal(gw, iw)
    assert (gp[:, 1] >= 0).mean() > 0.9


@pytest.mark.parametrize("n", [7, 40])
def test_ordered_births_skewed_source_bitexact(octx, n):
    """A source whose weight sits on a few states across 16 decades, most
    states unweighted: many guide buckets hold one state or none and one
    holds many, and the births resolved after the bucket sort must still be
    the states the walk's own first draw selects."""
    ctx = octx
    prob = capi.build_problem(capi.make_spec(capi.DIFFUSION3D_7PT, n, seed=5, p0=4))
    L, nh = 40, 20000
    _setup(ctx, prob, False, {"History Length": L, "Weight Cutoff": 1e-3, "Number of Histories": nh,
                              "Random Seed": 5})
    dom, dinv = _oracle(prob, False)
    rng = np.random.default_rng(n)
    rhs = np.zeros(prob.n)
    hot = rng.choice(prob.n, size=min(prob.n, 37), replace=False)
    rhs[hot] = np.logspace(-8, 8, hot.size) * rng.choice([-1.0, 1.0], size=hot.size)
    rhs[hot[:5]] = 1e-300  # weights far below the rest round to no births at all
    ctx.set_params({"Birth Order": "State"})
    gp, gw = ctx.record_paths(rhs, 3, 0, nh, L)
    mp = po.mc_params(nh, L, 1e-3, 64, po.EST_COLLISION, po.SRC_RANDOM, seed=5, sampler=po.SAMPLER_CDF)
    op, ow = po.record_paths(dom, rhs, mp, 0, nh, L, solve_index=3)
    assert np.array_equal(gp, op) and np.array_equal(gw, ow)
    born = np.unique(gp[:, 0])
    assert set(born.tolist()) <= set(hot.tolist())
    xg = ctx.mc_solve_raw(rhs, 3)
    xo, _ = po.mc_solve(dom, rhs, mp, solve_index=3)
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("tile", [0, 1024])
@pytest.mark.parametrize("est", ["Collision", "Expected Value"])
def test_ordered_births_tally_matches_oracle(octx, est, tile):
    ctx = octx
    prob = capi.build_problem(capi.make_spec(capi.DIFFUSION3D_7PT, 12, seed=5, p0=4))
    nh, L, cutoff = 40000, 50, 1e-2
    _setup(ctx, prob, False, {"Estimator Type": est, "History Length": L, "Weight Cutoff": cutoff,
                              "Number of Histories": nh, "Random Seed": 99, "Tally Tile": tile,
                              "Birth Order": "State"})
    dom, dinv = _oracle(prob, False)
    rhs = prob.b * dinv
    xg = ctx.mc_solve_raw(rhs, 5)
    st = ctx.stats()
    mp = po.mc_params(nh, L, cutoff, 64, po.EST_EXPECTED_VALUE if est == "Expected Value" else po.EST_COLLISION,
                      po.SRC_RANDOM, seed=99, sampler=po.SAMPLER_CDF)
    xo, steps = po.mc_solve(dom, rhs, mp, solve_index=5)
    assert st.steps == steps
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
    # a window of history ids keeps its meaning: the same histories, walked in
    # their state order, against the oracle's window
    xw = ctx.mc_solve_window(rhs, 5, 7000, 9000)
    xow, _ = po.mc_solve_window(dom, rhs, mp, 7000, 9000, solve_index=5)
    assert np.linalg.norm(xw - xow) <= 1e-10 * np.linalg.norm(xow)


def test_ordered_births_autotune_and_mcsa(octx):
    ctx = octx
    prob = _s27(16)
    _setup(ctx, prob, True, {"Walk Precision": "FP32", "Transition Sampler": "Alias Strided",
                             "History Length": 1000, "Weight Cutoff": 1e-4, "Number of Histories": 200000,
                             "Random Seed": 3, "Convergence Tolerance": 1e-8, "Maximum Iterations": 60,
                             "Birth Order": "State"})
    dom, dinv = _oracle(prob, True)
    rhs = prob.b * dinv
    t = ctx.autotune(prob.b, pilot_histories=1 << 15)
    assert t["ms_best"] > 0.0 and t["ms_confirmed"] > 0.0
    xs = ctx.mc_solve_raw(rhs, 11)
    ctx.set_params({"Birth Order": "Id"})
    xi = ctx.mc_solve_raw(rhs, 11)
    assert np.linalg.norm(xs - xi) <= 1e-12 * np.linalg.norm(xi)  # FP64 tallies: summation order only
    ctx.set_params({"Birth Order": "State"})
    x = ctx.mcsa_solve(prob.b)
    st = ctx.stats()
    assert st.converged == 1 and st.achieved_tol <= 1e-8
    S = sp.csr_matrix((prob.val32.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    xd = sla.spsolve(S.tocsc(), prob.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def test_ordered_births_refused_in_a_captured_iteration(octx):
    ctx = octx
    prob = capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, 16, 16, seed=1))
    _setup(ctx, prob, False, {"Number of Histories": 4096, "History Length": 100, "Weight Cutoff": 1e-3,
                              "Convergence Tolerance": 1e-6, "Maximum Iterations": 5, "Use Graph": 1,
                              "Birth Order": "State"})
    with pytest.raises(capi.GpuError):
        ctx.mcsa_solve(prob.b)
    with pytest.raises(capi.GpuError):
        ctx.set_params({"Birth Order": "Sorted"})
