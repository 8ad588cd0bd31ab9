"""Tally precision of FP32 walks on an adversarial tally: the walks run
from birth weights down to 1e-9 of them (weight cutoff 1e-9 on rows whose
|H| sums are 0.6-0.9), with signs of both kinds, so every LDS tile slot adds
weights eight decades apart.  With the paths bit-identical to the oracle's
(which sums the same FP32 weights in FP64), the tally's error is the error of
its summation alone:

* FP64 HBM tally (untiled FP32 walk), FP64 tile partial sums and
  compensated FP32 partial sums: within 1e-12 of the oracle;
* plain FP32 partial sums ("Walk"): the rounding of each grab's FP32 sums
  shows, far above 1e-12 (the reason the compensated form exists)."""
import numpy as np
import pytest

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def adversarial():
    prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 14, seed=6, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)
    A = po.Csr(prob.rp, prob.col, prob.val32.astype(np.float64))
    dinv = po.jacobi_dinv(A)
    rhs = (prob.b - 0.45) * dinv
    rhs[::97] *= 1e3  # a few states with far larger sources
    return prob, A, dinv, rhs


@pytest.mark.parametrize("source", ["Stratified", "Random"])
def test_tally_precision_against_oracle(adversarial, source):
    prob, A, dinv, rhs = adversarial
    nh = 64 * prob.n
    params = {"Walk Precision": "FP32", "Transition Sampler": "Alias Strided", "Source Sampling Type": source,
              "Estimator Type": "Collision", "Weight Cutoff": 1e-9, "History Length": 200,
              "Number of Histories": nh, "Random Seed": 8, "Walk Block Size": 64, "Walk Batch": 512}
    dom = po.build_domain(A, dinv)
    mp = po.mc_params(nh, 200, 1e-9, 32, po.EST_COLLISION,
                      po.SRC_STRATIFIED if source == "Stratified" else po.SRC_RANDOM, seed=8, num_threads=16,
                      sampler=po.SAMPLER_ALIAS_STRIDED)
    xo, steps_o = po.mc_solve(dom, rhs, mp, solve_index=2)
    err = {}
    with capi.Context(0) as ctx:
        ctx.set_problem(prob, "f32")
        ctx.set_point_jacobi()
        ctx.set_params(params)
        for tile, kind in ((0, "FP64"), (1024, "FP64"), (1024, "Compensated"), (1024, "Walk")):
            ctx.set_params({"Tally Tile": tile, "Tally Tile Precision": kind})
            x = ctx.mc_solve_raw(rhs, 2)
            assert ctx.stats().steps == steps_o
            err[(tile, kind)] = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert steps_o > 40 * nh  # long walks: weights span nine decades
    assert err[(0, "FP64")] <= 1e-12, err
    assert err[(1024, "FP64")] <= 1e-12, err
    assert err[(1024, "Compensated")] <= 1e-12, err
    assert err[(1024, "Walk")] > 1e-10, err
