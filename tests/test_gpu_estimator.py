"""Correctness, second way: the GPU's per-iteration MC correction has the
mean of the reference estimator.  Over independent solves, the mean of the
GPU estimate (build semantics: quantised source, per-history streams, the
form's sampler and precision) and the mean of the reference-semantics
restatement of MCLS's estimator (one stream, |b| CDF with lower_bound, LIFO
source, weight sign(b) ||b||_1) agree within their standard errors (also for
the strided alias tables with the LDS tally tile), and both
match the truncated Neumann series sum_{k<L} H^k b that the adjoint
collision estimator targets."""
import numpy as np
import pytest

from mcls_amd import capi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FORMS = {
    "chosen": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "Alias"},
    "plain": {"Walk Precision": "FP64", "Source Sampling Type": "Random", "Transition Sampler": "CDF"},
    "fp32-cdf": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "CDF"},
    "strided": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Transition Sampler": "Alias Strided",
                "Tally Tile": 256, "Tally Tile Precision": "Walk"},
}


def _system():
    # a 27-point operator on 5^3 (rows of |H| sum to [0.6, 0.9]) and a signed
    # source, like a residual
    prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 5, seed=8, p0=0.6, p1=0.9, p2=0.8, p3=1.2))
    rng = np.random.default_rng(3)
    return prob, rng.random(prob.n) - 0.4


@pytest.mark.parametrize("form", sorted(FORMS))
def test_mc_correction_mean_matches_reference_estimator(form):
    prob, v = _system()
    n = prob.n
    A = po.Csr(prob.rp, prob.col, prob.values())
    dinv = po.jacobi_dinv(A)
    rhs = v * dinv
    L, nh, reps = 8, 20000, 48
    H = np.eye(n) - dinv[:, None] * A.to_dense()
    exact = sum(np.linalg.matrix_power(H, k) @ rhs for k in range(L))
    with capi.Context(0) as ctx:
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(dict(FORMS[form], **{"Estimator Type": "Collision", "Weight Cutoff": 0.0, "History Length": L,
                                            "Number of Histories": nh, "Random Seed": 31}))
        xg = np.array([ctx.mc_solve_raw(rhs, 500 + r) for r in range(reps)])
    dom = po.build_domain(A, dinv)
    xr = np.array([po.ref_mc_solve(dom, rhs, po.mc_params(nh, L), rng_state=r + 1)[0] for r in range(reps)])
    mg, sg = xg.mean(0), xg.std(0, ddof=1) / np.sqrt(reps)
    mr, sr = xr.mean(0), xr.std(0, ddof=1) / np.sqrt(reps)
    z = (mg - mr) / np.sqrt(sg ** 2 + sr ** 2)
    assert np.all(np.abs(z) < 5.0), np.abs(z).max()
    assert 0.2 < np.mean(z * z) < 3.0
    for m, s in ((mg, sg), (mr, sr)):
        ze = (m - exact) / np.maximum(s, 1e-300)
        assert np.all(np.abs(ze) < 5.0) and 0.2 < np.mean(ze * ze) < 3.0
    # a plausible mistake (walking H instead of H^T) is many errors away
    wrong = sum(np.linalg.matrix_power(H.T, k) @ rhs for k in range(L))
    assert np.max(np.abs(mg - wrong) / sg) > 8
