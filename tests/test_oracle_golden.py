"""The CPU oracle pinned against the reference's known-answer tests and
published vectors (no GPU).  Each test names the reference test it
transcribes; paths are relative to the MCLS source root.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po


def test_xorshift_marsaglia_vector():
    # Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003: the 64-bit
    # (13,7,17) generator from x = 88172645463325252 starts with these
    # outputs.  Same recurrence as src/MCLS_Xorshift.hpp:103-109.
    s = C.c_uint64(88172645463325252)
    got = [po.lib().orc_xorshift64(C.byref(s)) for _ in range(3)]
    assert got == [8748534153485358512, 3040900993826735515, 3453997556048239312]


def test_splitmix64_published_vector():
    # SplitMix64 from state 1234567 (the values used by the reference
    # implementations of Vigna's xoshiro family seeding).
    st, out = 1234567, []
    for _ in range(5):
        out.append(po.lib().orc_splitmix64(st))
        st = (st + 0x9E3779B97F4A7C15) % 2**64
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_uniform_mappings_are_exact():
    L = po.lib()
    assert L.orc_u01_f64(0) == 0.0
    assert L.orc_u01_f64(2**64 - 1) == 1.0 - 2.0**-53
    assert L.orc_u01_f32(2**64 - 1) == np.float32(1.0 - 2.0**-24)
    assert L.orc_u01_f64(1 << 63) == 0.5


def test_sampling_tools_multi_bin():
    # test/tstSamplingTools.cpp:65-93
    cdf = np.array([0.13, 0.27, 0.44, 0.69, 1.00])
    cases = [(0.13, 0), (0.27, 1), (0.44, 2), (0.69, 3), (1.00, 4), (0.00, 0),
             (0.12999999, 0), (0.13000001, 1), (0.26999999, 1), (0.27000001, 2),
             (0.43999999, 2), (0.44000001, 3), (0.68999999, 3), (0.69000001, 4),
             (0.99999999, 4)]
    for u, k in cases:
        assert po.lib().orc_lower_bound_f64(cdf, 5, u) == k, u
        # FP32 tables: the 1e-8 offsets vanish in single precision, so only
        # the exact edges and the interior points carry over.
        if np.min(np.abs(cdf - u)) == 0.0 or np.min(np.abs(cdf - u)) > 1e-6:
            assert po.lib().orc_lower_bound_f32(cdf.astype(np.float32), 5, np.float32(u)) == k, u


def test_sampling_tools_one_bin():
    # test/tstSamplingTools.cpp:96-101
    cdf = np.array([1.0])
    for u in (0.13, 0.27, 0.44, 0.69, 1.0):
        assert po.lib().orc_lower_bound_f64(cdf, 1, u) == 0


def test_lower_bound_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(1, 30))
        cdf = np.sort(rng.random(n))
        cdf[-1] = 1.0
        u = float(rng.random())
        assert po.lib().orc_lower_bound_f64(cdf, n, u) == int(np.searchsorted(cdf, u, side="left"))
    q = np.cumsum(rng.integers(0, 5, 50)).astype(np.uint64)
    for t in range(int(q[-1])):
        assert po.lib().orc_upper_bound_u64(q, 50, t) == int(np.searchsorted(q, t, side="right"))


def _diag(n, v):
    return po.Csr(np.arange(n + 1), np.arange(n), np.full(n, v))


def test_domain_diagonal_transition_weight():
    # tpetra/test/tstTpetraAlmostOptimalDomain.cpp:204-259: A = diag(3), one
    # rank: processTransition keeps the state and sets w -> -w(3p - 1).
    n = 10
    dom = po.build_domain(_diag(n, 3.0))
    for i in range(n):
        ev, state, w, steps, _ = po.ref_transport(dom, i, 3.0, history_length=1, rng_state=i + 1)
        assert state == i and steps == 1 and w == -3.0 * (3 - 1)


def test_domain_transporter_cutoff():
    # tpetra/test/tstTpetraDomainTransporter.cpp:97-172: A = diag(0.5), L = 1,
    # w = 3: tally x_i = 3, final weight 1.5, event CUTOFF, state unchanged.
    n = 10
    dom = po.build_domain(_diag(n, 0.5))
    x = np.zeros(n)
    for i in range(n):
        ev, state, w, steps, x = po.ref_transport(dom, i, 3.0, history_length=1, x=x)
        assert ev == 1 and state == i and w == 1.5
    assert np.all(x == 3.0)


def test_adjoint_tally_and_normalize():
    # tpetra/test/tstTpetraAdjointTally.cpp:83-167: tally 2 per state, then
    # normalize by nh = 10 -> 0.2.  A walk of length 0 is exactly one tally.
    n = 10
    dom = po.build_domain(_diag(n, 0.5))
    b = np.full(n, 2.0)
    mp = po.mc_params(10, history_length=1)
    x, nh, steps = po.ref_mc_solve(dom, b, mp, stratified=True)
    # stratified: one history per state, birth weight ||b||_1 = 20, tally
    # 20 at birth; normalisation by the 10 emitted histories
    assert nh == 10 and np.allclose(x, 20.0 / 10)


def _shift_matrix(n):
    # tpetra/test/tstTpetraUniformAdjointSource.cpp:95-107 (one rank)
    rp, col, val = [0], [], []
    for i in range(n):
        if i >= 1:
            col.append(i - 1)
            val.append(-0.5)
        if i == n - 1:
            if col and col[-1] == n - 1:
                pass
            col.append(n - 1)
            val.append(-0.5)
        rp.append(len(col))
    return po.Csr(rp, col, val, n)


@pytest.mark.parametrize("mult", [1, 10])
def test_uniform_adjoint_source_counts(mult):
    # tpetra/test/tstTpetraUniformAdjointSource.cpp:75-251 with b = -1:
    # numToTransport = mult * N, every history weight = -N (sign(b) ||b||_1).
    n = 10
    b = np.full(n, -1.0)
    nh, counts = po.ref_build_source(b, mult * n, stratified=False, ref_quirk=False)
    assert nh == mult * n and counts.sum() == mult * n
    # reference defect (src/MCLS_UniformAdjointSource_impl.hpp:185): the signed
    # b[0] seeds the CDF, so the count drops to 8 * mult for mult = 1.
    nq, _ = po.ref_build_source(b, mult * n, stratified=False, ref_quirk=True)
    assert nq < mult * n
    dom = po.build_domain(_shift_matrix(n))
    mp = po.mc_params(mult * n, history_length=1)
    x, nh2, _ = po.ref_mc_solve(dom, b, mp)
    assert nh2 == mult * n


def test_richardson_fifteen_iterations():
    # tpetra/test/tstTpetraFixedPointSolverManager.cpp:72-190: 10x10 matrix,
    # tol 1e-8 (infinity norm), Richardson converges in exactly 15 iterations
    # for b = -1, 2, -2, with solutions of the sign of b.
    n = 10
    A = np.zeros((n, n))
    A[0, 0:3] = [0.14, 0.14, 1.0]
    for i in range(1, n - 1):
        A[i, i - 1:i + 2] = [0.14, 1.0, 0.14]
    A[n - 1, n - 3:n] = [0.14, 0.14, 1.0]
    Ac = po.Csr.from_dense(A)
    for bv in (-1.0, 2.0, -2.0):
        x, res = po.fixed_point_solve(Ac, np.full(n, bv), params=po.solver_params(tol=1e-8, max_iters=100))
        assert res.iters == 15 and res.converged == 1 and res.achieved_tol > 0
        assert np.all(np.sign(x) == np.sign(bv))


def test_point_jacobi():
    # tpetra/test/tstTpetraPointJacobiPreconditioner.cpp:80-189: diag 2 -> 0.5
    n = 10
    assert np.all(po.jacobi_dinv(_diag(n, 2.0)) == 0.5)
    A = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, 2.0), 1) + np.diag(np.full(n - 1, 2.0), -1)
    A[0, 1] = A[n - 1, n - 2] = 0.0
    assert np.all(po.jacobi_dinv(po.Csr.from_dense(A)) == 0.5)


def test_composite_operator_and_domain_closed_form():
    # The domain rows are the rows of H^T, H = I - omega D^-1 A, with
    # |h| < DBL_EPSILON dropped (src/MCLS_AlmostOptimalDomain_impl.hpp:296-364);
    # checked against dense linear algebra on a random nonsymmetric matrix.
    rng = np.random.default_rng(5)
    n = 40
    A = (rng.random((n, n)) < 0.15) * rng.normal(size=(n, n))
    A += np.diag(3.0 + rng.random(n))
    Ac = po.Csr.from_dense(A)
    for omega in (1.0, 0.6):
        dinv = po.jacobi_dinv(Ac)
        dom = po.build_domain(Ac, dinv, omega)
        H = np.eye(n) - omega * (dinv[:, None] * A)
        HT = H.T
        for i in range(n):
            cols = dom.col[dom.rp[i]:dom.rp[i + 1]]
            hs = dom.h[dom.rp[i]:dom.rp[i + 1]]
            nz = np.nonzero(np.abs(HT[i]) >= np.finfo(float).eps)[0]
            stored = set(np.nonzero(A[:, i])[0]) | {i} if A[i, i] else set(np.nonzero(A[:, i])[0])
            expect = [j for j in nz if j in stored]
            assert list(cols) == expect
            assert np.allclose(hs, HT[i, cols], rtol=1e-15, atol=0)
            W = np.abs(hs).sum()
            assert abs(dom.W[i] - W) <= 1e-14 * W
            cdf = dom.cdf[dom.rp[i]:dom.rp[i + 1]]
            assert cdf[-1] == 1.0 and np.all(np.diff(cdf) >= 0)
            assert np.allclose(cdf, np.cumsum(np.abs(hs)) / W, rtol=1e-14)


def test_composite_threshold_filter():
    # tpetra/src/MCLS_TpetraHelpers.hpp:196-263: entries with |m| <= max(row)
    # * threshold dropped from each row of the composite (only with a
    # preconditioner).  Row max is the signed maximum.
    A = np.array([[4.0, -1.0, 0.0], [-0.1, 4.0, -2.0], [0.0, -1.0, 4.0]])
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    M = (dinv[:, None] * A).T  # composite rows
    dom = po.build_domain(Ac, dinv, 1.0, 0.3)
    for i in range(3):
        thr = M[i][A[:, i] != 0].max() * 0.3
        keep = [j for j in range(3) if A[j, i] != 0 and abs(M[i, j]) > thr and j != i]
        assert list(dom.col[dom.rp[i]:dom.rp[i + 1]]) == keep
    # without a preconditioner nothing is filtered
    dom0 = po.build_domain(Ac, None, 1.0, 0.3)
    assert dom0.nnz == Ac.rp[-1]


def test_domain_errors():
    bad = po.Csr([0, 1, 2], [0, 7], [1.0, 1.0], 2)
    with pytest.raises(ValueError):
        po.build_domain(bad)
    dup = po.Csr([0, 2, 3], [0, 0, 1], [1.0, 2.0, 1.0], 2)
    with pytest.raises(ValueError):
        po.build_domain(dup)
    with pytest.raises(ValueError):
        po.jacobi_dinv(po.Csr([0, 1, 2], [1, 0], [1.0, 1.0], 2))
