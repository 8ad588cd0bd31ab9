"""The oracle's estimators pinned by what the mathematics fixes (no GPU).

On a small nonsymmetric system the adjoint estimators must reproduce
  collision:       E[x] = sum_{k=0}^{L-1} H^k b
  expected value:  E[x] = sum_{k=0}^{L}   H^k b
with H = I - omega D^-1 A (the walk samples rows of H^T), for every source
mode, sampler and precision.  A transposed operand, a tally after (instead
of before) the transition, or a dropped sign each move the mean by many
standard errors.  Also: alias tables reproduce the row probabilities, the
reference-semantics solve is unbiased, MCSA converges to the direct solution.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _system(n=12, seed=3):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 4.0 + rng.random()
        for j in rng.choice([k for k in range(n) if k != i], 3, replace=False):
            A[i, j] = -rng.random() * (1.0 if rng.random() < 0.7 else -1.0)
    b = rng.random(n) - 0.3
    return A, b


def _replicas(dom, b, mp, reps):
    xs = []
    for r in range(reps):
        x, _ = po.mc_solve(dom, b, mp, solve_index=1000 + r)
        xs.append(x)
    xs = np.array(xs)
    return xs.mean(0), xs.std(0, ddof=1) / np.sqrt(reps)


def _check_z(mean, se, exact):
    z = (mean - exact) / np.maximum(se, 1e-300)
    assert np.all(np.abs(z) < 5.0), z
    assert 0.2 < np.mean(z * z) < 3.0, np.mean(z * z)


@pytest.mark.parametrize("est", [po.EST_COLLISION, po.EST_EXPECTED_VALUE])
@pytest.mark.parametrize("source", [po.SRC_RANDOM, po.SRC_STRATIFIED])
@pytest.mark.parametrize("sampler", [po.SAMPLER_CDF, po.SAMPLER_ALIAS])
@pytest.mark.parametrize("prec", [64, 32])
def test_estimator_expectation(est, source, sampler, prec):
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    omega = 0.9
    dom = po.build_domain(Ac, dinv, omega)
    H = np.eye(n) - omega * dinv[:, None] * A
    L = 6
    terms = L + 1 if est == po.EST_EXPECTED_VALUE else L
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(terms))
    mp = po.mc_params(3000, L, 0.0, prec, est, source, seed=prec + 7 * est + 3 * source + sampler,
                      sampler=sampler)
    mean, se = _replicas(dom, b, mp, 60)
    _check_z(mean, se, exact)
    # a plausible mistake: walking H instead of H^T moves the mean away
    wrong = sum(np.linalg.matrix_power(H.T, k) @ b for k in range(terms))
    assert np.max(np.abs(mean - wrong) / np.maximum(se, 1e-300)) > 8


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("est", [po.EST_COLLISION, po.EST_EXPECTED_VALUE])
def test_learned_proposal_is_unbiased(prec, est):
    # P_ij ∝ |h_ij| g_j with weight factor S_i / g_j keeps E[x] for any g > 0
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv)
    H = np.eye(n) - dinv[:, None] * A
    L = 6
    terms = L + 1 if est == po.EST_EXPECTED_VALUE else L
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(terms))
    g = np.exp(np.random.default_rng(4).normal(size=n)).astype(np.float32)
    mp = po.mc_params(3000, L, 0.0, prec, est, po.SRC_STRATIFIED, seed=11, importance=g)
    mean, se = _replicas(dom, b, mp, 60)
    _check_z(mean, se, exact)
    # the proposal tables are a distribution and the factors invert it
    lcdf, lfac = po.build_proposal(dom, g)
    for i in range(n):
        s, e = dom.rp[i], dom.rp[i + 1]
        if e == s:
            continue
        assert lcdf[e - 1] == 1.0
        p = np.diff(np.concatenate([[0.0], lcdf[s:e]]))
        assert np.allclose(p * lfac[s:e], np.abs(dom.h[s:e]), rtol=1e-12)


@pytest.mark.parametrize("prec", [64, 32])
def test_learned_alias_proposal_is_unbiased(prec):
    """The learned proposal drawn by alias tables keeps E[x] too, and its
    tables reproduce P_ij = |h_ij| g_j / S_i with factors S_i / g_j."""
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv)
    H = np.eye(n) - dinv[:, None] * A
    L = 6
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(L))
    g = np.exp(np.random.default_rng(5).normal(size=n)).astype(np.float32)
    mp = po.mc_params(3000, L, 0.0, prec, po.EST_COLLISION, po.SRC_STRATIFIED, seed=13, importance=g,
                      sampler=po.SAMPLER_LEARNED_ALIAS)
    mean, se = _replicas(dom, b, mp, 60)
    _check_z(mean, se, exact)
    thr, alias, lfac = po.build_proposal_alias(dom, g)
    _, lfac_cdf = po.build_proposal(dom, g)
    assert np.array_equal(lfac, lfac_cdf)  # the same factors as the CDF form
    for i in range(n):
        s, e = dom.rp[i], dom.rp[i + 1]
        z = e - s
        if z == 0:
            continue
        wts = np.abs(dom.h[s:e]) * g[dom.col[s:e]].astype(np.float64)
        p = wts / wts.sum()
        q = np.zeros(z)
        for k in range(z):
            q[k] += thr[s + k] / z
            q[alias[s + k]] += (1.0 - thr[s + k]) / z
        assert np.max(np.abs(q - p)) < 1e-14
        assert np.allclose(p * lfac[s:e], np.abs(dom.h[s:e]), rtol=1e-12)


def test_reference_semantics_unbiased():
    A, b = _system(10, seed=8)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv)
    H = np.eye(10) - dinv[:, None] * A
    L = 5
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(L))
    xs = []
    for r in range(80):
        x, nh, _ = po.ref_mc_solve(dom, b, po.mc_params(2000, L), rng_state=r + 1)
        xs.append(x)
    xs = np.array(xs)
    _check_z(xs.mean(0), xs.std(0, ddof=1) / np.sqrt(len(xs)), exact)


def test_alias_tables_reproduce_row_probabilities():
    A, b = _system(30, seed=1)
    Ac = po.Csr.from_dense(A)
    dom = po.build_domain(Ac, po.jacobi_dinv(Ac))
    thr, alias = po.build_alias(dom)
    for i in range(dom.n):
        s, e = dom.rp[i], dom.rp[i + 1]
        z = e - s
        if z == 0:  # absorbing row (column of A with only its diagonal)
            continue
        p = np.abs(dom.h[s:e]) / dom.W[i]
        q = np.zeros(z)
        for k in range(z):
            q[k] += thr[s + k] / z
            q[alias[s + k]] += (1.0 - thr[s + k]) / z
        assert np.max(np.abs(q - p)) < 1e-14


def test_source_quantisation_is_exact_and_proportional():
    rng = np.random.default_rng(2)
    b = rng.normal(size=1000) * np.exp(rng.normal(size=1000) * 3)
    b[::50] = 0.0
    Q, q, cdf = po.source_quantize(b)
    assert Q == int(cdf[-1]) and np.all(np.diff(cdf.astype(object)) >= 0)
    nz = b != 0
    assert np.all(q[~nz] == 0) and np.all(q[nz] >= 1)
    # q_i = ceil(|b_i| 2^(S-e)): relative error below 2^(S-e)/|b_i|-scale
    ratio = q[nz].astype(float) / np.abs(b[nz])
    big = np.abs(b[nz]) > np.abs(b).max() * 1e-6
    assert np.allclose(ratio[big], ratio[big][0], rtol=1e-9)


def test_mcsa_oracle_converges_to_direct_solution():
    A, b = _system(40, seed=5)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    mp = po.mc_params(4000, 50, 1e-3, 64, seed=4)
    x, res, hist = po.mcsa_solve(Ac, b, dinv=dinv, params=po.solver_params(tol=1e-10, max_iters=60, mc=mp))
    assert res.converged == 1 and res.achieved_tol <= 1e-10
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    # reference semantics, 4 sets in threads
    sp = po.solver_params(tol=1e-10, max_iters=60, num_sets=4, reference_mode=1,
                          mc=po.mc_params(1000, 50, 1e-3, 64, seed=9))
    x2, res2, _ = po.mcsa_solve(Ac, b, dinv=dinv, params=sp)
    assert res2.converged == 1
    assert np.linalg.norm(x2 - xd) <= 1e-8 * np.linalg.norm(xd)


@pytest.mark.parametrize("sampler", [po.SAMPLER_CDF, po.SAMPLER_ALIAS])
@pytest.mark.parametrize("prec", [64, 32])
def test_forward_estimator_expectation(sampler, prec):
    """Forward method (rows of H, histories from every state, each history's
    sum of w b[state] averaged at its start state): E[x] = sum_{k<L} H^k b;
    walking H^T instead (the adjoint domain) moves the mean away."""
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    omega = 0.9
    dom = po.build_domain(Ac, dinv, omega, forward=True)
    H = np.eye(n) - omega * dinv[:, None] * A
    L = 6
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(L))
    mp = po.mc_params(250 * n, L, 0.0, prec, po.EST_FORWARD, seed=11 + prec + sampler, sampler=sampler)
    mean, se = _replicas(dom, b, mp, 60)
    _check_z(mean, se, exact)
    wrong = sum(np.linalg.matrix_power(H.T, k) @ b for k in range(L))
    assert np.max(np.abs(mean - wrong) / np.maximum(se, 1e-300)) > 8


def test_forward_domain_is_rows_of_h():
    A, b = _system()
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv, 0.8, forward=True)
    M = A * dinv[:, None]  # (PL A)_ij = a_ij / a_ii, rounded as the build does
    H = -(0.8 * M)
    H[np.diag_indices_from(H)] = 1.0 - 0.8 * np.diag(M)
    dense = np.zeros_like(H)
    for i in range(len(b)):
        for k in range(dom.rp[i], dom.rp[i + 1]):
            dense[i, dom.col[k]] = dom.h[k]
        assert np.all(np.diff(dom.col[dom.rp[i]:dom.rp[i + 1]]) > 0)  # ascending columns
    H[np.abs(H) < np.finfo(float).eps] = 0.0
    assert np.array_equal(dense, H)
    # and the adjoint domain is its transpose
    adj = po.build_domain(Ac, dinv, 0.8)
    dense_t = np.zeros_like(H)
    for i in range(len(b)):
        for k in range(adj.rp[i], adj.rp[i + 1]):
            dense_t[i, adj.col[k]] = adj.h[k]
    assert np.array_equal(dense_t, dense.T)


def test_forward_source_and_one_step_tally():
    """tstTpetraUniformForwardSource (histories = N x ceil(Sample Ratio),
    emitted state by state, weight 1) and tstTpetraForwardTally (a history
    that collides once tallies w b at its start state)."""
    assert po.lib().orc_forward_per_state(10, 10) == 1
    assert po.lib().orc_forward_per_state(10, 25) == 3
    assert po.lib().orc_forward_per_state(10, 0) == 1
    A = po.Csr.from_dense(np.array([[1.0, -0.25, 0.0], [-0.25, 1.0, -0.25], [0.0, -0.25, 1.0]]))
    dinv = po.jacobi_dinv(A)
    dom = po.build_domain(A, dinv, forward=True)
    b = np.array([3.0, -2.0, 0.5])
    mp = po.mc_params(3 * 4, 1, 0.0, 64, po.EST_FORWARD, seed=9)
    x, steps = po.mc_solve(dom, b, mp)
    assert np.array_equal(x, b) and steps == 12
    paths, w = po.record_paths(dom, b, mp, 0, 12, 1)
    assert np.array_equal(paths[:, 0], np.repeat(np.arange(3), 4))
    # one transition: w = W_start sign(h) with W = 0.25, 0.5, 0.25 and h > 0
    assert np.array_equal(w, np.repeat([0.25, 0.5, 0.25], 4))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("est", [po.EST_COLLISION, po.EST_FORWARD])
def test_strided_alias_is_unbiased(prec, est):
    """Alias tables padded to a fixed stride (zero-probability slots) keep
    E[x]; their slots reproduce |h_ij| / W_i, padding slots are never kept,
    and the stride is the longest row's length."""
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    H = np.eye(n) - dinv[:, None] * A
    forward = est == po.EST_FORWARD
    dom = po.build_domain(Ac, dinv, forward=forward)
    L = 6
    # adjoint collision: sum_k (H^T)^k-walks of the H^T domain give sum_k H^k b;
    # forward: rows of H, the same series
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(L))
    mp = po.mc_params(3000, L, 0.0, prec, est, po.SRC_STRATIFIED, seed=17, sampler=po.SAMPLER_ALIAS_STRIDED)
    mean, se = _replicas(dom, b, mp, 60)
    _check_z(mean, se, exact)
    stride, thr, alias = po.build_alias_strided(dom)
    lens = np.diff(dom.rp)
    assert stride == max(lens.max(), 1)
    for i in range(n):
        z = lens[i]
        if z == 0:
            continue
        p = np.zeros(stride)
        for k in range(stride):
            p[k] += thr[i, k] / stride
            p[alias[i, k]] += (1.0 - thr[i, k]) / stride
        target = np.zeros(stride)
        target[:z] = np.abs(dom.h[dom.rp[i]:dom.rp[i + 1]]) / dom.W[i]
        assert np.max(np.abs(p - target)) < 1e-14
        assert np.all(thr[i, z:] == 0.0) and np.all(alias[i] < z)


@pytest.mark.parametrize("est", [po.EST_COLLISION, po.EST_EXPECTED_VALUE])
@pytest.mark.parametrize("source", [po.SRC_RANDOM, po.SRC_STRATIFIED])
def test_windows_partition_the_solve(est, source):
    # the windows [0, a), [a, c), [c, end) of one solve add up to the whole
    # solve (minus the expected-value birth term), and a window past the end
    # walks nothing
    A, b = _system(30, seed=5)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv)
    mp = po.mc_params(4000, 8, 0.0, 64, est, source, seed=19)
    full, steps = po.mc_solve(dom, b, mp, solve_index=4)
    parts = [po.mc_solve_window(dom, b, mp, f, c, solve_index=4) for f, c in ((0, 1000), (1000, 1700), (2700, 10**6))]
    total = sum(x for x, _ in parts)
    if est == po.EST_EXPECTED_VALUE:
        total = total + b
    assert np.allclose(total, full, rtol=1e-12, atol=1e-12)
    assert sum(s for _, s in parts) == steps
    empty, s0 = po.mc_solve_window(dom, b, mp, 10**7, 50, solve_index=4)
    assert s0 == 0 and not np.any(empty)


@pytest.mark.parametrize("source", [po.SRC_RANDOM, po.SRC_STRATIFIED])
def test_birth_magnitudes_keep_the_estimator_unbiased(source):
    # births drawn by |mag| instead of |b| (a batch's reallocation) with
    # weights b_k Q / q_k: the same expectation; mag = |b| is the plain solve
    A, b = _system()
    n = len(b)
    Ac = po.Csr.from_dense(A)
    dinv = po.jacobi_dinv(Ac)
    dom = po.build_domain(Ac, dinv)
    H = np.eye(n) - dinv[:, None] * A
    L = 5
    exact = sum(np.linalg.matrix_power(H, k) @ b for k in range(L))
    mp = po.mc_params(3000, L, 0.0, 64, po.EST_COLLISION, source, seed=23)
    x0, _ = po.mc_solve(dom, b, mp, solve_index=2)
    x1, _ = po.mc_solve_window(dom, b, mp, 0, 10**9, solve_index=2, mag=np.abs(b))
    assert np.array_equal(x0, x1)
    mag = np.abs(b) * np.linspace(0.2, 3.0, n)
    xs = np.array([po.mc_solve_window(dom, b, mp, 0, 10**9, solve_index=500 + r, mag=mag)[0] for r in range(80)])
    _check_z(xs.mean(0), xs.std(0, ddof=1) / np.sqrt(len(xs)), exact)
    # mag decides where histories are born: all of it on state k with one
    # collision per history (L = 1) tallies b_k at k and nothing elsewhere
    k = 7
    one = np.zeros(n)
    one[k] = 2.5
    m1 = po.mc_params(500, 1, 0.0, 64, po.EST_COLLISION, source, seed=29)
    x, steps = po.mc_solve_window(dom, b, m1, 0, 10**9, solve_index=1, mag=one)
    expect = np.zeros(n)
    expect[k] = b[k]
    assert np.allclose(x, expect, rtol=1e-14, atol=0) and steps == 500


@pytest.mark.parametrize("kind", ["stencil", "ragged"])
def test_residual_equals_scipy_bit_for_bit(kind):
    """orc_prec_residual (updatePrecResidual without PR): rows summed left to
    right in FP64 without contraction, then PL (b - s).  scipy's CSR
    product sums each row in stored order too, so the two agree bit for bit
    (a dropped entry, a wrong column or a misplaced PL would not)."""
    import scipy.sparse as sp

    rng = np.random.default_rng(7)
    if kind == "stencil":
        from mcls_amd import capi
        prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 9, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2))
        rp, col, val, n = prob.rp, prob.col, prob.values(), prob.n
    else:
        n = 300
        lens = rng.integers(0, 30, n)
        rp = np.zeros(n + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        col = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lens]).astype(np.int32)
        val = rng.standard_normal(int(rp[-1]))
    A = po.Csr(rp, col, val)
    S = sp.csr_matrix((val, col, rp), shape=(n, n))
    b = rng.standard_normal(n)
    x = rng.standard_normal(n)
    d = rng.random(n) + 0.5
    assert np.array_equal(po.prec_residual(A, b, x, None), b - S @ x)
    assert np.array_equal(po.prec_residual(A, b, x, d), d * (b - S @ x))
