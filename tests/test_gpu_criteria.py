"""Convergence criteria of the adjoint domain against dense linear algebra:
spectral radii of H, |H| and H* (numpy eigenvalues of the domain the device
built), ||H||_inf, the smallest row sum and the balance ratios
(src/MCLS_AlmostOptimalDomain_impl.hpp:439-576)."""
import numpy as np
import pytest
import scipy.sparse as sp

from mcls_amd import capi

pytestmark = pytest.mark.gpu


def _random_nonsym(n=60, seed=4):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.08, random_state=rng, format="lil")
    A = A.multiply(rng.choice([-1.0, 1.0], size=A.shape)).tolil()
    for i in range(n):
        A[i, i] = 1.5 * np.abs(A[i]).sum() + 0.5
    return sp.csr_matrix(A)


def _cases():
    out = []
    p = capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, 12, 12, seed=0))
    out.append(("poisson", sp.csr_matrix((p.values(), p.col, p.rp), shape=(p.n, p.n))))
    p = capi.build_problem(capi.make_spec(capi.ANISO2D_5PT, 10, 9, seed=4, p0=37.0))
    out.append(("aniso", sp.csr_matrix((p.values(), p.col, p.rp), shape=(p.n, p.n))))
    p = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 6, seed=2, p0=0.6, p1=0.9, p2=0.8, p3=1.2))
    out.append(("s27", sp.csr_matrix((p.values(), p.col, p.rp), shape=(p.n, p.n))))
    out.append(("random", _random_nonsym()))
    return out


def _dense_domain(ctx):
    rp, col, h, cdf, W = ctx.domain()
    n = len(rp) - 1
    M = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            M[i, col[k]] = h[k]
    return M, W


@pytest.mark.parametrize("name,A", _cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_criteria_match_dense(name, A):
    A.sort_indices()
    with capi.Context(0) as ctx:
        ctx.set_matrix(A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data)
        ctx.set_point_jacobi()
        cr = ctx.convergence_criteria(max_iters=40000, tol=1e-10)
        M, W = _dense_domain(ctx)
    P = np.abs(M)
    S = P * W[:, None]
    rho = [np.max(np.abs(np.linalg.eigvals(X))) for X in (M, P, S)]
    print(name, cr, rho)
    # |H| and H* are non-negative: a real Perron root, power iteration is exact
    assert abs(cr["rho_habs"] - rho[1]) <= 1e-5 * rho[1]
    assert abs(cr["rho_hstar"] - rho[2]) <= 1e-5 * rho[2]
    # H: a +-rho pair (bipartite stencils) is handled by the window average; a
    # complex pair of a non-normal H only to the window's oscillation
    tol_h = 1e-5 if name != "random" else 5e-2
    assert abs(cr["rho_h"] - rho[0]) <= tol_h * rho[0]
    rows = P.sum(axis=1)
    assert np.isclose(cr["inf_norm"], rows.max(), rtol=1e-14)
    assert np.isclose(cr["min_row_sum"], rows.min(), rtol=1e-14)
    nz = P.max(axis=1) > 0
    with np.errstate(divide="ignore"):  # an empty column gives inf, as in the reference
        ratio = P.max(axis=1)[nz] / P.max(axis=0)[nz]
    assert np.isclose(cr["balance_min"], ratio.min(), rtol=1e-14)
    assert np.isclose(cr["balance_max"], ratio.max(), rtol=1e-14)
    assert np.isclose(cr["balance_ave"], ratio.mean(), rtol=1e-12)


def test_criteria_printed_after_domain_build(capfd):
    p = capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, 8, 8, seed=0))
    with capi.Context(0) as ctx:
        ctx.set_problem(p)
        ctx.set_point_jacobi()
        ctx.set_params({"Compute Convergence Criteria": 1})
        ctx.build_domain()
        ctx.synchronize()
    out = capfd.readouterr().out
    assert "Neumann/Ulam Convergence Criteria" in out and "rho(H*) =" in out and "||H||_inf =" in out
    # Poisson 8x8 under Jacobi: rho(H) = rho(|H|) = cos(pi/9)
    line = [ln for ln in out.splitlines() if ln.startswith("rho(H+)")][0]
    assert abs(float(line.split("=")[1]) - np.cos(np.pi / 9)) <= 1e-5
