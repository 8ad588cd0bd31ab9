"""mcls_gpu_mc_solve_noalloc (the MC solve that allocates nothing, its
per-call buffers in a static device array) against mcls_gpu_mc_solve on
twin contexts: same problem, same parameters, same solve numbers, so the
same histories walk the same paths.  Paths, step and history counts are
identical; the tallies differ only by the order of the atomic additions
(FP64 tallies to 1e-12 relative, tiles of FP32 partial sums to 2e-5).
Two calls in a row, the second on a smaller system, so that what the first
call left in the shared array must not leak into the second."""
import numpy as np
import pytest

from mcls_amd import capi

pytestmark = pytest.mark.gpu

PLAIN = {"Walk Precision": "FP64", "Transition Sampler": "CDF", "Source Sampling Type": "Random",
         "Weight Cutoff": 1e-4, "History Length": 2000, "Random Seed": 11}
CHOSEN = {"Walk Precision": "FP32", "Transition Sampler": "Alias Strided", "Source Sampling Type": "Stratified",
          "Weight Cutoff": 1e-1, "History Length": 500, "Tally Tile": 1024, "Tally Tile Precision": "Walk",
          "Walk Block Size": 64, "Walk Batch": 512, "Random Seed": 12}


def _poisson(m):
    return capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, m, m, seed=m))


def _s27(m):
    return capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, m, seed=m, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)


def _setup(ctx, prob, dt, params, order):
    if order is not None:
        ctx.set_state_order(order)
    ctx.set_problem(prob, dt)
    ctx.set_point_jacobi()
    ctx.set_params(params)
    ctx.build_domain()


def _twin_solves(prob, dt, params, order, calls):
    """x from mcls_gpu_mc_solve and from the allocation-free entry point,
    `calls` solves in a row on each context, device vectors both ways."""
    out = []
    with capi.Context(0) as ca, capi.Context(0) as cb:
        _setup(ca, prob, dt, params, order)
        _setup(cb, prob, dt, params, order)
        b = capi.DeviceArray(prob.n)
        xa, xb = capi.DeviceArray(prob.n), capi.DeviceArray(prob.n)
        b.upload(prob.b)
        for _ in range(calls):
            rc = ca.L.mcls_gpu_mc_solve(ca.h, b.ptr, xa.ptr, capi.MCLS_DEVICE)
            assert rc == capi.MCLS_OK, ca.L.mcls_gpu_last_error()
            cb.mc_solve_noalloc(b.ptr, xb.ptr)
            sa, sb = ca.stats(), cb.stats()
            out.append((xa.download(), xb.download(), sa, sb))
        for d in (b, xa, xb):
            d.close()
    return out


def _check(out, rel):
    for xa, xb, sa, sb in out:
        assert np.all(np.isfinite(xb))
        assert sb.steps == sa.steps and sb.histories == sa.histories and sb.solve_index == sa.solve_index
        assert sb.walk_launches == sa.walk_launches == 1
        err = np.linalg.norm(xb - xa) / np.linalg.norm(xa)
        assert err <= rel, err


def test_noalloc_matches_mc_solve_plain_then_smaller():
    big = _twin_solves(_poisson(48), "f64", PLAIN, None, 2)
    small = _twin_solves(_poisson(20), "f64", PLAIN, None, 1)
    _check(big, 1e-12)
    _check(small, 1e-12)
    # consecutive calls are different solves (fresh streams)
    assert not np.array_equal(big[0][1], big[1][1])


def test_noalloc_matches_mc_solve_chosen_ordered_then_smaller():
    m = 16
    order = capi.nested_brick_order(m, m, m, [(8, 8, 8), (4, 4, 4)])
    big = _twin_solves(_s27(m), "f32", CHOSEN, order, 2)
    k = 8
    small = _twin_solves(_s27(k), "f32", CHOSEN, capi.brick_order(k, k, k, 4, 4, 4), 1)
    _check(big, 2e-5)
    _check(small, 2e-5)


def test_noalloc_untiled_fp32_then_smaller():
    params = dict(CHOSEN, **{"Tally Tile": 0, "Transition Sampler": "Alias", "Source Sampling Type": "Random"})
    _check(_twin_solves(_s27(12), "f32", params, None, 1), 2e-5)
    _check(_twin_solves(_s27(6), "f32", params, None, 1), 2e-5)


def test_noalloc_errors():
    L = capi.gpu_lib()
    prob = _poisson(12)
    with capi.Context(0) as ctx:
        b, x = capi.DeviceArray(prob.n), capi.DeviceArray(prob.n)
        b.upload(prob.b)
        ctx.set_problem(prob)
        ctx.set_point_jacobi()
        ctx.set_params(PLAIN)
        # no domain yet: the entry point builds nothing
        assert L.mcls_gpu_mc_solve_noalloc(ctx.h, b.ptr, x.ptr) == capi.MCLS_ERR_STATE
        ctx.build_domain()
        assert L.mcls_gpu_mc_solve_noalloc(ctx.h, None, x.ptr) == capi.MCLS_ERR_INVALID_ARG
        # a parameter that changes the domain makes it stale again
        ctx.set_params({"Neumann Relaxation": 0.9})
        assert L.mcls_gpu_mc_solve_noalloc(ctx.h, b.ptr, x.ptr) == capi.MCLS_ERR_STATE
        ctx.build_domain()
        ctx.mc_solve_noalloc(b.ptr, x.ptr)
        assert np.all(np.isfinite(x.download()))
        b.close()
        x.close()
    # more unknowns than the scratch array holds
    n = capi.MCLS_NOALLOC_MAX_N + 1
    rp = np.arange(n + 1, dtype=np.int64)
    col = np.arange(n, dtype=np.int32)
    with capi.Context(0) as ctx:
        ctx.set_matrix(rp, col, np.full(n, 2.0))
        # refused on the host before either vector is touched
        b, x = capi.DeviceArray(1), capi.DeviceArray(1)
        rc = L.mcls_gpu_mc_solve_noalloc(ctx.h, b.ptr, x.ptr)
        assert rc == capi.MCLS_ERR_INVALID_ARG
        assert b"scratch limit" in L.mcls_gpu_last_error()
        b.close()
        x.close()
