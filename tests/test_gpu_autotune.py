"""Walk autotuner: it times the candidate launch shapes, stores the fastest,
and the estimates do not depend on the shape (paths are bit-identical; FP64
tallies agree to rounding of the atomic summation order)."""
import numpy as np
import pytest

from mcls_amd import capi

pytestmark = pytest.mark.gpu


def _poisson(ctx, sampler, prec):
    prob = capi.build_problem(capi.make_spec(capi.POISSON2D_5PT, 64, 64, seed=3))
    ctx.set_problem(prob)
    ctx.set_point_jacobi()
    ctx.set_params({"Walk Precision": f"FP{prec}", "Transition Sampler": sampler, "Source Sampling Type": "Stratified",
                    "History Length": 200, "Weight Cutoff": 1e-3, "Number of Histories": 400000,
                    "Random Seed": 21})
    return prob, prob.b * ctx.get_left_prec_diag()


@pytest.mark.parametrize("sampler,prec", [("Alias", 32), ("CDF", 64)])
def test_autotune_choice_and_shape_independence(ctx, sampler, prec):
    prob, rhs = _poisson(ctx, sampler, prec)
    before = ctx.stats().solve_index
    t = ctx.autotune(prob.b, pilot_histories=1 << 16)
    print("tuning", t)
    assert ctx.stats().solve_index == before
    assert t["candidates"] == 13
    assert t["block"] in (64, 128, 256) and t["batch"] in (64, 128, 256, 512)
    assert t["blocks_per_cu"] * t["block"] // 64 in (8, 16, 32)
    assert 0.0 < t["ms_best"] <= t["ms_default"]
    ref_p, ref_w = ctx.record_paths(rhs, 5, 0, 3000, 200)
    ref_x = ctx.mc_solve_raw(rhs, 6)
    for blk, bpc, bat in ((64, 32, 512), (256, 2, 64), (128, 8, 128)):
        ctx.set_params({"Walk Block Size": blk, "Walk Blocks Per CU": bpc, "Walk Batch": bat})
        p, w = ctx.record_paths(rhs, 5, 0, 3000, 200)
        assert np.array_equal(p, ref_p) and np.array_equal(w, ref_w)
        x = ctx.mc_solve_raw(rhs, 6)
        tol = 1e-12 if prec == 64 else 1e-4  # FP32 atomics: summation order
        assert np.linalg.norm(x - ref_x) <= tol * np.linalg.norm(ref_x)
    ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64})


def test_walk_batch_validation(ctx):
    _poisson(ctx, "CDF", 64)
    for bad in (0, 32, 100, 8192):
        with pytest.raises(capi.GpuError):
            ctx.set_params({"Walk Batch": bad})
