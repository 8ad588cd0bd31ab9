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


@pytest.mark.parametrize("sampler,prec,method", [("Alias", 32, "Adjoint"), ("CDF", 64, "Adjoint"),
                                                ("Alias", 32, "Forward")])
def test_autotune_choice_and_shape_independence(ctx, sampler, prec, method):
    prob, rhs = _poisson(ctx, sampler, prec)
    ctx.set_params({"Monte Carlo Type": method})
    before = ctx.stats().solve_index
    t = ctx.autotune(prob.b, pilot_histories=1 << 16)
    print("tuning", t)
    assert ctx.stats().solve_index == before
    # shapes, batches, tiles (not for forward walks), one full walk per block
    # size, one per neighbouring batch of the chosen one (one or two), and a
    # window pilot of the final shape when the full walks changed it
    base = (16 if method == "Adjoint" else 13) + 3
    assert base + 1 <= t["candidates"] <= base + 3
    assert t["tally_tile"] in (0, 128, 512, 1024)
    assert t["block"] in (64, 128, 256) and t["batch"] in (64, 128, 256, 512)
    assert t["blocks_per_cu"] * t["block"] // 64 in (8, 16, 32)
    assert t["ms_best"] > 0.0 and t["ms_confirmed"] > 0.0
    ref_p, ref_w = ctx.record_paths(rhs, 5, 0, 3000, 200)
    ref_x = ctx.mc_solve_raw(rhs, 6)
    for blk, bpc, bat in ((64, 32, 512), (256, 2, 64), (128, 8, 128)):
        ctx.set_params({"Walk Block Size": blk, "Walk Blocks Per CU": bpc, "Walk Batch": bat})
        p, w = ctx.record_paths(rhs, 5, 0, 3000, 200)
        assert np.array_equal(p, ref_p) and np.array_equal(w, ref_w)
        x = ctx.mc_solve_raw(rhs, 6)
        tol = 1e-12  # FP64 tallies (every walk precision): summation order only
        assert np.linalg.norm(x - ref_x) <= tol * np.linalg.norm(ref_x)
    ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64, "Tally Tile": 0,
                    "Monte Carlo Type": "Adjoint"})


def test_autotune_entry_tile_over_lds(ctx):
    """A caller's tile that does not fit the LDS at the entry block size (4
    waves x 2048 entries x 12 bytes > 64 KiB) is not timed: the search starts
    untiled and returns a shape whose tile fits."""
    prob, rhs = _poisson(ctx, "Alias", 32)
    ctx.set_params({"Tally Tile": 2048, "Tally Tile Precision": "FP64", "Walk Block Size": 256,
                    "Walk Blocks Per CU": 0})
    try:
        t = ctx.autotune(prob.b, pilot_histories=1 << 16)
        assert t["ms_default"] == -1.0 and t["ms_best"] > 0.0 and t["ms_confirmed"] > 0.0
        assert t["tally_tile"] in (0, 128, 512, 1024)
        assert t["block"] // 64 * t["tally_tile"] * 12 <= 65536
        ctx.mc_solve_raw(rhs, 1)  # the returned configuration runs
    finally:
        ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64, "Tally Tile": 0,
                        "Tally Tile Precision": "FP64"})


def test_walk_batch_validation(ctx):
    _poisson(ctx, "CDF", 64)
    for bad in (0, 32, 100, 8192):
        with pytest.raises(capi.GpuError):
            ctx.set_params({"Walk Batch": bad})


@pytest.mark.parametrize("tile_prec", ["FP64", "Walk", "Compensated"])
@pytest.mark.parametrize("sampler,prec", [("Alias", 32), ("CDF", 64), ("CDF", 32)])
def test_tally_tile_matches_direct_tally(ctx, sampler, prec, tile_prec):
    """The LDS tile changes only where the additions happen: same paths, same
    tallies to the rounding of the summation order.  Its partial sums are
    FP64, or in the walk's precision ("Walk"), and are flushed into an FP64
    tally either way, which every walk precision uses; compensated FP32
    partial sums ("Compensated") are as good as FP64 ones."""
    prob, rhs = _poisson(ctx, sampler, prec)
    ctx.set_params({"Tally Tile": 0, "Tally Tile Precision": tile_prec})
    ref = ctx.mc_solve_raw(rhs, 3)
    for tile, blk in ((64, 256), (512, 128), (1024, 64)):
        ctx.set_params({"Tally Tile": tile, "Walk Block Size": blk, "Walk Blocks Per CU": 0})
        x = ctx.mc_solve_raw(rhs, 3)
        tol = 1e-4 if (prec == 32 and tile_prec == "Walk") else 1e-12
        assert np.linalg.norm(x - ref) <= tol * np.linalg.norm(ref), tile
    ctx.set_params({"Tally Tile": 0, "Walk Block Size": 256})
    with pytest.raises(capi.GpuError):
        ctx.set_params({"Tally Tile": 100})
    with pytest.raises(capi.GpuError):
        ctx.set_params({"Tally Tile Precision": "FP16"})
    # 4 waves x 2048 entries x 12 bytes (FP64 or compensated partial sums)
    # > 64 KiB; with FP32 partial sums 8 bytes an entry fit exactly
    ctx.set_params({"Tally Tile": 2048, "Walk Block Size": 256})
    if tile_prec != "Walk" or prec == 64:
        with pytest.raises(capi.GpuError):
            ctx.mc_solve_raw(rhs, 4)
    else:
        x = ctx.mc_solve_raw(rhs, 3)
        assert np.linalg.norm(x - ref) <= 1e-4 * np.linalg.norm(ref)
    ctx.set_params({"Tally Tile": 0, "Tally Tile Precision": "FP64"})


@pytest.mark.parametrize("tile_prec", ["FP64", "Walk", "Compensated"])
def test_mcsa_with_tally_tile_converges(ctx, tile_prec):
    prob = capi.build_problem(capi.make_spec(capi.STENCIL3D_27PT, 16, seed=3, p0=0.6, p1=0.9, p2=0.8, p3=1.2),
                              want64=False, want32=True)
    ctx.set_problem(prob, "f32")
    ctx.set_point_jacobi()
    ctx.set_params({"Walk Precision": "FP32", "Transition Sampler": "Alias", "Source Sampling Type": "Stratified",
                    "History Length": 1000, "Weight Cutoff": 1e-2, "Number of Histories": 16 * prob.n,
                    "Random Seed": 3, "Convergence Tolerance": 1e-8, "Maximum Iterations": 50,
                    "Tally Tile": 256, "Walk Block Size": 256, "Walk Blocks Per CU": 0,
                    "Tally Tile Precision": tile_prec})
    x = ctx.mcsa_solve(prob.b)
    st = ctx.stats()
    ctx.set_params({"Tally Tile": 0, "Tally Tile Precision": "FP64"})
    assert st.converged and st.achieved_tol <= 1e-8
