"""Benchmark: MCSA time-to-tolerance and histories/s on one MI355X.

Workload (default): 3-D 27-point stencil on a 256^3 grid (16,777,216 unknowns,
449,455,096 nonzeros, CSR with FP32 values and int32 columns), point-Jacobi
preconditioning, 2^28 histories per MCSA iteration, tolerance 1e-8 on
||rp||_2 / ||PL b||_2.  One step = one full MCSA solve from x = 0 to the
tolerance, with the matrix, the domain and b resident in HBM.

Forms:
  chosen: FP32 walks and tallies, FP64 residual/update/convergence test,
          stratified source, relative weight cutoff 1e-1, and the learned
          proposal with plain almost-optimal (alias) sampling as the
          fallback: before the timed solves a network is trained on this
          operator's own adjoint tally and the lowest variance x walk time
          of plain / diagonal importance / network on pilot solves is kept.
  plain:  FP64 walks and tallies, random source, CDF inversion (the
          reference's sampler), weight cutoff 1e-4.
  forward: the forward method with the chosen walk (reported beside).
value = histories/s of the chosen form over the K timed solves.

The walk-kernel roofline uses the per-step algorithmic bytes of the walk
(bytes_per_step below) times the transitions of one launch, divided by that
launch's HIP-event duration measured by the library on its own stream.

cpu_baseline: the oracle's restatement of the MCLS reference algorithm
(FP64, per-row CDF + lower_bound, one Xorshift stream per set, multi-set
threads summed / numSets) timed on one MCSA iteration of a bounded history
sample on the same operator (see the printed sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mcls_amd import capi  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "MCSA wall-time to 1e-8 rel. residual (plain vs chosen form); histories/sec"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


WORKLOADS = {
    "cfg4": dict(kind=capi.STENCIL3D_27PT, nx=256, seed=4, p=(0.6, 0.9, 0.8, 1.2), f32=True,
                 histories=1 << 28, bricks=((32, 32, 32), (8, 8, 8), (4, 4, 4)), label="3-D 27-point stencil 256^3 (16.8M unknowns, 449M nnz, "
                 "FP32 CSR), point Jacobi, 2^28 histories/iteration, tol 1e-8"),
    "cfg4-64": dict(kind=capi.STENCIL3D_27PT, nx=64, seed=4, p=(0.6, 0.9, 0.8, 1.2), f32=True,
                    histories=16 * 64 ** 3, bricks=((32, 32, 32), (4, 4, 4)), label="3-D 27-point stencil 64^3, FP32 CSR, point Jacobi, "
                    "16 histories/unknown, tol 1e-8"),
}

FORMS = {
    # weight cutoff 0.1: 8 iterations of 8.5 steps per history at 2^28
    # histories per iteration (1e-2 takes 6 of 16.5); larger cutoffs run
    # faster only as the walks shrink towards a one-step stochastic Jacobi
    # sweep (0.6: 2.1 steps, 21 iterations), so the MC correction stays a
    # multi-term Neumann estimate here
    "chosen": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified",
               "Estimator Type": "Collision", "Weight Cutoff": 1e-1, "History Length": 1000,
               # alias tables padded to the row stride: one 16-byte record per step
               "Transition Sampler": "Alias Strided", "Monte Carlo Type": "Adjoint",
               # LDS tally tile partial sums in FP64 (gfx950's LDS FP64 atomic
               # add, issued without waiting for a return): measured 24.7 ms per
               # 256^3 walk launch against 30.7 ms for FP32 partial sums with a
               # TwoSum compensation, which needs each LDS add's returned value
               "Tally Tile Precision": "FP64"},
    "plain": {"Walk Precision": "FP64", "Source Sampling Type": "Random",
              "Estimator Type": "Collision", "Weight Cutoff": 1e-4, "History Length": 1000,
              "Transition Sampler": "CDF", "Monte Carlo Type": "Adjoint", "Tally Tile Precision": "FP64"},
    # the forward method with the chosen form's walk (no per-step tally atomics)
    "forward": {"Walk Precision": "FP32", "Source Sampling Type": "Stratified",
                "Estimator Type": "Collision", "Weight Cutoff": 1e-1, "History Length": 1000,
                # the strided records: 0.151 s against 0.202 s for the
                # carried-header ones on the 256^3 workload (19.5 against 26.7
                # ms per walk launch)
                "Transition Sampler": "Alias Strided", "Monte Carlo Type": "Forward"},
}


def make_problem(w):
    p0, p1, p2, p3 = w["p"]
    spec = capi.make_spec(w["kind"], w["nx"], seed=w["seed"], p0=p0, p1=p1, p2=p2, p3=p3)
    return capi.build_problem(spec, want64=not w["f32"], want32=w["f32"], nthreads=16)


class DeviceVectors:
    """b and the solution vector resident in HBM, so that a timed solve
    starts with its inputs on the device and leaves its output there."""

    def __init__(self, b):
        self.b = capi.DeviceArray(len(b))
        self.b.upload(b)
        self.x = capi.DeviceArray(len(b))

    def reset(self):
        self.x.zero()


def solve_once(ctx, dev):
    """One MCSA solve from x = 0 on device-resident b and x; the timed
    region is the library call (its own stream is synchronised inside)."""
    dev.reset()
    t0 = time.perf_counter()
    ctx.mcsa_solve_device(dev.b.ptr.value, dev.x.ptr.value)
    dt = time.perf_counter() - t0
    return dt, ctx.stats()


def solve_once_host(ctx, b):
    """The same solve with b and x in host memory (PCIe copies included)."""
    t0 = time.perf_counter()
    ctx.mcsa_solve(b)
    ctx.synchronize()
    return time.perf_counter() - t0


def bytes_per_step(prec, z, sampler, forward=False):
    """Algorithmic bytes of one walk step (one tally + one transition), the
    per-step figure of the measurement plan, whatever the record layout:

    CDF (and the learned CDF rows): 8 (row_ptr[i], row_ptr[i+1]) + p z (CDF
        row) + 4 (col) + p (W_i, or the entry's weight factor) + 2p (tally
        read-modify-write)
    Alias (all alias forms): the CDF row replaced by 8 (one slot)
    Forward: the tally read-modify-write replaced by 8 (b[state])
    p = 4 (FP32 walk) or 8 (FP64 walk); z = mean domain row length.
    """
    p = 4 if prec == "FP32" else 8
    pick = 8 if "Alias" in sampler else p * z
    return 8 + pick + 4 + p + (8 if forward else 2 * p)


# the learned proposal drawn by alias records: the same per-step cost as the
# chosen form's plain alias walk
LEARNED_SAMPLER = "Learned Alias"


def learned_stage(ctx, prob, w, pilot_histories=1 << 24, reps=4, train_rows=1 << 18, epochs=20):
    """The chosen form's learned proposal on this instance, with plain
    almost-optimal sampling as the fallback: train the network on the
    solver's own adjoint tally of this operator (features and targets of a
    row sample), then run `reps` pilot MC solves of the first MCSA source
    with each candidate - plain (the chosen alias walk), the diagonal
    importance g = |a_jj| and the network - and keep the one with the lowest
    variance x walk time per solve (the cost of a fixed error).  Leaves the
    context configured with it; returns the record for the bench line."""
    from mcls_amd import learned

    t0 = time.perf_counter()
    nh = w["histories"]
    rhs = prob.b * ctx.get_left_prec_diag()
    ctx.set_params({"Number of Histories": min(pilot_histories, nh)})
    X, T = learned.gather(ctx, rhs, solve_index=1 << 40)
    rng = np.random.default_rng(1)
    rows = rng.choice(prob.n, size=min(prob.n, train_rows), replace=False)
    net = capi.Mlp(0, seed=1)
    loss = learned.train(net, X[rows], T[rows], epochs=epochs, batch_rows=512, lr=2e-3, seed=1)
    cands = {"plain": None, "diagonal": "diagonal", "network": net}
    var, ms = {}, {}
    for name, g in cands.items():
        if g is None:
            ctx.set_params({"Transition Sampler": FORMS["chosen"]["Transition Sampler"]})
        else:
            ctx.set_params({"Transition Sampler": LEARNED_SAMPLER})
            if g == "diagonal":
                ctx.set_importance_diagonal()
            else:
                ctx.set_importance_model(net)
        samples, t = [], 0.0
        for k in range(reps):
            samples.append(ctx.mc_solve_raw(rhs, (1 << 41) + k))
            t += ctx.stats().ms_walk
        var[name] = learned.relative_variance(np.stack(samples))
        ms[name] = t / reps
        if g is net:
            ctx.set_importance_model(None)
        del samples
    eff = {k: var[k] * ms[k] for k in var}
    choice = min(eff, key=eff.get)
    if choice == "plain":
        ctx.set_params({"Transition Sampler": FORMS["chosen"]["Transition Sampler"]})
    else:
        ctx.set_params({"Transition Sampler": LEARNED_SAMPLER})
        if choice == "diagonal":
            ctx.set_importance_diagonal()
        else:
            ctx.set_importance_model(net)
    ctx.set_params({"Number of Histories": nh})
    return {"choice": choice, "relative_variance": var, "walk_ms_per_solve": ms,
            "variance_x_ms": eff, "pilot_histories": min(pilot_histories, nh), "pilot_solves": reps,
            "train_rows": int(len(rows)), "train_epochs": epochs, "train_loss": float(loss),
            "seconds": time.perf_counter() - t0}, net


def measured_traffic(workload, form, steps_per_launch, tally_tile=None, state_order=None):
    """HBM bytes per launch of a form's walk kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload
    (calibrated against a copy kernel), when its transitions per launch match
    this run's to 2% and it was taken with the same sampler, LDS tally tile
    (size and partial-sum precision) and state order; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"walk_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("workload") != workload or form not in rec.get("forms", {}):
        return None
    rec = rec["forms"][form]
    if abs(rec["steps_per_launch"] - steps_per_launch) > 0.02 * steps_per_launch:
        return None
    if rec.get("tally_tile") != tally_tile or rec.get("state_order") != state_order:
        return None
    if tally_tile and rec.get("tally_tile_precision") != FORMS[form].get("Tally Tile Precision", "FP64"):
        return None
    if rec.get("sampler") != FORMS[form]["Transition Sampler"]:
        return None
    return float(rec["traffic_bytes_per_launch"])


# The attribution ladder: from the plain form to the chosen one a change at
# a time (each line autotuned, same operator, numbering and tolerance).
LADDER = [
    # the plain estimator with its random births walked in birth-state order
    # (same histories, streams and paths; only the order in which they are
    # walked changes): what locality alone buys each kernel
    ("plain + births walked in state order (FP64 CDF, random, cutoff 1e-4)",
     {"Walk Precision": "FP64", "Transition Sampler": "CDF", "Source Sampling Type": "Random",
      "Weight Cutoff": 1e-4, "Tally Tile Precision": "FP64", "Birth Order": "State"}),
    ("chosen kernel at the plain estimator, births walked in state order",
     {"Walk Precision": "FP32", "Transition Sampler": "Alias Strided", "Source Sampling Type": "Random",
      "Weight Cutoff": 1e-4, "Tally Tile Precision": "FP64", "Birth Order": "State"}),
    ("plain estimator -> chosen estimator (stratified, cutoff 1e-1), FP64 CDF walk",
     {"Walk Precision": "FP64", "Transition Sampler": "CDF", "Source Sampling Type": "Stratified",
      "Weight Cutoff": 1e-1, "Tally Tile Precision": "FP64"}),
    ("+ FP32 walk (CDF)",
     {"Walk Precision": "FP32", "Transition Sampler": "CDF", "Source Sampling Type": "Stratified",
      "Weight Cutoff": 1e-1, "Tally Tile Precision": "FP64"}),
    ("chosen kernel at the plain estimator (random, cutoff 1e-4)",
     {"Walk Precision": "FP32", "Transition Sampler": "Alias Strided", "Source Sampling Type": "Random",
      "Weight Cutoff": 1e-4, "Tally Tile Precision": "FP64"}),
]


def error_bound(prob, x, dinv):
    """FP64 error of x against the exact solution without solving: with
    H = I - D^-1 A and e = x* - x, D^-1 (b - A x) = (I - H) e, so
    ||e||_inf <= ||D^-1 r||_inf / (1 - ||H||_inf) whenever ||H||_inf < 1
    (the 27-point workload's rows of |H| sum to s_i <= 0.9).  r in FP64 on
    the host from the FP32 matrix values widened exactly."""
    import scipy.sparse as sp

    val = prob.val32 if prob.val32 is not None else prob.values()
    A = sp.csr_matrix((val.astype(np.float64), prob.col, prob.rp), shape=(prob.n, prob.n))
    r = prob.b - A @ x
    absA = abs(A)
    rowsum = np.asarray(absA.sum(axis=1)).ravel()
    diag = np.abs(A.diagonal())
    hnorm = float(np.max((rowsum - diag) / diag))
    del absA, A
    bound = float(np.max(np.abs(dinv * r))) / (1.0 - hnorm)
    xinf = float(np.max(np.abs(x)))
    return {"bound_inf_abs": bound, "bound_inf_rel": bound / xinf, "norm_H_inf": hnorm,
            "rel_residual_2": float(np.linalg.norm(dinv * r) / np.linalg.norm(dinv * prob.b)),
            "method": "||x - x*||_inf <= ||D^-1 (b - A x)||_inf / (1 - ||I - D^-1 A||_inf), "
                      "residual in FP64 on the host"}


def cpu_baseline(prob, w, plain_iterations=None, nh_sample=1 << 24):
    """The oracle's restatement of the MCLS algorithm in reference semantics
    (oracle/cpu_baseline.py: FP64 CDF lower_bound, one Xorshift stream per
    set, LIFO emission, sets summed / numSets) on the same operator, one set
    per host thread, with a bounded history sample: one MCSA iteration of
    nh_sample histories.  Histories/s from its MC transport time; the
    per-iteration time and the time to tolerance at the workload's history
    count are extrapolated from it."""
    from oracle import cpu_baseline as cb

    nh_sample = min(nh_sample, w["histories"])
    r = cb.run(prob, w["f32"], nh_sample, iters=1, workload_histories=w["histories"])
    iters = plain_iterations or 7
    per_iter = r["extrapolated"]["seconds_mc_per_iteration"]
    return {"value": r["histories_per_s"], "unit": "histories/s", "cores": r["cores"], "kind": "port",
            "host": r["host"],
            "sample": f"1 MCSA iteration of the MCLS-semantics CPU restatement (FP64, CDF lower_bound, "
                      f"{r['sets']} sets x {r['histories_per_set']} histories, weight cutoff 1e-4) on the same "
                      f"operator ({prob.n} unknowns); MC transport time only",
            "seconds_mc": r["seconds_mc"],
            "steps_per_history": r["steps_per_history"],
            "seconds_per_iteration_extrapolated": per_iter,
            "time_to_tol_extrapolated_s": iters * per_iter,
            "extrapolation": f"{w['histories']} histories per iteration x {iters} iterations "
                             f"(the plain form's count), MC transport only"}


def residual_probe(w, prob, calls=10):
    """r = PL (b - A x), point-Jacobi PL, on the workload's operator in the
    bench's state order: `calls` calls per residual format (host b and x, so
    the boundary's copies and gathers run too; a kernel trace separates the
    residual kernels).  Returns {format: host ms per call}."""
    ctx = capi.Context(0)
    if w.get("bricks"):
        ctx.set_state_order(capi.nested_brick_order(w["nx"], w["nx"], w["nx"], w["bricks"]))
    ctx.set_problem(prob, "f32" if w["f32"] else "f64")
    ctx.set_point_jacobi()
    x = np.random.default_rng(1).standard_normal(prob.n)
    out = {}
    for fmt in ("Sliced ELL", "CSR"):
        ctx.set_params({"Residual Format": fmt})
        ctx.prec_residual(prob.b, x)
        t0 = time.perf_counter()
        for _ in range(calls):
            _, nrm = ctx.prec_residual(prob.b, x)
        out[fmt] = 1e3 * (time.perf_counter() - t0) / calls
        log(f"residual probe: {fmt}: {calls} calls, {out[fmt]:.2f} ms per call with host vectors, |r| = {nrm:.6e}")
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 reference solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--form", default="chosen", choices=sorted(FORMS))
    ap.add_argument("--no-tune", action="store_true", help="skip the walk autotuner (default launch shape)")
    ap.add_argument("--no-forward", action="store_true", help="skip the forward-method solve")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-proposal selection of the chosen form (plain sampling)")
    ap.add_argument("--order", default="bricks", choices=["bricks", "natural"],
                    help="state order of the context: the grid's unknowns in nested bricks (the workload's "
                    "brick sizes) or the generator's x-fastest numbering")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy kernel (HBM counter calibration)")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="a solver parameter over the form's own (profiling variants; the line's config "
                    "records it and no committed traffic record matches it)")
    ap.add_argument("--residual-probe", action="store_true",
                    help="only run the preconditioned residual 10 times in each residual format on the "
                    "workload's operator in the bench's state order (residual kernel profiles)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 3 x 1 GiB read + 1 GiB write (elementwise copy kernel)")
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(set replication over RCCL is a design note only); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    log(f"generating {args.workload}")
    prob = make_problem(w)
    log(f"n={prob.n} nnz={prob.nnz}")
    if args.residual_probe:
        residual_probe(w, prob)
        return
    ctx = capi.Context(0)
    # The context numbers the unknowns brick by brick (P A P^T; b and x stay
    # in the generator's numbering at the API): the walks born close together
    # then share rows in L2 and LDS tile slots.  Same operator, same
    # histories, same tolerance.
    order_label = "natural"
    if args.order == "bricks" and w.get("bricks"):
        lv = w["bricks"]
        ctx.set_state_order(capi.nested_brick_order(w["nx"], w["nx"], w["nx"], lv))
        order_label = " in ".join(f"{b[0]}^3 bricks" for b in reversed(lv))
        log(f"state order: {order_label}")
    ctx.set_problem(prob, "f32" if w["f32"] else "f64")
    ctx.set_point_jacobi()
    base = {"Convergence Tolerance": 1e-8, "Maximum Iterations": 60, "Number of Histories": w["histories"],
            "Random Seed": 2024}
    ctx.set_params(base)
    overrides = {}
    for kv in args.set:
        k, _, v = kv.partition("=")
        try:
            overrides[k] = int(v)
        except ValueError:
            try:
                overrides[k] = float(v)
            except ValueError:
                overrides[k] = v
    FORMS[args.form].update(overrides)
    ctx.set_params(FORMS[args.form])
    setup = {}
    t0 = time.perf_counter()
    ctx.build_domain()
    ctx.synchronize()
    st = ctx.stats()
    setup["domain_and_tables_s"] = time.perf_counter() - t0
    log(f"domain built: nnz={st.domain_nnz} in {time.perf_counter() - t0:.2f}s (kernel {st.ms_domain:.1f} ms)")
    z = st.domain_nnz / prob.n
    learned_rec, net = None, None
    if args.form == "chosen" and not args.no_learned:
        learned_rec, net = learned_stage(ctx, prob, w)
        setup["learned_stage_s"] = learned_rec["seconds"]
        log(f"learned proposal: {learned_rec['choice']} (relative variance "
            f"{ {k: round(v, 5) for k, v in learned_rec['relative_variance'].items()} }, walk ms "
            f"{ {k: round(v, 2) for k, v in learned_rec['walk_ms_per_solve'].items()} })")
    sampler = LEARNED_SAMPLER if learned_rec and learned_rec["choice"] != "plain" else FORMS[args.form]["Transition Sampler"]
    tuning = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(prob.b)
        setup["autotune_s"] = time.perf_counter() - t0
        log(f"autotune: {tuning}")
    setup["total_s"] = sum(setup.values())
    ctx_tile = tuning["tally_tile"] if tuning else 0  # "Tally Tile" defaults to 0

    dev = DeviceVectors(prob.b)
    for i in range(args.warmup):
        dt, st = solve_once(ctx, dev)
        log(f"warmup {i}: {dt:.3f}s iters={st.iterations} tol={st.achieved_tol:.2e}")
    times, hist, steps, iters, walk_ms, launches, tols = [], 0, 0, [], 0.0, 0, []
    for i in range(args.steps):
        dt, st = solve_once(ctx, dev)
        times.append(dt)
        hist += st.histories
        steps += st.steps
        iters.append(st.iterations)
        walk_ms += st.ms_walk
        launches += st.walk_launches
        tols.append(st.achieved_tol)
        log(f"step {i}: {dt:.3f}s iters={st.iterations} tol={st.achieved_tol:.2e} "
            f"walk={st.ms_walk:.1f}ms src={st.ms_source:.1f}ms res={st.ms_residual:.1f}ms")
    # the same solve from host arrays, for the PCIe-inclusive time (not the value)
    host_io_s = solve_once_host(ctx, prob.b)
    log(f"with host b and x (PCIe copies included): {host_io_s:.3f}s")
    total = sum(times)
    hps = hist / total
    prec = FORMS[args.form]["Walk Precision"]
    bstep = bytes_per_step(prec, z, sampler, FORMS[args.form].get("Monte Carlo Type") == "Forward")
    avg_launch_s = walk_ms / 1e3 / max(launches, 1)
    steps_per_launch = steps / max(launches, 1)
    achieved = bstep * steps_per_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": hps,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == "FP32" else "fp64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['label']}", "form": args.form, "n": prob.n,
                   "nnz": prob.nnz, "histories_per_iteration": w["histories"],
                   "state_order": order_label, "parallelism": "1 GPU",
                   **({"overrides": overrides} if overrides else {})},
        "time_to_tol_s": total / args.steps,
        "time_to_tol_host_io_s": host_io_s,
        "mcsa_iterations": iters,
        "achieved_tol": max(tols),
        "steps_per_history": steps / max(hist, 1),
        "walk_steps_per_s": steps / (walk_ms / 1e3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if overrides else measured_traffic(args.workload, args.form,
                                                                        steps_per_launch, ctx_tile, order_label),
                     # not counted in this run (counters need their own profiler
                     # passes): the committed rocprofv3 FETCH_SIZE + WRITE_SIZE record
                     # of this form, calibrated on a copy kernel, when it matches
                     "traffic_from": f"profiles/walk_{args.workload}_traffic.json",
                     "kernel": f"k_walk<{'float' if prec == 'FP32' else 'double'}>",
                     "bytes_per_step": bstep, "steps_per_launch": steps_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3},
        "tuning": tuning,
        "learned_proposal": learned_rec,
        "setup_s": setup,
    }
    # the FP64 error of the chosen solution, bounded without a solve
    if args.workload.startswith("cfg4"):
        dev.reset()
        ctx.mcsa_solve_device(dev.b.ptr.value, dev.x.ptr.value)
        out["error_vs_fp64"] = error_bound(prob, dev.x.download(), ctx.get_left_prec_diag())
        log(f"error bound: {out['error_vs_fp64']['bound_inf_rel']:.2e} (relative, inf-norm)")
    if not args.no_forward and args.form == "chosen":
        # the forward method with the chosen form's walk (reported beside
        # the headline, which stays the adjoint solve)
        ctx.set_params(FORMS["forward"])
        ftune = None
        if not args.no_tune:
            ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64, "Tally Tile": 0})
            ftune = ctx.autotune(prob.b)
        solve_once(ctx, dev)  # forward domain build and warm-up
        fdt, fst = solve_once(ctx, dev)
        log(f"forward FP32: {fdt:.3f}s iters={fst.iterations} tol={fst.achieved_tol:.2e}")
        fbytes = bytes_per_step("FP32", z, FORMS["forward"]["Transition Sampler"], True)
        fsteps_per_launch = fst.steps / max(fst.walk_launches, 1)
        out["forward"] = {"time_to_tol_s": fdt, "mcsa_iterations": fst.iterations,
                          "traffic": measured_traffic(args.workload, "forward", fsteps_per_launch,
                                                      0, order_label),
                          "achieved_tol": fst.achieved_tol, "histories_per_s": fst.histories / fdt,
                          "walk_steps_per_s": fst.steps / (fst.ms_walk / 1e3),
                          "roofline_frac": fbytes * fst.steps / (fst.ms_walk / 1e3) / 1e9 / HBM_PEAK_GBS,
                          "bytes_per_step": fbytes, "steps_per_launch": fsteps_per_launch,
                          "avg_launch_ms": fst.ms_walk / max(fst.walk_launches, 1),
                          "tuning": ftune}
    if not args.no_plain and args.form == "chosen":
        # the plain form in the same context (same numbering of the unknowns)
        ctx.set_params(FORMS["plain"])
        ptune = None
        if not args.no_tune:
            ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64, "Tally Tile": 0})
            ptune = ctx.autotune(prob.b)
            log(f"autotune (plain): {ptune}")
        dt, st = solve_once(ctx, dev)
        log(f"plain FP64: {dt:.3f}s iters={st.iterations} tol={st.achieved_tol:.2e}")
        pbytes = bytes_per_step("FP64", z, FORMS["plain"]["Transition Sampler"])
        psteps_per_launch = st.steps / max(st.walk_launches, 1)
        out["plain"] = {"time_to_tol_s": dt, "mcsa_iterations": st.iterations,
                        "achieved_tol": st.achieved_tol, "histories_per_s": st.histories / dt,
                        "walk_steps_per_s": st.steps / (st.ms_walk / 1e3),
                        "roofline_frac": pbytes * st.steps / (st.ms_walk / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_step": pbytes, "steps_per_launch": psteps_per_launch,
                        "avg_launch_ms": st.ms_walk / max(st.walk_launches, 1),
                        "traffic": measured_traffic(args.workload, "plain", psteps_per_launch,
                                                    (ptune or {}).get("tally_tile", 0), order_label),
                        "tuning": ptune}
        out["speedup_vs_plain"] = dt / out["time_to_tol_s"]
        if "forward" in out:
            out["forward"]["speedup_vs_plain"] = dt / out["forward"]["time_to_tol_s"]
        # the attribution ladder: what each change buys on its own
        ladder = []
        for label, params in LADDER:
            ctx.set_params(dict(FORMS["chosen"], **dict({"Birth Order": "Id"}, **params)))
            if not args.no_tune:
                ctx.set_params({"Walk Block Size": 256, "Walk Blocks Per CU": 0, "Walk Batch": 64, "Tally Tile": 0})
                ctx.autotune(prob.b)
            solve_once(ctx, dev)
            ldt, lst = solve_once(ctx, dev)
            ladder.append({"form": label, "time_to_tol_s": ldt, "mcsa_iterations": lst.iterations,
                           "achieved_tol": lst.achieved_tol, "steps_per_history": lst.steps / max(lst.histories, 1),
                           "speedup_vs_plain": dt / ldt})
            log(f"ladder: {label}: {ldt:.3f}s iters={lst.iterations}")
        ctx.set_params({"Birth Order": "Id"})
        out["attribution"] = {"plain_s": dt, "chosen_s": out["time_to_tol_s"], "ladder": ladder}
        out["plain_same_estimator"] = next(r for r in ladder if r["form"].startswith("plain estimator -> chosen"))
    dev.b.close()
    dev.x.close()
    ctx.close()
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(prob, w, out.get("plain", {}).get("mcsa_iterations"))
        log(f"cpu baseline {out['cpu_baseline']['value']:.3e} histories/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
