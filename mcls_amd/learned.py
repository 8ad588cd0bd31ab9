"""Training and evaluation of the learned transition proposal.

The network (include/mcls_mlp.h) maps per-state features of (A, source) to an
importance g; the walk then samples P_ij = |h_ij| g_j / S_i with weight
factor S_i / g_j (unbiased for any g > 0).  Everything here drives the
C-ABI: features, targets (from the solver's own adjoint tallies), training
(the network's own forward/backward/Adam kernels) and the variance
comparison against plain almost-optimal sampling, which is the fallback
wherever g raises variance.  The reference has no counterpart: its
transitions are always P_ij = |h_ij| / W_i
(src/MCLS_AlmostOptimalDomain_impl.hpp:342-363).
"""
import os

import numpy as np


def gather(ctx, rhs, solve_index, target_source=None):
    """Features and training targets for the system loaded in ``ctx``.

    ``rhs`` is the MC source the network sees at inference (the preconditioned
    right-hand side or residual).  The targets come from one plain adjoint MC
    solve of ``target_source`` (default |rhs|, which keeps the tally positive):
    t_j = log |a_jj x_j| normalised and clipped to [-4, 4]."""
    X = ctx.importance_features(rhs)
    src = np.abs(rhs) if target_source is None else target_source
    T = ctx.importance_targets(src, solve_index)
    return X, T


def train(net, X, T, epochs=30, batch_rows=512, lr=2e-3, seed=0):
    """Replace the network's dataset with (X, T) and train; returns the final
    epoch's mean loss."""
    net.dataset_clear()
    net.dataset_append(X, T)
    return net.train(epochs=epochs, batch_rows=batch_rows, lr=lr, seed=seed)


def solve_samples(ctx, rhs, reps, first_index):
    """``reps`` independent MC solves of rhs (solve numbers first_index ...)
    with the context's current sampler; returns a (reps, n) array."""
    return np.stack([ctx.mc_solve_raw(rhs, first_index + k) for k in range(reps)])


def relative_variance(samples):
    """Per-solve variance of the estimator relative to its mean's squared
    norm: sum_j var_j / sum_j mean_j^2 (the estimators are unbiased, so this
    is the expected squared relative error of one solve)."""
    mean = samples.mean(axis=0)
    var = samples.var(axis=0, ddof=1)
    return float(var.sum() / max(float((mean * mean).sum()), 1e-300))


def variance_mask(plain_samples, learned_samples):
    """Per-state switch of the importance: 1 where the learned proposal's
    sample variance is at most the plain one's, else 0 (plain fallback)."""
    vp = plain_samples.var(axis=0, ddof=1)
    vl = learned_samples.var(axis=0, ddof=1)
    return (vl <= vp).astype(np.int8)


def compare(ctx, rhs, reps=8, first_index=1000, net=None, diagonal=True, samplers=("CDF", "Learned")):
    """Relative per-solve variance of plain almost-optimal sampling against
    the learned proposal (network ``net``) and the diagonal importance
    g = |a_jj|, all on the same source and solve numbers.  ``samplers``
    names the plain and the importance-reshaped "Transition Sampler" (by CDF
    inversion, or "Alias" / "Learned Alias" for FP32 walks).  Leaves the
    context on the plain sampler with no network attached."""
    out = {}
    ctx.set_params({"Transition Sampler": samplers[0]})
    plain = solve_samples(ctx, rhs, reps, first_index)
    out["plain"] = relative_variance(plain)
    ctx.set_params({"Transition Sampler": samplers[1]})
    if net is not None:
        ctx.set_importance_model(net)
        ctx.set_importance_mask(None)
        learned = solve_samples(ctx, rhs, reps, first_index)
        out["learned"] = relative_variance(learned)
        ctx.set_importance_model(None)
    if diagonal:
        ctx.set_importance_diagonal()
        out["diagonal"] = relative_variance(solve_samples(ctx, rhs, reps, first_index))
    ctx.set_params({"Transition Sampler": samplers[0]})
    return out


def select_proposal(ctx, rhs, net=None, reps=4, first_index=2000, samplers=("CDF", "Learned")):
    """Instance-level fallback: measure the pilot variance of plain sampling,
    the diagonal importance and (when given) the network on ``rhs``, then
    leave the context configured with the lowest.  Returns (choice,
    variances).  Switching g state by state (``variance_mask``) makes g
    discontinuous, which raises variance on diffusion operators; choosing
    per instance keeps each candidate's g whole."""
    v = compare(ctx, rhs, reps=reps, first_index=first_index, net=net, diagonal=True, samplers=samplers)
    choice = min(v, key=v.get)
    if choice == "plain":
        ctx.set_params({"Transition Sampler": samplers[0]})
    else:
        ctx.set_params({"Transition Sampler": samplers[1]})
        if choice == "diagonal":
            ctx.set_importance_diagonal()
        else:
            ctx.set_importance_model(net)
    return choice, v


# The learned-proposal family of the workloads: 2-D variable-coefficient
# diffusion, kappa = exp(Gaussian random field), correlation length 0.1,
# standard deviation 1 (512 random Fourier features), b ~ U[0, 1).
def family_instance(seed, n=256):
    from mcls_amd import capi

    return capi.build_problem(capi.make_spec(capi.VARCOEF2D_5PT, n, n, seed=seed, p0=0.1, p1=1.0, p2=512))


# Weights (include/mcls_mlp.h layout, FP32) of the network trained on 2,000
# instances of the family at 256 x 256 (seeds 100000-101999, 1,024 sampled
# states each, targets from 8 histories per unknown, 30 epochs).
FAMILY_NETWORK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "importance_varcoef2d_256.npy")


def load_network(path=FAMILY_NETWORK, device=0):
    """An importance network with the weights stored at ``path``."""
    from mcls_amd import capi

    net = capi.Mlp(device, seed=1)
    net.set_params(np.load(path))
    return net


STUDY_PARAMS = {"Walk Precision": "FP32", "Source Sampling Type": "Stratified", "Estimator Type": "Collision",
                "Weight Cutoff": 1e-4, "History Length": 20000, "Random Seed": 7}
STUDY_SAMPLERS = ("Alias", "Learned Alias")


def _load(ctx, prob, per_unknown):
    ctx.set_problem(prob)
    ctx.set_point_jacobi()
    ctx.set_params(dict(STUDY_PARAMS, **{"Number of Histories": per_unknown * prob.n,
                                         "Transition Sampler": STUDY_SAMPLERS[0]}))
    return prob.b * ctx.get_left_prec_diag()


def _prefetch(seeds, n):
    """Yield (seed, problem) for ``seeds``, building the next instance on a
    host thread while the caller's GPU work on the current one runs."""
    from concurrent.futures import ThreadPoolExecutor

    seeds = list(seeds)
    if not seeds:
        return
    with ThreadPoolExecutor(1) as pool:
        fut = pool.submit(family_instance, seeds[0], n)
        for k, s in enumerate(seeds):
            prob = fut.result()
            if k + 1 < len(seeds):
                fut = pool.submit(family_instance, seeds[k + 1], n)
            yield s, prob


def family_train(n_train, n=256, rows_per_instance=1024, per_unknown=8, epochs=30, batch_rows=4096, lr=2e-3,
                 seed0=100000, log=print):
    """Train a network on ``n_train`` instances of the family: per instance,
    features and targets (one plain MC solve's tally at ``per_unknown``
    histories per unknown) of ``rows_per_instance`` sampled states.  Returns
    (network, record)."""
    import time

    from mcls_amd import capi

    t0 = time.perf_counter()
    net = capi.Mlp(0, seed=1)
    Xs, Ts = [], []
    with capi.Context(0) as ctx:
        for k, (s, prob) in enumerate(_prefetch(range(seed0, seed0 + n_train), n)):
            rhs = _load(ctx, prob, per_unknown)
            X, T = gather(ctx, rhs, solve_index=k)
            rows = np.random.default_rng(k).choice(prob.n, size=min(prob.n, rows_per_instance), replace=False)
            Xs.append(X[rows])
            Ts.append(T[rows])
            if k % 100 == 0:
                log(f"train instance {k}: {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    loss = train(net, np.concatenate(Xs), np.concatenate(Ts), epochs=epochs, batch_rows=batch_rows, lr=lr, seed=1)
    rec = {"n": n, "train_instances": n_train, "train_seeds": [seed0, seed0 + n_train - 1],
           "rows_per_instance": rows_per_instance, "histories_per_unknown": per_unknown, "epochs": epochs,
           "batch_rows": batch_rows, "lr": lr, "train_loss": float(loss), "gather_seconds": t1 - t0,
           "train_seconds": time.perf_counter() - t1, "params": STUDY_PARAMS}
    log(f"trained on {len(Xs) * rows_per_instance} rows: loss {loss:.4f} in {rec['train_seconds']:.1f}s")
    return net, rec


def family_eval(net, n_eval, n=256, per_unknown=4, reps=4, pilot_reps=3, seed0=200000, log=print):
    """Evaluate ``net`` on ``n_eval`` held-out instances: per instance, the
    relative variance of one MC solve with plain almost-optimal sampling, the
    diagonal importance and the network (``reps`` solves each on the same
    streams), and the instance-level selection (``pilot_reps`` pilot solves on
    other streams; it keeps the lowest pilot variance).  Returns a record with
    the per-instance rows and a summary."""
    import time

    from mcls_amd import capi

    t0 = time.perf_counter()
    rows_out = []
    with capi.Context(0) as ctx:
        for h, (s, prob) in enumerate(_prefetch(range(seed0, seed0 + n_eval), n)):
            rhs = _load(ctx, prob, per_unknown)
            v = compare(ctx, rhs, reps=reps, first_index=1000, net=net, diagonal=True, samplers=STUDY_SAMPLERS)
            choice, pilot = select_proposal(ctx, rhs, net=net, reps=pilot_reps, first_index=5000,
                                            samplers=STUDY_SAMPLERS)
            ctx.set_importance_model(None)
            rows_out.append({"seed": s, "variance": v, "selected": choice, "pilot": pilot})
            if h % 20 == 0:
                log(f"eval instance {h}: {v} -> {choice} ({time.perf_counter() - t0:.1f}s)")
    rec = {"n": n, "eval_instances": n_eval, "eval_seeds": [seed0, seed0 + n_eval - 1],
           "histories_per_unknown": per_unknown, "reps": reps, "pilot_reps": pilot_reps, "eval": rows_out,
           "seconds": time.perf_counter() - t0}
    if rows_out:
        q = [0.0, 0.1, 0.5, 0.9, 1.0]
        names = ["min", "p10", "median", "p90", "max"]
        r_net = np.array([r["variance"]["plain"] / r["variance"]["learned"] for r in rows_out])
        r_diag = np.array([r["variance"]["plain"] / r["variance"]["diagonal"] for r in rows_out])
        rec["summary"] = {
            "variance_reduction_network": dict(zip(names, np.quantile(r_net, q).tolist())),
            "variance_reduction_diagonal": dict(zip(names, np.quantile(r_diag, q).tolist())),
            "network_below_plain_fraction": float(np.mean(r_net > 1.0)),
            "network_beats_diagonal_fraction": float(np.mean(r_net > r_diag)),
            "selected": {c: int(sum(r["selected"] == c for r in rows_out)) for c in ("plain", "diagonal", "learned")},
            "selection_kept_worse_than_plain": int(sum(r["pilot"][r["selected"]] > r["pilot"]["plain"]
                                                      for r in rows_out)),
        }
    return rec


def _direct_error(prob, x):
    """Relative 2-norm error of x against an FP64 direct solve of A x = b."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    A = sp.csr_matrix((prob.values(), prob.col, prob.rp), shape=(prob.n, prob.n))
    xd = sla.spsolve(A.tocsc(), prob.b)
    return float(np.linalg.norm(x - xd) / np.linalg.norm(xd))


def family_mcsa(net, seeds, n=256, per_unknown=32, iters=30, relaxation=0.7, smoother_steps=10, chunk=20,
                log=print):
    """MCSA (damped Richardson smoothing, ||r|| <= 1e-8 ||PL b||) on the
    family instances ``seeds`` with plain sampling, the diagonal importance
    and the network, up to ``iters`` iterations.  The solve runs in calls of
    ``chunk`` iterations, each starting from the previous one's x: MCSA
    carries nothing else between iterations, the MC streams follow the
    context's solve counter and the test is relative to ||PL b||, so this is
    the same iteration sequence.  A diverging solve ends in a non-finite
    residual and is recorded with its residual norms."""
    import time

    from mcls_amd import capi

    out = []
    with capi.Context(0) as ctx:
        for s, prob in _prefetch(seeds, n):
            res = {"seed": s}
            for name in ("plain", "diagonal", "network"):
                _load(ctx, prob, per_unknown)
                if name != "plain":
                    ctx.set_params({"Transition Sampler": STUDY_SAMPLERS[1]})
                    if name == "diagonal":
                        ctx.set_importance_diagonal()
                    else:
                        ctx.set_importance_model(net)
                x = np.zeros(prob.n)
                hist, done, t0 = [], 0, time.perf_counter()
                rec = {}
                while done < iters:
                    ctx.set_params({"Convergence Tolerance": 1e-8, "Maximum Iterations": min(chunk, iters - done),
                                    "Richardson Relaxation": relaxation, "Smoother Steps": smoother_steps})
                    try:
                        x = ctx.mcsa_solve(prob.b, x)
                    except capi.GpuError as e:  # a diverging solve overflows
                        h = [float(v) for v in ctx.residual_history()]
                        hist += h if not hist else h[1:]
                        rec = {"error": str(e)}
                        break
                    st = ctx.stats()
                    h = [float(v) for v in ctx.residual_history()]
                    hist += h if not hist else h[1:]  # a call's first norm is the previous call's last
                    done += int(st.iterations)
                    rec = {"iterations": done, "converged": bool(st.converged), "achieved_tol": hist[-1]}
                    log(f"mcsa seed {s} {name}: {done} iterations, relative residual {hist[-1]:.3e}")
                    if st.converged or int(st.iterations) == 0:
                        break
                rec["seconds"] = time.perf_counter() - t0
                if rec.get("converged"):
                    rec["error_vs_direct"] = _direct_error(prob, x)
                rec["residual_history"] = hist
                res[name] = rec
                ctx.set_importance_model(None)
                log(f"mcsa seed {s} {name}: {({k: v for k, v in rec.items() if k != 'residual_history'})}")
            out.append(res)
    return {"n": n, "histories_per_unknown": per_unknown, "max_iterations": iters, "relaxation": relaxation,
            "smoother_steps": smoother_steps, "chunk": chunk, "instances": out}


__all__ = ["gather", "train", "solve_samples", "relative_variance", "variance_mask", "compare",
           "select_proposal", "family_instance", "load_network", "family_train", "family_eval", "family_mcsa"]


if __name__ == "__main__":
    import argparse
    import json
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mcls_amd import capi

    ap = argparse.ArgumentParser(description="train and evaluate the learned proposal on its workload family")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--train", type=int, default=2000, help="training instances (0: load --net)")
    ap.add_argument("--eval", type=int, default=200, help="held-out instances")
    ap.add_argument("--rows", type=int, default=1024, help="sampled states per training instance")
    ap.add_argument("--train-per-unknown", type=int, default=8, help="histories per unknown of the target solves")
    ap.add_argument("--eval-per-unknown", type=int, default=4, help="histories per unknown of the compared solves")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--reps", type=int, default=4)
    ap.add_argument("--mcsa", type=int, default=0, help="held-out instances solved by MCSA with each proposal")
    ap.add_argument("--mcsa-per-unknown", type=int, default=32)
    ap.add_argument("--mcsa-iters", type=int, default=30)
    ap.add_argument("--net", default=None, help="network weights (.npy): written after training, read when --train 0")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    rec = {}
    if a.train:
        net, rec["train"] = family_train(a.train, n=a.n, rows_per_instance=a.rows, per_unknown=a.train_per_unknown,
                                         epochs=a.epochs, log=log)
        if a.net:
            np.save(a.net, net.params())
    else:
        net = load_network(a.net or FAMILY_NETWORK)
        rec["network"] = os.path.basename(a.net or FAMILY_NETWORK)
    if a.eval:
        rec["eval"] = family_eval(net, a.eval, n=a.n, per_unknown=a.eval_per_unknown, reps=a.reps, log=log)
    if a.mcsa:
        rec["mcsa"] = family_mcsa(net, range(200000, 200000 + a.mcsa), n=a.n, per_unknown=a.mcsa_per_unknown,
                                  iters=a.mcsa_iters, log=log)
    if a.out:
        with open(a.out, "w") as f:
            f.write(json.dumps(rec))
    print(json.dumps({k: v.get("summary", {k2: v2 for k2, v2 in v.items() if k2 != "instances"})
                      if isinstance(v, dict) else v for k, v in rec.items()}))
