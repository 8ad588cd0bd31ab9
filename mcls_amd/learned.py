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


def compare(ctx, rhs, reps=8, first_index=1000, net=None, diagonal=True):
    """Relative per-solve variance of plain almost-optimal sampling against
    the learned proposal (network ``net``) and the diagonal importance
    g = |a_jj|, all on the same source and solve numbers.  Leaves the
    context on the plain sampler with no network attached."""
    out = {}
    ctx.set_params({"Transition Sampler": "CDF"})
    plain = solve_samples(ctx, rhs, reps, first_index)
    out["plain"] = relative_variance(plain)
    ctx.set_params({"Transition Sampler": "Learned"})
    if net is not None:
        ctx.set_importance_model(net)
        ctx.set_importance_mask(None)
        learned = solve_samples(ctx, rhs, reps, first_index)
        out["learned"] = relative_variance(learned)
        ctx.set_importance_model(None)
    if diagonal:
        ctx.set_importance_diagonal()
        out["diagonal"] = relative_variance(solve_samples(ctx, rhs, reps, first_index))
    ctx.set_params({"Transition Sampler": "CDF"})
    return out


def select_proposal(ctx, rhs, net=None, reps=4, first_index=2000):
    """Instance-level fallback: measure the pilot variance of plain sampling,
    the diagonal importance and (when given) the network on ``rhs``, then
    leave the context configured with the lowest.  Returns (choice,
    variances).  Switching g state by state (``variance_mask``) makes g
    discontinuous, which raises variance on diffusion operators; choosing
    per instance keeps each candidate's g whole."""
    v = compare(ctx, rhs, reps=reps, first_index=first_index, net=net, diagonal=True)
    choice = min(v, key=v.get)
    if choice == "plain":
        ctx.set_params({"Transition Sampler": "CDF"})
    else:
        ctx.set_params({"Transition Sampler": "Learned"})
        if choice == "diagonal":
            ctx.set_importance_diagonal()
        else:
            ctx.set_importance_model(net)
    return choice, v


__all__ = ["gather", "train", "solve_samples", "relative_variance", "variance_mask", "compare",
           "select_proposal"]
