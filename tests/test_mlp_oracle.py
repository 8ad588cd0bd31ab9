"""The importance network's CPU oracle pinned by finite differences and
closed forms (no GPU)."""
import numpy as np

from oracle import mlp_oracle as mo


def _data(rows=40, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(rows, mo.N_IN)), rng.normal(size=rows)


def test_exact_gradient_matches_finite_differences():
    X, T = _data()
    p = mo.init_params(1)
    p += np.random.default_rng(2).normal(size=p.shape) * 0.05  # non-zero biases
    loss, g = mo.loss_grad(p, X, T, emulate=False)
    rng = np.random.default_rng(3)
    idx = rng.choice(mo.N_PARAMS, 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mo.loss_grad(p + e, X, T, emulate=False)
        lm, _ = mo.loss_grad(p - e, X, T, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_gradient_close_to_exact():
    X, T = _data(200, 4)
    p = mo.init_params(5)
    _, ge = mo.loss_grad(p, X, T, emulate=False)
    _, gq = mo.loss_grad(p, X, T, emulate=True)
    assert np.linalg.norm(gq - ge) <= 0.05 * np.linalg.norm(ge)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -3.140625, 65504.0], np.float32)
    y = mo.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0 + 2**-7)
    assert y[3] == np.float32(1.0 + 2**-6)
    assert y[4] == np.float32(-3.140625)
    assert mo.bf16(np.float32(0.1)) == np.float32(0.10009765625)


def test_adam_matches_closed_form_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = mo.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first bias-corrected Adam step moves each parameter by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)


def test_training_reduces_loss_on_a_learnable_target():
    rng = np.random.default_rng(7)
    X = rng.normal(size=(512, mo.N_IN))
    T = np.tanh(X[:, 0]) + 0.5 * X[:, 1] ** 2
    p = mo.init_params(8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0, _ = mo.loss_grad(p, X, T, emulate=False)
    for t in range(1, 301):
        _, g = mo.loss_grad(p, X, T, emulate=False)
        p, m, v = mo.adam(p, g, m, v, t, lr=3e-3)
    l1, _ = mo.loss_grad(p, X, T, emulate=False)
    assert l1 < 0.2 * l0
