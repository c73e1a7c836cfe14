"""The float64 MLP reference (learned CG guess) against finite differences and Adam's closed form (CPU)."""

import numpy as np

from oracle import mlp


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    params = mlp.init([7, 9, 9, 3], seed=2)
    params = [(W, rng.normal(scale=0.3, size=b.shape)) for W, b in params]
    x, t = rng.normal(size=(11, 7)), rng.normal(size=(11, 3))
    _, grads = mlp.loss_grad(params, x, t)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            flat, gflat = arr.reshape(-1), g.reshape(-1)
            for k in rng.choice(flat.size, size=min(6, flat.size), replace=False):
                old = flat[k]
                flat[k] = old + h
                lp = mlp.loss_grad(params, x, t)[0]
                flat[k] = old - h
                lm = mlp.loss_grad(params, x, t)[0]
                flat[k] = old
                assert abs((lp - lm) / (2 * h) - gflat[k]) < 1e-7 * max(1.0, abs(gflat[k]))


def test_silu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    assert np.allclose((mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h), mlp.dsilu(z), atol=1e-9)


def test_first_adam_step_moves_by_lr_times_sign():
    """At step 1 the bias-corrected update is lr * g / (|g| + eps): every weight moves by ~lr against its gradient."""
    params = mlp.init([4, 5, 2], seed=3)
    grads = [(np.full_like(W, 0.5), np.full_like(b, -2.0)) for W, b in params]
    new, _ = mlp.adam_step(params, grads, mlp.zeros_state(params), 1, lr=1e-3)
    for (W0, b0), (W1, b1) in zip(params, new):
        assert np.allclose(W1 - W0, -1e-3 * 0.5 / (0.5 + 1e-8))
        assert np.allclose(b1 - b0, 1e-3 * 2.0 / (2.0 + 1e-8))
