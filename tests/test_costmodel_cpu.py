"""Pins the CPU restatement of the cost model: backpropagation against central
finite differences, and a fit that drives the loss down on a learnable target."""
import numpy as np

from oracle import costmodel_ref as R


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, n = 13, 40
    x = rng.normal(size=(n, d))
    t = rng.normal(size=n)
    p = rng.normal(scale=0.3, size=R.n_params(d))
    _, g = R.grad(p, x, t)
    h = 1e-6
    idx = rng.choice(len(p), 120, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (R.loss(p + e, x, t) - R.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_fit_learns_a_smooth_target():
    rng = np.random.default_rng(1)
    d, n = 4, 64
    x = rng.uniform(-1, 1, size=(n, d))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - x[:, 2] * x[:, 3]
    p0 = rng.normal(scale=0.3, size=R.n_params(d))
    l0 = R.loss(p0, x, t)
    p1 = R.adam_fit(p0, x, t, steps=400, lr=3e-3)
    assert R.loss(p1, x, t) < 0.1 * l0
