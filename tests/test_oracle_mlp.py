"""Pins the FP64 MLP oracle (oracle/mlp64.py) that checks the surrogate's device kernels.

Its gradient is compared with central finite differences of its own loss (which a dropped term,
a transposed weight matrix or a wrong activation derivative fails, the dimensions being all
different), and its outputs and gradient at all-zero parameters with their closed forms.
"""
import numpy as np

from oracle import mlp64

DIMS = (3, 5, 2, 4)  # n_in, width, depth, n_out


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    n = mlp64.n_params(*DIMS)
    p = rng.normal(0, 0.7, n)
    x = rng.normal(size=(7, DIMS[0]))
    t = rng.random((7, DIMS[3]))
    mean, scale = rng.normal(size=DIMS[0]), rng.uniform(0.5, 2.0, DIMS[0])
    loss, g = mlp64.loss_grad(p, x, t, *DIMS, mean=mean, scale=scale)
    h = 1e-6
    fd = np.zeros(n)
    for k in range(n):
        e = np.zeros(n)
        e[k] = h
        lp, _ = mlp64.loss_grad(p + e, x, t, *DIMS, mean=mean, scale=scale)
        lm, _ = mlp64.loss_grad(p - e, x, t, *DIMS, mean=mean, scale=scale)
        fd[k] = (lp - lm) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-9), np.max(np.abs(g - fd))
    assert loss > 0


def test_zero_parameters_closed_form():
    rng = np.random.default_rng(2)
    n = mlp64.n_params(*DIMS)
    x = rng.normal(size=(6, DIMS[0]))
    t = rng.random((6, DIMS[3]))
    acts = mlp64.forward(np.zeros(n), x, *DIMS)
    assert np.all(acts[-1] == 0.5)  # sigmoid(0)
    loss, g = mlp64.loss_grad(np.zeros(n), x, t, *DIMS)
    assert np.isclose(loss, np.mean((0.5 - t) ** 2))
    # only the output bias has a gradient: sum_b 2 (0.5 - t) * 0.25 / count
    want = np.zeros(n)
    want[-DIMS[3]:] = np.sum(2.0 * (0.5 - t) * 0.25, axis=0) / t.size
    assert np.allclose(g, want, atol=1e-15)


def test_layout_size():
    n_in, width, depth, n_out = 12, 256, 3, 64
    assert mlp64.n_params(n_in, width, depth, n_out) == (12 * 256 + 256) + 2 * (256 * 256 + 256) + (256 * 64 + 64)
