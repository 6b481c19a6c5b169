"""Pins of the FP64 MLP restatement: parameter gradient and input Jacobian against central
differences, and the SiLU derivative against its definition."""
import numpy as np

import mlp_oracle as mo


def _setup(seed=0, n_in=5, width=16, n_out=3, n=7):
    rng = np.random.default_rng(seed)
    _, npar = mo.layout(n_in, width, n_out)
    p = rng.normal(scale=0.4, size=npar)
    X = rng.normal(size=(n, n_in))
    T = rng.normal(size=(n, n_out))
    return p, X, T, (n_in, width, n_out)


def test_silu_derivative():
    z = np.linspace(-6, 6, 41)
    h = 1e-6
    assert np.allclose(mo.dsilu(z), (mo.silu(z + h) - mo.silu(z - h)) / (2 * h), atol=1e-9)


def test_parameter_gradient_fd():
    p, X, T, dims = _setup()
    _, g = mo.loss_and_grad(p, X, T, *dims)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, X, T, *dims)[0] - mo.loss_and_grad(pm, X, T, *dims)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_input_jacobian_fd():
    p, X, _, dims = _setup(2)
    J = mo.jacobian(p, X, *dims)
    h = 1e-6
    for i in range(dims[0]):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, i] += h
        Xm[:, i] -= h
        fd = (mo.forward(p, Xp, *dims)[0] - mo.forward(p, Xm, *dims)[0]) / (2 * h)
        assert np.allclose(J[:, :, i], fd, atol=1e-8)


def test_weighted_parameter_gradient_fd():
    p, X, T, dims = _setup(3)
    W = np.random.default_rng(4).uniform(0.0, 3.0, size=T.shape)
    loss, g = mo.loss_and_grad(p, X, T, *dims, W=W)
    Y = mo.forward(p, X, *dims)[0]
    assert np.isclose(loss, np.sum(W * (Y - T) ** 2) / Y.size)
    rng = np.random.default_rng(5)
    h = 1e-6
    for i in rng.choice(p.size, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, X, T, *dims, W=W)[0] - mo.loss_and_grad(pm, X, T, *dims, W=W)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
