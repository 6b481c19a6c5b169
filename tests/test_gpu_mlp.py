"""The surrogate MLP's GPU kernels against the FP64 numpy restatement (tests/mlp_oracle.py).

Bars: the BF16-MFMA training gradient agrees with the FP64 gradient to 2e-2 per layer
(norm-wise; BF16 operands carry 8 mantissa bits), the FP32 inference output and input Jacobian
to 1e-5 relative; Adam training on a smooth function reduces the loss by 20x.
"""
import numpy as np
import pytest

import mlp_oracle as mo
from ceos_amd import gpu, hiprt

pytestmark = pytest.mark.gpu

DIMS = (8, 256, 6)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, DIMS[0])).astype(np.float32)
    T = rng.normal(size=(n, DIMS[2])).astype(np.float32)
    return X, T


def _dev(a):
    return hiprt.DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float32))


@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_training_gradient_matches_fp64(n):
    m = gpu.Mlp(*DIMS, max_batch=4096, seed=3)
    p = m.params()
    X, T = _data(n, n)
    g, loss = m.gradients(_dev(X), _dev(T), n)
    lo, go = mo.loss_and_grad(p.astype(np.float64), X, T, *DIMS)
    print("loss", loss, lo, flush=True)
    assert abs(loss - lo) < 2e-2 * lo
    lay, _ = mo.layout(*DIMS)
    for l, (o, ob, din, dout) in enumerate(lay):
        for a, b, what in ((o, ob, "W"), (ob, ob + dout, "b")):
            e = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            print(l, what, e, flush=True)
            assert e < 2e-2


def test_fp32_inference_and_jacobian():
    n = 777
    m = gpu.Mlp(*DIMS, max_batch=1024, seed=5)
    p = m.params().astype(np.float64)
    X, _ = _data(9, n)
    dX = _dev(X)
    dY = hiprt.DeviceArray(n * DIMS[2], np.float32)
    dJ = hiprt.DeviceArray(n * DIMS[2] * DIMS[0], np.float32)
    m.forward(dX, n, dY, dJ)
    Y = dY.numpy().reshape(n, DIMS[2])
    J = dJ.numpy().reshape(n, DIMS[2], DIMS[0])
    Yo, _, _ = mo.forward(p, X, *DIMS)
    Jo = mo.jacobian(p, X, *DIMS)
    assert np.linalg.norm(Y - Yo) / np.linalg.norm(Yo) < 1e-5
    assert np.linalg.norm(J - Jo) / np.linalg.norm(Jo) < 1e-5


def test_adam_training_fits_a_smooth_map():
    rng = np.random.default_rng(0)
    n = 4096
    X = rng.uniform(-1, 1, size=(n, DIMS[0])).astype(np.float32)
    A = rng.normal(size=(DIMS[0], DIMS[2])) / 3
    T = (np.tanh(X @ A) + 0.3 * (X[:, :DIMS[2]] ** 2)).astype(np.float32)
    m = gpu.Mlp(*DIMS, max_batch=n, seed=1)
    dX, dT = _dev(X), _dev(T)
    l0 = m.train_step(dX, dT, n, 1e-3, want_loss=True)
    for _ in range(300):
        m.train_step(dX, dT, n, 1e-3)
    l1 = m.train_step(dX, dT, n, 1e-3, want_loss=True)
    print("loss", l0, l1, flush=True)
    assert l1 < l0 / 20
    # the trained parameters reproduce the loss in FP64
    lo, _ = mo.loss_and_grad(m.params().astype(np.float64), X, T, *DIMS)
    assert lo < l0 / 20


def test_weighted_gradient_matches_fp64():
    n = 1000
    m = gpu.Mlp(*DIMS, max_batch=n, seed=4)
    p = m.params()
    X, T = _data(17, n)
    W = np.random.default_rng(18).uniform(0.0, 5.0, size=T.shape).astype(np.float32)
    g, loss = m.gradients(_dev(X), _dev(T), n, dW=_dev(W))
    lo, go = mo.loss_and_grad(p.astype(np.float64), X, T, *DIMS, W=W)
    print("weighted loss", loss, lo, flush=True)
    assert abs(loss - lo) < 2e-2 * lo
    e = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert e < 2e-2
    # a weight of zero removes a sample: same gradient as the batch without it
    W2 = np.ones_like(W)
    W2[::2] = 0.0
    g2, _ = m.gradients(_dev(X), _dev(T), n, dW=_dev(W2))
    _, go2 = mo.loss_and_grad(p.astype(np.float64), X[1::2], T[1::2], *DIMS)
    assert np.linalg.norm(2 * g2 - go2) / np.linalg.norm(go2) < 2e-2


def test_fp32_training_mode_gradient_matches_fp64():
    """The reference training mode (FP32 GEMMs) reproduces the FP64 gradient to FP32 rounding."""
    n = 512
    m = gpu.Mlp(*DIMS, max_batch=n, seed=6)
    m.set_training_fp32(True)
    p = m.params()
    X, T = _data(23, n)
    g, loss = m.gradients(_dev(X), _dev(T), n)
    lo, go = mo.loss_and_grad(p.astype(np.float64), X, T, *DIMS)
    assert abs(loss - lo) < 1e-5 * lo
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4
