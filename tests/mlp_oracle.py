"""FP64 numpy restatement of the surrogate MLP (test infrastructure, checker only).

Same network and parameter layout as the library's ceos_mlp: n_in -> width -> width -> width ->
n_out, SiLU after the hidden layers, flat parameters [W1 b1 W2 b2 W3 b3 W4 b4] with W row-major
[out][in]; loss = mean over the batch and outputs of (Y - T)^2.
"""
import numpy as np


def layout(n_in, width, n_out):
    dims = [(n_in, width), (width, width), (width, width), (width, n_out)]
    out, off = [], 0
    for din, dout in dims:
        out.append((off, off + din * dout, din, dout))
        off += din * dout + dout
    return out, off


def unpack(p, n_in, width, n_out):
    lay, _ = layout(n_in, width, n_out)
    p = np.asarray(p, dtype=np.float64)
    return [(p[o:o + din * dout].reshape(dout, din), p[o + din * dout:o + din * dout + dout]) for o, _, din, dout in lay]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, n_in, width, n_out):
    Ws = unpack(p, n_in, width, n_out)
    a = np.asarray(X, dtype=np.float64)
    Zs, As = [], [a]
    for W, b in Ws[:3]:
        z = a @ W.T + b
        a = silu(z)
        Zs.append(z)
        As.append(a)
    W, b = Ws[3]
    return a @ W.T + b, Zs, As


def loss_and_grad(p, X, T, n_in, width, n_out, W=None):
    """Loss sum W (Y - T)^2 / (batch n_out) (W = 1 when None) and its parameter gradient."""
    Y, Zs, As = forward(p, X, n_in, width, n_out)
    Ws = unpack(p, n_in, width, n_out)
    n = Y.size
    d = Y - T
    w = np.ones_like(d) if W is None else np.asarray(W, dtype=np.float64)
    loss = np.sum(w * d * d) / n
    g = []
    dz = 2.0 * w * d / n
    for l in range(3, -1, -1):
        W, _ = Ws[l]
        g.append((dz.T @ As[l], dz.sum(axis=0)))
        if l > 0:
            dz = (dz @ W) * dsilu(Zs[l - 1])
    g = g[::-1]
    return loss, np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in g])


def jacobian(p, X, n_in, width, n_out):
    """dY/dX per sample: [n][n_out][n_in]."""
    _, Zs, _ = forward(p, X, n_in, width, n_out)
    Ws = unpack(p, n_in, width, n_out)
    J = np.einsum("on,bn->bon", Ws[3][0], dsilu(Zs[2]))
    for l in (2, 1):
        J = np.einsum("bon,ni->boi", J, Ws[l][0]) * dsilu(Zs[l - 1])[:, None, :]
    return np.einsum("bon,ni->boi", J, Ws[0][0])
