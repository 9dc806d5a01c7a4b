"""CPU restatement of the closure network -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may import
this module, as the checker of the HIP kernels in libheom_amd/csrc/mlp.hip.

The network has no counterpart in the reference (which truncates the
hierarchy at max_depth, heom.h:77-91); its definition is the mathematics:

    x' = (x - x_shift) / x_scale
    h1 = tanh(W1 x' + b1),  h2 = tanh(W2 h1 + b2),  y' = W3 h2 + b3
    y  = y' * y_scale + y_shift
    loss = mean over rows and outputs of (y' - (t - y_shift) / y_scale)^2

with the parameters flattened as W1 (n_hid x n_in, row-major), b1, W2, b2,
W3 (n_out x n_hid), b3, and Adam (Kingma & Ba 2015) with bias correction.
Everything here is float64.  Pinned by: central finite differences of the
loss (tests/test_closure_cpu.py), which a transposed weight, a missing
activation derivative or a wrong normalisation fails.
"""
import numpy as np


def unpack(p, n_in, n_hid, n_out):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in [(n_hid, n_in), (n_hid,), (n_hid, n_hid), (n_hid,), (n_out, n_hid), (n_out,)]:
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, n_in, n_hid, n_out, X, scaling=None):
    """Raw rows in, raw rows out."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_hid, n_out)
    xm, xs, ym, ys = _scaling(scaling, n_in, n_out)
    x = (np.asarray(X, np.float64) - xm) / xs
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return (h2 @ W3.T + b3) * ys + ym


def loss_grad(p, n_in, n_hid, n_out, X, T, scaling=None):
    """Mean squared error of the normalised outputs and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_hid, n_out)
    xm, xs, ym, ys = _scaling(scaling, n_in, n_out)
    x = (np.asarray(X, np.float64) - xm) / xs
    t = (np.asarray(T, np.float64) - ym) / ys
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    d = y - t
    loss = np.mean(d * d)
    dy = 2.0 * d / d.size
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (1.0 - h2 * h2)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (1.0 - h1 * h1)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def adam(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def column_scaling(X):
    """Per-column mean and standard deviation (scale 1 for a constant column)."""
    X = np.asarray(X, np.float64)
    mu = X.mean(0)
    sd = np.sqrt(((X - mu) ** 2).mean(0))
    return mu, np.where(sd > 0.0, sd, 1.0)


def _scaling(s, n_in, n_out):
    if s is None:
        return np.zeros(n_in), np.ones(n_in), np.zeros(n_out), np.ones(n_out)
    return [np.asarray(a, np.float64) for a in s]
