"""CPU restatement of the learned-CDF network (test infrastructure only).

The network is the one the learned sort trains on the device: input x (the
normalised key word), 32 ReLU units, 32 ReLU units, one linear output; loss
= mean squared error against the CDF targets y.  Parameter layout (1153
floats): w1[32] b1[32] w2[32][32] (row = output unit) b2[32] w3[32] b3.

`forward` / `grad` are written in float64 from the chain rule; they are
pinned by central finite differences in tests/test_oracle_golden.py, and
the device kernels (nfh_mlp_grad) are checked against them.
"""

import numpy as np

H = 32
W1, B1, W2, B2, W3, B3 = 0, 32, 64, 64 + 1024, 64 + 1024 + 32, 64 + 1024 + 64
P = B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1], p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, x):
    """Network output for every x (float64)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)[:, None]
    h1 = np.maximum(x * w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss(p, x, y):
    r = forward(p, x) - np.asarray(y, np.float64)
    return float(np.mean(r * r))


def grad(p, x, y):
    """d loss / d p (float64, layout as p)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    S = x.size
    pre1 = x[:, None] * w1 + b1
    h1 = np.maximum(pre1, 0.0)
    a2 = h1 @ w2.T + b2
    h2 = np.maximum(a2, 0.0)
    out = h2 @ w3 + b3
    dout = 2.0 * (out - y) / S
    g = np.zeros(P)
    g[B3] = dout.sum()
    g[W3:B3] = h2.T @ dout
    dh2 = (dout[:, None] * w3) * (a2 > 0)
    g[B2:W3] = dh2.sum(0)
    g[W2:B2] = (dh2.T @ h1).reshape(-1)
    dh1 = (dh2 @ w2) * (pre1 > 0)
    g[B1:W2] = dh1.sum(0)
    g[W1:B1] = dh1.T @ x
    return g


def finite_diff(p, x, y, idx, eps=1e-6):
    """Central difference of `loss` along parameter idx."""
    p = np.asarray(p, np.float64).copy()
    p[idx] += eps
    a = loss(p, x, y)
    p[idx] -= 2 * eps
    b = loss(p, x, y)
    return (a - b) / (2 * eps)
