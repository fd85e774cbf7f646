"""CPU restatement of the group-by strategy selector (test infrastructure).

features(keys): the 64 numbers the device kernel computes from a strided
sample of up to 4096 keys (56-bin histogram over the sample's [min, max],
then log2(n)/32, log2(span)/32, distinct fraction, largest key share, share
of rows whose key occurs >= 4 times, ascending-neighbour fraction,
log2(distinct)/16, 1).

forward / loss / grad: the 64 -> 64 (ReLU) -> 5 softmax classifier, float64,
parameter layout W1[64][64] (row = hidden unit), b1[64], W2[5][64], b2[5];
pinned by finite differences in tests/test_oracle_golden.py.
"""

import numpy as np

F, H, C = 64, 64, 5
W1, B1 = 0, F * H
W2 = B1 + H
B2 = W2 + C * H
P = B2 + C
SAMPLE, HBINS = 4096, 56


def features(keys):
    keys = np.asarray(keys, np.int32)
    n = keys.size
    S = min(n, SAMPLE)
    idx = (np.arange(S, dtype=np.uint64) * np.uint64(n)) // np.uint64(S)
    s = keys[idx.astype(np.int64)].astype(np.int64)
    mn, mx = int(s.min()), int(s.max())
    span = mx - mn + 1
    hist = np.bincount(((s - mn) * HBINS) // span, minlength=HBINS)[:HBINS]
    _, cnt = np.unique(s, return_counts=True)
    asc = int(np.count_nonzero(s[:-1] <= s[1:]))
    out = np.zeros(F)
    out[:HBINS] = hist / S
    out[HBINS + 0] = np.log2(n) / 32
    out[HBINS + 1] = np.log2(span) / 32
    out[HBINS + 2] = cnt.size / S
    out[HBINS + 3] = cnt.max() / S
    out[HBINS + 4] = cnt[cnt >= 4].sum() / S
    out[HBINS + 5] = asc / (S - 1) if S > 1 else 1.0
    out[HBINS + 6] = np.log2(cnt.size) / 16
    out[HBINS + 7] = 1.0
    return out


def unpack(p):
    p = np.asarray(p, np.float64)
    return p[W1:B1].reshape(H, F), p[B1:W2], p[W2:B2].reshape(C, H), p[B2:P]


def forward(p, X):
    w1, b1, w2, b2 = unpack(p)
    h = np.maximum(np.asarray(X, np.float64) @ w1.T + b1, 0.0)
    return h @ w2.T + b2


def loss(p, X, y):
    z = forward(p, X)
    z = z - z.max(1, keepdims=True)
    lse = np.log(np.exp(z).sum(1))
    return float(np.mean(lse - z[np.arange(len(y)), y]))


def grad(p, X, y):
    w1, b1, w2, b2 = unpack(p)
    X = np.asarray(X, np.float64)
    y = np.asarray(y)
    N = X.shape[0]
    a = X @ w1.T + b1
    h = np.maximum(a, 0.0)
    z = h @ w2.T + b2
    z = z - z.max(1, keepdims=True)
    pr = np.exp(z)
    pr /= pr.sum(1, keepdims=True)
    dz = pr.copy()
    dz[np.arange(N), y] -= 1.0
    dz /= N
    g = np.zeros(P)
    g[W2:B2] = (dz.T @ h).reshape(-1)
    g[B2:P] = dz.sum(0)
    dh = (dz @ w2) * (a > 0)
    g[W1:B1] = (dh.T @ X).reshape(-1)
    g[B1:W2] = dh.sum(0)
    return g


def finite_diff(p, X, y, idx, eps=1e-6):
    p = np.asarray(p, np.float64).copy()
    p[idx] += eps
    a = loss(p, X, y)
    p[idx] -= 2 * eps
    b = loss(p, X, y)
    return (a - b) / (2 * eps)


def loss_cost(p, X, T):
    """Cost model: mean over samples of 1/2 sum_c (z_c - T_c)^2."""
    d = forward(p, X) - np.asarray(T, np.float64)
    return float(0.5 * (d * d).sum() / d.shape[0])


def grad_cost(p, X, T):
    w1, b1, w2, b2 = unpack(p)
    X = np.asarray(X, np.float64)
    N = X.shape[0]
    a = X @ w1.T + b1
    h = np.maximum(a, 0.0)
    dz = (h @ w2.T + b2 - np.asarray(T, np.float64)) / N
    g = np.zeros(P)
    g[W2:B2] = (dz.T @ h).reshape(-1)
    g[B2:P] = dz.sum(0)
    dh = (dz @ w2) * (a > 0)
    g[W1:B1] = (dh.T @ X).reshape(-1)
    g[B1:W2] = dh.sum(0)
    return g


def finite_diff_cost(p, X, T, idx, eps=1e-6):
    p = np.asarray(p, np.float64).copy()
    p[idx] += eps
    a = loss_cost(p, X, T)
    p[idx] -= 2 * eps
    b = loss_cost(p, X, T)
    return (a - b) / (2 * eps)
