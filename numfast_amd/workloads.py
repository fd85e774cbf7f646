"""Seeded synthetic inputs for the relational hot path, generated on the GPU.

Shapes follow the benchmark configurations of the reference path
(group-by + unique, stable argsort, batched group-by); the generators run
with torch on ``cuda:0`` so 2**28-row inputs are ready in milliseconds and
stay resident in HBM.  Seed 0 unless stated.
"""

import math

import torch


def _gen(seed, device="cuda"):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def zipf_ids(n, ids, s, g, device="cuda"):
    """n int32 keys: bounded Zipf(s) over ranks 1..ids (inverse CDF), mapped
    to ids through a seeded permutation of [0, ids)."""
    ranks = torch.arange(1, ids + 1, device=device, dtype=torch.float64)
    cdf = torch.cumsum(ranks.pow(-s), 0)
    cdf /= cdf[-1].clone()
    u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    r = torch.searchsorted(cdf, u, right=True).clamp_(max=ids - 1)
    perm = torch.randperm(ids, generator=g, device=device).to(torch.int32)
    return perm[r]


def groupby_config(n=1 << 28, ids=1 << 24, dist="zipf", s=1.1, null_frac=0.05, seed=0,
                   device="cuda"):
    """High-cardinality group-by + unique: keys, int32 values in
    [-2**15, 2**15), a value validity mask and a separate key validity mask
    (for unique), both Bernoulli(1 - null_frac)."""
    g = _gen(seed, device)
    if dist == "zipf":
        keys = zipf_ids(n, ids, s, g, device)
    else:
        keys = torch.randint(0, ids, (n,), generator=g, device=device, dtype=torch.int32)
    vals = torch.randint(-(1 << 15), 1 << 15, (n,), generator=g, device=device,
                         dtype=torch.int32)
    valid = (torch.rand(n, generator=g, device=device) >= null_frac).to(torch.uint8)
    key_valid = (torch.rand(n, generator=g, device=device) >= null_frac).to(torch.uint8)
    return keys, vals, valid, key_valid


def sort_config(n=1 << 28, dist="uniform", seed=0, device="cuda"):
    """Stable argsort inputs: 'uniform' over the full int32 range, 'swaps'
    (sorted, then 10% of positions permuted among themselves) or
    'distinct1024' (1024 distinct uniform int32 values)."""
    g = _gen(seed, device)
    full = lambda m: torch.randint(-(1 << 31), 1 << 31, (m,), generator=g, device=device,
                                   dtype=torch.int64).to(torch.int32)
    if dist == "uniform":
        return full(n)
    if dist == "swaps":
        k = torch.sort(full(n)).values
        m = n // 10
        idx = torch.randperm(n, generator=g, device=device)[:m]
        k[idx] = k[idx[torch.randperm(m, generator=g, device=device)]]
        return k
    if dist == "distinct1024":
        pool = full(1024)
        return pool[torch.randint(0, 1024, (n,), generator=g, device=device)]
    raise ValueError(dist)


def batch_config(nq=4096, total=1 << 28, seed=0, device="cuda"):
    """Batched group-by inputs on the device: nq queries, row counts
    log-uniform in [1e3, 1e7] rescaled to `total` (factor returned, counts
    floored at 1e3), key cardinality log-uniform in [10, 1e6], keys drawn
    from a bounded Zipf(s) by the continuous inverse CDF, s uniform in
    [0, 1.5] (s = 0: uniform), ranks scattered over [0, card) by an affine
    map with a prime multiplier above 1e6; int32 values uniform in
    [-2^15, 2^15).  Returns keys, vals, row_off (int64, nq + 1), card
    (int32), scale."""
    g = _gen(seed, device)
    u = lambda m: torch.rand(m, generator=g, device=device, dtype=torch.float64)
    lo, hi = math.log(1e3), math.log(1e7)
    n = torch.exp(lo + (hi - lo) * u(nq))
    scale = float(total / n.sum().item())
    n = torch.clamp(torch.floor(n * scale), min=1e3).to(torch.int64)
    card = torch.exp(math.log(10) + (math.log(1e6) - math.log(10)) * u(nq)).to(torch.int64)
    s = 1.5 * u(nq)
    row_off = torch.zeros(nq + 1, dtype=torch.int64, device=device)
    row_off[1:] = torch.cumsum(n, 0)
    N = int(row_off[-1].item())
    qid = torch.repeat_interleave(torch.arange(nq, device=device), n)
    c = card[qid].to(torch.float64)
    sq = s[qid]
    x = u(N)
    one = (sq - 1.0).abs() < 1e-9
    e = torch.where(one, torch.ones_like(sq), 1.0 - sq)
    # inverse CDF of the density r^-s on [1, c + 1)
    r = torch.where(one, torch.exp(x * torch.log(c + 1.0)),
                    torch.pow(1.0 + x * (torch.pow(c + 1.0, e) - 1.0), 1.0 / e))
    rank = torch.clamp(torch.floor(r) - 1.0, min=0.0)
    rank = torch.minimum(rank, c - 1.0).to(torch.int64)
    keys = ((rank * 1000003 + qid * 7919) % card[qid]).to(torch.int32)
    vals = torch.randint(-(1 << 15), 1 << 15, (N,), generator=g, device=device,
                         dtype=torch.int32)
    return keys, vals, row_off, card.to(torch.int32), scale
