"""Parity at the benchmark sizes (2^28 rows), checked on the device.

The C oracle is too slow at these sizes, so these tests check the library's
device-pointer calls against torch's own routines on the same GPU -- a
stable sort (torch.sort(stable=True)), torch.unique, index_add_ and
scatter_reduce -- and against size-independent properties: a permutation
whose keys are non-decreasing with ties in row order; counts that add up
to the valid rows; sums that add up to the valid values; min <= mean <=
max; uniq[inv] == keys on valid rows and -1 on NULL rows.

Inputs are the benchmark's own seeded generators: 2^28 int32 keys Zipf(1.1)
over 2^24 ids (and the uniform variant) with values in [-2^15, 2^15) and
5% NULL values / keys; 2^28 int32 sort keys uniform over int32, sorted with
10% of positions permuted, and 1024 distinct values.
"""

import ctypes

import pytest

pytestmark = pytest.mark.gpu

N = 1 << 28
IDS = 1 << 24


@pytest.fixture(scope="module")
def torch_mod(hip):
    import torch
    return torch


def _groupby_reference(torch, keys, vals, mask):
    """Groups of the rows with mask set: ukeys, counts, int64 sums, min, max
    (torch.unique + index_add_ + scatter_reduce)."""
    kk = keys[mask]
    vv = vals[mask].to(torch.int64)
    u, inv, cnt = torch.unique(kk, sorted=True, return_inverse=True, return_counts=True)
    del kk
    ng = u.numel()
    sums = torch.zeros(ng, dtype=torch.int64, device="cuda").index_add_(0, inv, vv)
    mn = torch.full((ng,), 2 ** 31 - 1, dtype=torch.int64, device="cuda")
    mn.scatter_reduce_(0, inv, vv, reduce="amin")
    mx = torch.full((ng,), -(2 ** 31), dtype=torch.int64, device="cuda")
    mx.scatter_reduce_(0, inv, vv, reduce="amax")
    return u.to(torch.int64), cnt.to(torch.int64), sums, mn.to(torch.int32), mx.to(torch.int32)


@pytest.mark.parametrize("dist", ["zipf", "uniform"])
def test_groupby_and_unique_2p28(hip, torch_mod, dist):
    torch = torch_mod
    from numfast_amd import workloads as W
    print("group-by + unique, 2^28 rows, %s keys over 2^24 ids" % dist, flush=True)
    keys, vals, valid, kvalid = W.groupby_config(n=N, ids=IDS, dist=dist)
    outs = [torch.empty(IDS, dtype=dt, device="cuda") for dt in
            (torch.int64, torch.int64, torch.int64, torch.int32, torch.int32, torch.float64)]
    ng = hip.nfh_group_agg_i32(keys.data_ptr(), vals.data_ptr(), valid.data_ptr(), None, N,
                               *[o.data_ptr() for o in outs], IDS)
    torch.cuda.synchronize()
    assert ng > 0
    uk, cnt, sm, mn, mx, mean = [o[:ng] for o in outs]
    m = valid.bool()
    ru, rc, rs, rmn, rmx = _groupby_reference(torch, keys, vals, m)
    assert ng == ru.numel()
    assert torch.equal(uk, ru) and torch.equal(cnt, rc) and torch.equal(sm, rs)
    assert torch.equal(mn, rmn) and torch.equal(mx, rmx)
    # size-independent properties
    assert bool((uk[1:] > uk[:-1]).all())
    assert int(cnt.sum()) == int(m.sum())
    assert int(sm.sum()) == int(vals[m].to(torch.int64).sum())
    assert torch.equal(mean, sm.to(torch.float64) / cnt.to(torch.float64))
    assert bool((mn.to(torch.float64) <= mean).all() and (mean <= mx.to(torch.float64)).all())
    del uk, cnt, sm, mn, mx, mean, ru, rc, rs, rmn, rmx, outs
    # unique over the key mask
    uniq = torch.empty(IDS, dtype=torch.int32, device="cuda")
    inv = torch.empty(N, dtype=torch.int32, device="cuda")
    nu = hip.nfh_unique_valid_i32(keys.data_ptr(), kvalid.data_ptr(), N, uniq.data_ptr(),
                                  inv.data_ptr())
    torch.cuda.synchronize()
    km = kvalid.bool()
    ref_u = torch.unique(keys[km], sorted=True)
    assert nu == ref_u.numel() and torch.equal(uniq[:nu], ref_u)
    assert bool((inv[~km] == -1).all())
    iv = inv[km].to(torch.int64)
    assert bool((iv >= 0).all() and (iv < nu).all())
    assert torch.equal(uniq[:nu][iv], keys[km])


def _check_perm(torch, keys, perm):
    n = keys.numel()
    ref = torch.sort(keys, stable=True).indices.to(torch.int32)
    assert torch.equal(perm, ref)
    # the properties, independently of torch's sort
    sk = keys[perm.to(torch.int64)]
    assert bool((sk[1:] >= sk[:-1]).all())
    tie = sk[1:] == sk[:-1]
    assert bool((perm[1:][tie] > perm[:-1][tie]).all())
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_fill_(0, perm.to(torch.int64), 1)
    assert int(seen.sum()) == n


@pytest.mark.parametrize("dist", ["uniform", "swaps", "distinct1024"])
def test_argsort_and_learned_sort_2p28(hip, torch_mod, dist):
    torch = torch_mod
    from numfast_amd import native
    from numfast_amd import workloads as W
    print("argsort, 2^28 int32 keys, %s" % dist, flush=True)
    keys = W.sort_config(N, dist)
    perm = torch.empty(N, dtype=torch.int32, device="cuda")
    t0, t1, tp = (torch.empty(N, dtype=torch.int32, device="cuda") for _ in range(3))
    rc = hip.nf_sort_perm_i32(keys.data_ptr(), N, 0, perm.data_ptr(), t0.data_ptr(),
                              t1.data_ptr(), tp.data_ptr())
    torch.cuda.synchronize()
    assert rc == 0
    _check_perm(torch, keys, perm)
    # the learned-CDF sort gives the same permutation
    L = native.lib("learned")
    model = native.CdfModel()
    assert L.nfh_cdf_train_i32(keys.data_ptr(), N, 100, 0, ctypes.byref(model), None) == 0
    perm2 = torch.empty_like(perm)
    assert L.nfh_sort_perm_learned_i32(keys.data_ptr(), N, ctypes.byref(model),
                                       perm2.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(perm2, perm)
    # descending: stable on the inverted word, ties in row order
    rc = hip.nf_sort_perm_i32(keys.data_ptr(), N, 1, perm.data_ptr(), t0.data_ptr(),
                              t1.data_ptr(), tp.data_ptr())
    torch.cuda.synchronize()
    assert rc == 0
    ref = torch.sort(keys, stable=True, descending=True).indices.to(torch.int32)
    assert torch.equal(perm, ref)


def test_filter_derive_sort_limit_2p28(hip, torch_mod):
    torch = torch_mod
    from numfast_amd import native
    from numfast_amd import workloads as W
    print("filter -> derive -> sort -> limit, 2^28 rows", flush=True)
    L = native.lib("learned")
    v = W.sort_config(N, "uniform")
    lim = 1000
    w_out = torch.empty(lim, dtype=torch.int32, device="cuda")
    r_out = torch.empty(lim, dtype=torch.int32, device="cuda")
    got = L.nfh_filter_derive_sort_limit_i32(v.data_ptr(), N, 3, lim, None, w_out.data_ptr(),
                                             r_out.data_ptr())
    torch.cuda.synchronize()
    assert got == lim
    rows = torch.nonzero(v > 0).squeeze(1)
    w = v[rows].to(torch.int64) * 3
    w = ((w + 2 ** 31) % 2 ** 32 - 2 ** 31).to(torch.int32)  # int32 wrap
    order = torch.sort(w, stable=True).indices[:lim]
    assert torch.equal(w_out, w[order]) and torch.equal(r_out, rows[order].to(torch.int32))
