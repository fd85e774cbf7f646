"""GPU parity of the batched group-by (one hipGraph-captured persistent
launch, and the plain per-query form) against the oracle's per-query dense
group-by."""

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu
BATCH_UNIT_KEYS = 4096  # keys per unit of a query wider than 4096 keys


def _plan(torch, k, v, row_off, card, cap=None):
    kt = torch.from_numpy(k).cuda()
    vt = torch.from_numpy(v).cuda()
    return native.GroupBatch(torch, kt, vt, row_off, card, cap=cap), kt, vt


def _check(res, ref):
    ng, uk, ct, sm, go = res
    ruk, rct, rsm, rgo = ref
    assert ng == ruk.size
    assert np.array_equal(uk.cpu().numpy(), ruk)
    assert np.array_equal(ct.cpu().numpy(), rct)
    assert np.array_equal(sm.cpu().numpy(), rsm)
    assert np.array_equal(go.cpu().numpy(), rgo)


@pytest.mark.parametrize("nq,total,card_hi", [(8, 5000, 50), (64, 1 << 18, 10 ** 5),
                                              (300, 1 << 20, 10 ** 6)])
def test_batch_parity_both_modes(hip, nq, total, card_hi):
    import torch
    rng = np.random.default_rng(nq)
    k, v, row_off, card = ox.batch_workload(rng, nq, total, card_hi=card_hi, rows_lo=10,
                                            rows_hi=1e5)
    ref = ox.group_batch(k, v, row_off, card)
    b, kt, vt = _plan(torch, k, v, row_off, card)
    _check(b.run(0), ref)
    _check(b.run(1), ref)
    _check(b.run(2), ref)  # each query through the library's plain group-by
    b.close()


def test_graph_replays_with_eager_work_between(hip):
    """The captured graph replayed several times, with eager launches and
    synchronisations in between, equals the eager result every time; new
    data in the same buffers shows up in the next replay."""
    import torch
    rng = np.random.default_rng(1)
    k, v, row_off, card = ox.batch_workload(rng, 40, 1 << 17, card_hi=3000, rows_lo=10,
                                            rows_hi=1e5)
    ref = ox.group_batch(k, v, row_off, card)
    b, kt, vt = _plan(torch, k, v, row_off, card)
    for _ in range(3):
        _check(b.run(0), ref)
        _check(b.run(1), ref)  # eager per-query launches between replays
        torch.cuda.synchronize()
        _ = (kt.float() * 2).sum().item()  # unrelated eager work
    v2 = rng.integers(-(2 ** 15), 2 ** 15, v.size).astype(np.int32)
    vt.copy_(torch.from_numpy(v2))
    torch.cuda.synchronize()
    _check(b.run(0), ox.group_batch(k, v2, row_off, card))
    b.close()


def test_batch_values_outside_packed_range(hip):
    import torch
    rng = np.random.default_rng(2)
    k, v, row_off, card = ox.batch_workload(rng, 20, 50000, card_hi=5000, rows_lo=10,
                                            rows_hi=1e4)
    pick = rng.random(v.size) < 0.1
    v[pick] = rng.integers(-(2 ** 31), 2 ** 31 - 1, int(pick.sum()), dtype=np.int64).astype(
        np.int32)
    v[:5] = [32767, 32768, -32768, -32769, 2 ** 31 - 1]
    ref = ox.group_batch(k, v, row_off, card)
    b, _, _ = _plan(torch, k, v, row_off, card)
    _check(b.run(0), ref)
    _check(b.run(1), ref)
    b.close()


def test_batch_empty_queries_and_codes(hip):
    import torch
    k = np.array([0, 2, 2, 1, 0, 0, 5], np.int32)
    v = np.array([1, 2, 3, 4, 5, 6, 7], np.int32)
    row_off = np.array([0, 3, 3, 4, 7, 7], np.int64)  # queries 1 and 4 have no rows
    card = np.array([3, 4, 2, 6, 0], np.int32)
    ref = ox.group_batch(k, v, row_off, card)
    b, kt, vt = _plan(torch, k, v, row_off, card)
    _check(b.run(0), ref)
    _check(b.run(1), ref)
    _check(b.run(2), ref)
    b.close()
    # a key outside its query's range -> -2
    kb = k.copy()
    kb[3] = 2  # query 2 has card 2
    assert ox.group_batch(kb, v, row_off, card) == -2
    b, _, _ = _plan(torch, kb, v, row_off, card)
    assert b.run(0)[0] == -2 and b.run(1)[0] == -2 and b.run(2)[0] == -2
    b.close()
    # capacity below the group count -> -7
    b, _, _ = _plan(torch, k, v, row_off, card, cap=2)
    assert b.run(0)[0] == -7 and b.run(2)[0] == -7
    assert native.lib("batch").nfh_group_batch_run(b.plan, 3) == -3
    b.close()


def test_batch_plan_rejects_bad_offsets(hip):
    import ctypes
    import torch
    L = native.lib("batch")
    kt = torch.zeros(4, dtype=torch.int32, device="cuda")
    ro = np.array([0, 5, 4], np.int64)  # decreasing
    cd = np.array([3, 3], np.int32)
    out = torch.zeros(4, dtype=torch.int64, device="cuda")
    go = torch.zeros(3, dtype=torch.int64, device="cuda")
    rc = ctypes.c_int32(0)
    p = L.nfh_group_batch_plan(kt.data_ptr(), kt.data_ptr(), ro.ctypes.data, cd.ctypes.data, 2,
                               out.data_ptr(), out.data_ptr(), out.data_ptr(), go.data_ptr(), 4,
                               ctypes.byref(rc))
    assert not p and rc.value == -3


def test_batch_wide_queries_edges(hip):
    """Queries wider than 4096 keys go through the by-key-range path
    (4096-key units): a wide query with no rows, one whose last unit holds
    a single key, values outside the packed range in multi-unit queries, a
    wide query's key outside its range (-2) and a capacity below the group
    count (-7); both modes against the oracle."""
    import torch
    rng = np.random.default_rng(5)
    cards = [4097, 4096 * 12 + 1, 70000, 5000, 100, 250000]
    rows = [3000, 40000, 0, 9000, 500, 120000]
    ks, vs = [], []
    for c, n in zip(cards, rows):
        k = rng.integers(0, c, n).astype(np.int32)
        if c == 4096 * 12 + 1 and n:
            k[:3] = [c - 1, c - 1, 0]  # the last unit's only key
        ks.append(k)
        v = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
        pick = rng.random(n) < 0.05
        v[pick] = rng.integers(-(2 ** 31), 2 ** 31 - 1, int(pick.sum()), dtype=np.int64).astype(
            np.int32)
        vs.append(v)
    k, v = np.concatenate(ks), np.concatenate(vs)
    row_off = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    card = np.array(cards, np.int32)
    ref = ox.group_batch(k, v, row_off, card)
    b, _, _ = _plan(torch, k, v, row_off, card)
    info = b.info()
    unit = BATCH_UNIT_KEYS
    assert info["units"] == sum(-(-c // unit) if c > 4096 else 1 for c in cards)
    assert info["wide_tiles"] > 0
    for _ in range(2):
        _check(b.run(0), ref)
        _check(b.run(1), ref)
    b.close()
    kb = k.copy()
    kb[row_off[5] + 7] = 250000  # the wide last query's range is [0, 250000)
    assert ox.group_batch(kb, v, row_off, card) == -2
    b, _, _ = _plan(torch, kb, v, row_off, card)
    assert b.run(0)[0] == -2 and b.run(1)[0] == -2
    b.close()
    b, _, _ = _plan(torch, k, v, row_off, card, cap=int(ref[0].size) - 1)
    assert b.run(0)[0] == -7 and b.run(1)[0] == -7
    b.close()


def test_batch_plan_rejects_card_over_2p24(hip):
    import ctypes
    import torch
    L = native.lib("batch")
    kt = torch.zeros(4, dtype=torch.int32, device="cuda")
    ro = np.array([0, 2, 4], np.int64)
    out = torch.zeros(4, dtype=torch.int64, device="cuda")
    go = torch.zeros(3, dtype=torch.int64, device="cuda")
    rc = ctypes.c_int32(0)
    for cd, ok in ((np.array([3, 1 << 24], np.int32), True),
                   (np.array([3, (1 << 24) + 1], np.int32), False)):
        p = L.nfh_group_batch_plan(kt.data_ptr(), kt.data_ptr(), ro.ctypes.data, cd.ctypes.data,
                                   2, out.data_ptr(), out.data_ptr(), out.data_ptr(),
                                   go.data_ptr(), 4, ctypes.byref(rc))
        assert bool(p) == ok and rc.value == (0 if ok else -3)
        if p:
            assert L.nfh_group_batch_run(p, 0) == 2  # keys 0 in both queries
            L.nfh_group_batch_free(p)
