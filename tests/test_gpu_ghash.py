"""GPU parity of the partitioned pair-hash (nf_ghash_*) with the oracle,
bit for bit: fingerprints, per-chunk counts (accumulated), reservation
scatter under several thread counts, the insert's table layout (the
sequential linear-probing layout, also checked by its defining property),
occupancy counts / fills with the per-shard carving, the refusals, device
buffers, and the whole grouped COUNT DISTINCT against np.unique."""

import numpy as np
import pytest

from numfast_amd import ghash, native
from oracle.executor import load_c_oracle
from test_oracle_ghash import (_chunks, check_linear_probing_layout, distinct_ref,
                               driver_prefix)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ora():
    return load_c_oracle()[0]


@pytest.fixture(scope="module")
def lib(hip):
    return native.lib("ghash")


def _p(a):
    return a.ctypes.data


def _data(rng, n, kind):
    if kind == "wide":
        return (rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64),
                rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64))
    if kind == "dups":
        return (rng.integers(-100, 1000, n).astype(np.int64), rng.integers(0, 30, n).astype(np.int64))
    pool = rng.integers(-2**40, 2**40, max(1, n // 50), dtype=np.int64)
    return rng.choice(pool, n), rng.integers(0, 3, n).astype(np.int64)


@pytest.mark.parametrize("n,P", [(1, 16), (4097, 16), (100_000, 64), (1 << 20, 256)])
def test_fp_count_scatter(lib, ora, n, P):
    rng = np.random.default_rng(n)
    k, v = _data(rng, n, "wide" if n % 2 else "dups")
    outs = []
    for L in (lib, ora):
        fp = np.zeros(n, np.int64)
        assert L.nf_ghash_fp_i64(_p(k), _p(v), n, _p(fp), P - 1, 8) == 0
        outs.append(fp)
    assert np.array_equal(outs[0], outs[1])
    fp = outs[0]
    for nth, t in ((1, 1), (3, 16), (16, 16), (64, 64), (100, 7)):
        base = rng.integers(0, 9, t * P).astype(np.int64)
        tcs = []
        for L in (lib, ora):
            tc = base.copy()
            assert L.nf_ghash_count_i64(_p(fp), n, _p(tc), t, P, nth) == 0
            tcs.append(tc)
        assert np.array_equal(tcs[0], tcs[1]), (nth, t)
        T = min(t, 64)
        _, tpos = driver_prefix(fp, n, P, T, nth)
        res = []
        for L in (lib, ora):
            pk, pv = np.full(n, -5, np.int64), np.full(n, -5, np.int64)
            assert L.nf_ghash_scatter_i64(_p(k), _p(v), _p(fp), n, _p(tpos), T, P, _p(pk), _p(pv),
                                          nth) == 0
            res.append((pk, pv))
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_scatter_chunk_major_prefix(lib, ora):
    """A prefix laying later chunks first (not the driver's): the chunk
    boundaries decide the lanes, and they must match."""
    rng = np.random.default_rng(8)
    n, P, T = 50_001, 32, 5
    k, v = _data(rng, n, "dups")
    fp = np.zeros(n, np.int64)
    assert ora.nf_ghash_fp_i64(_p(k), _p(v), n, _p(fp), P - 1, 1) == 0
    cnt = np.array([np.bincount(fp[a:b], minlength=P) for a, b in _chunks(n, T, T)])
    tpos = np.zeros((T, P), np.int64)
    base = 0
    for q in range(P):
        for tt in reversed(range(T)):
            tpos[tt, q] = base
            base += cnt[tt, q]
    tp = np.ascontiguousarray(tpos.reshape(-1))
    res = []
    for L in (lib, ora):
        pk, pv = np.zeros(n, np.int64), np.zeros(n, np.int64)
        assert L.nf_ghash_scatter_i64(_p(k), _p(v), _p(fp), n, _p(tp), T, P, _p(pk), _p(pv), T) == 0
        res.append(pk)
    assert np.array_equal(res[0], res[1])


def _pins_setup(rng, n, P, kind, slotf=1.5):
    k, v = _data(rng, n, kind)
    fp = ((k.view(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
           + (v.view(np.uint64) ^ (v.view(np.uint64) >> np.uint64(33)))
           * np.uint64(0x00BF58476D1FF973)) & np.uint64(P - 1)).astype(np.int64)
    order = np.argsort(fp, kind="stable")
    pk, pv = np.ascontiguousarray(k[order]), np.ascontiguousarray(v[order])
    counts = np.bincount(fp, minlength=P)
    starts = np.zeros(P + 1, np.int64)
    starts[1:] = np.cumsum(counts)
    mp = np.array([ghash._next_pow2(max(64, int(np.ceil(c * slotf)))) for c in counts], np.int64)
    return pk, pv, starts, mp


def _pins(L, pk, pv, starts, mp, seed=7):
    P = starts.size - 1
    toff = np.zeros(P + 1, np.int64)
    toff[1:] = np.cumsum(mp)
    m = int(toff[-1])
    tk, tv = np.full(m, seed, np.int64), np.full(m, -seed, np.int64)
    used = np.zeros(m, np.uint8)
    tm = np.ascontiguousarray(mp - 1)
    rc = L.nf_ghash_pins_i64(_p(pk), _p(pv), _p(starts), _p(toff), _p(tm), _p(tk), _p(tv),
                             _p(used), P, 8)
    return rc, toff, tk, tv, used


@pytest.mark.parametrize("n,P,kind", [(10, 16, "dups"), (20_000, 16, "dups"), (20_000, 64, "pool"),
                                      (300_000, 256, "wide"), (1 << 21, 16, "pool")])
def test_pins_layout_identical(lib, ora, n, P, kind):
    rng = np.random.default_rng(n + P)
    pk, pv, starts, mp = _pins_setup(rng, n, P, kind)
    g = _pins(lib, pk, pv, starts, mp)
    o = _pins(ora, pk, pv, starts, mp)
    assert g[0] == 0 and o[0] == 0
    for a, b in zip(g[1:], o[1:]):
        assert np.array_equal(a, b)
    if n <= 20_000:
        check_linear_probing_layout(pk, pv, starts, *g[1:])


@pytest.mark.parametrize("n,P,slotf", [(1 << 20, 4, 1.1), (300_000, 2, 1.3), (60_000, 1, 1.0)])
def test_pins_windows_that_spill(lib, ora, n, P, slotf):
    """All-distinct pairs in few partitions at half to 0.92 load: tables of
    many windows whose clusters cross window ends (spilled ids, the
    re-solve walk) or outgrow the extension (the global form)."""
    rng = np.random.default_rng(n + P)
    pk, pv, starts, mp = _pins_setup(rng, n, P, "wide", slotf)
    g = _pins(lib, pk, pv, starts, mp)
    o = _pins(ora, pk, pv, starts, mp)
    assert g[0] == o[0] == 0
    for a, b in zip(g[1:], o[1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("slotf", [1.5, 0.1])
def test_pins_rows_before_first_partition(lib, ora, slotf):
    """Rows before starts[0] belong to no partition and are not inserted,
    with tables of at least as many slots as rows (windowed form) and with
    fewer (global form)."""
    rng = np.random.default_rng(17)
    pk, pv, starts, mp = _pins_setup(rng, 20_000, 16, "pool", slotf)
    junk = 37
    pk = np.concatenate([rng.integers(-5, 5, junk), pk]).astype(np.int64)
    pv = np.concatenate([rng.integers(-5, 5, junk), pv]).astype(np.int64)
    starts = starts + junk
    g = _pins(lib, pk, pv, starts, mp)
    o = _pins(ora, pk, pv, starts, mp)
    assert g[0] == o[0] == 0
    for a, b in zip(g[1:], o[1:]):
        assert np.array_equal(a, b)


def test_pins_dense_tables_and_refusals(lib, ora):
    """Tables at their limit (distinct == slots), and one past it (-3
    where the reference would probe forever)."""
    rng = np.random.default_rng(3)
    P = 16
    pk = rng.integers(0, 64, 4000).astype(np.int64)
    pv = np.zeros(4000, np.int64)
    fp = np.zeros(4000, np.int64)
    assert ora.nf_ghash_fp_i64(_p(pk), _p(pv), 4000, _p(fp), P - 1, 1) == 0
    order = np.argsort(fp, kind="stable")
    pk, pv = np.ascontiguousarray(pk[order]), pv
    counts = np.bincount(fp, minlength=P)
    starts = np.zeros(P + 1, np.int64)
    starts[1:] = np.cumsum(counts)
    distinct = np.array([np.unique(pk[starts[q]:starts[q + 1]]).size for q in range(P)])
    exact = np.array([max(1, ghash._next_pow2(max(1, d))) for d in distinct], np.int64)
    g = _pins(lib, pk, pv, starts, exact)
    o = _pins(ora, pk, pv, starts, exact)
    assert g[0] == o[0] == 0
    for a, b in zip(g[1:], o[1:]):
        assert np.array_equal(a, b)
    q = int(np.argmax(distinct))
    small = exact.copy()
    small[q] = ghash._next_pow2(distinct[q]) // 2 if distinct[q] > 1 else 1
    if distinct[q] > small[q]:
        assert _pins(lib, pk, pv, starts, small)[0] == -3
        assert _pins(ora, pk, pv, starts, small)[0] == -3


@pytest.mark.parametrize("nthreads", [1, 5, 64])
def test_occupancy(lib, ora, nthreads):
    rng = np.random.default_rng(nthreads)
    P = 64
    slots = rng.choice([64, 256, 4096], P)
    toff = np.zeros(P + 1, np.int64)
    toff[1:] = np.cumsum(slots)
    m = int(toff[-1])
    used = (rng.random(m) < 0.3).astype(np.uint8) * rng.integers(1, 3, m).astype(np.uint8)
    tk = rng.integers(-10**15, 10**15, m, dtype=np.int64)
    cnts = []
    for L in (lib, ora):
        c = np.full(P, -1, np.int64)
        assert L.nf_ghash_occ_count(_p(used), _p(toff), P, _p(c), nthreads) == 0
        cnts.append(c)
    assert np.array_equal(cnts[0], cnts[1])
    ostart = np.zeros(P + 1, np.int64)
    ostart[1:] = np.cumsum(cnts[0])
    nocc = int(ostart[-1])
    for shift in (0, 3):
        oks = []
        for L in (lib, ora):
            ok = np.full(nocc + shift, 99, np.int64)
            os_ = ostart + shift
            assert L.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(os_), _p(ok), P,
                                       nthreads) == 0
            oks.append(ok)
        assert np.array_equal(oks[0], oks[1])
    short = ostart.copy()
    short[P] -= 1
    ok = np.zeros(nocc, np.int64)
    assert lib.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(short), _p(ok), P, nthreads) == -2
    assert ora.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(short), _p(ok), P, nthreads) == -2


def test_refusals_match(lib, ora):
    fp = np.array([0, 1, 9], np.int64)
    tc = np.zeros(64, np.int64)
    for L in (lib, ora):
        assert L.nf_ghash_count_i64(_p(fp), 3, _p(tc), 1, 16, 1) == 0
        assert L.nf_ghash_count_i64(_p(fp), 3, _p(tc), 1, 4, 1) == -2
        assert L.nf_ghash_count_i64(_p(fp), 3, _p(tc), 1, 3, 1) == -3
        assert L.nf_ghash_count_i64(_p(fp), 3, _p(tc), 0, 4, 1) == -3
        assert L.nf_ghash_fp_i64(None, None, 0, None, 15, 1) == 0
        assert L.nf_ghash_fp_i64(None, _p(fp), 3, _p(fp), 15, 1) == -1
        assert L.nf_ghash_occ_count(None, None, 4, None, 1) == -1


def test_device_buffers(lib, ora):
    """Device pointers pass straight through (torch tensors)."""
    import torch
    rng = np.random.default_rng(5)
    n, P = 70_000, 32
    k, v = _data(rng, n, "pool")
    dk, dv = torch.from_numpy(k).cuda(), torch.from_numpy(v).cuda()
    dfp = torch.empty(n, dtype=torch.int64, device="cuda")
    assert lib.nf_ghash_fp_i64(dk.data_ptr(), dv.data_ptr(), n, dfp.data_ptr(), P - 1, 4) == 0
    torch.cuda.synchronize()
    fp = np.zeros(n, np.int64)
    assert ora.nf_ghash_fp_i64(_p(k), _p(v), n, _p(fp), P - 1, 4) == 0
    assert np.array_equal(dfp.cpu().numpy(), fp)
    dtc = torch.zeros(4 * P, dtype=torch.int64, device="cuda")
    assert lib.nf_ghash_count_i64(dfp.data_ptr(), n, dtc.data_ptr(), 4, P, 4) == 0
    tc = np.zeros(4 * P, np.int64)
    assert ora.nf_ghash_count_i64(_p(fp), n, _p(tc), 4, P, 4) == 0
    assert np.array_equal(dtc.cpu().numpy(), tc)


@pytest.mark.parametrize("n,kind", [(1, "dups"), (777, "wide"), (1 << 20, "pool"),
                                    ((1 << 22) + 4321, "dups"), (1 << 24, "pool")])
def test_grouped_distinct_end_to_end(lib, n, kind):
    rng = np.random.default_rng(n)
    k, v = _data(rng, n, kind)
    uk, nu, _, _ = ghash.grouped_distinct_hash_native(k, v, nthreads=16, lib=lib)
    if kind == "wide":
        rk, rc = distinct_ref(k, v)
    else:
        # small value range: one int64 code per pair keeps np.unique 1-D
        code = np.unique(k * 64 + v)
        rk, rc = np.unique(code >> 6, return_counts=True)
    assert np.array_equal(uk, rk) and np.array_equal(nu, rc)
