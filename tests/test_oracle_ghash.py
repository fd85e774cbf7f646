"""The oracle's partitioned pair-hash (nf_ghash_*, lib.rs:2266-2470,
groupby/hash_grouped.rs) pinned without retyping it:

* fingerprint: hand-computed values (constants, sign, the >> 33 fold) and
  NumPy's wrapping uint64 arithmetic;
* count: per-chunk np.bincount, totals, accumulation into the caller's
  buffer;
* scatter: with the driver's prefix the lanes are np.argsort(fp, stable);
* pins: the occupied pairs of a partition are its distinct pairs, and the
  layout is the one sequential linear probing gives -- every slot from a
  pair's home up to its own slot is occupied by a pair that first occurred
  earlier (this property fixes the layout uniquely);
* occupancy: counts and slot-order fills, the per-shard carving;
* the whole grouped COUNT DISTINCT against np.unique.
No reference vectors exist for this path; these checks are its pins.
"""

import ctypes

import numpy as np
import pytest

from numfast_amd import ghash
from oracle.executor import load_c_oracle

C1, C2 = 0x9E3779B97F4A7C15, 0x00BF58476D1FF973
M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def lib():
    return load_c_oracle()[0]


def _p(a):
    return a.ctypes.data


def _s64(x):
    x &= M64
    return x - (1 << 64) if x >> 63 else x


def fp_of(lib, k, v, pmask=-1, nthreads=1):
    k, v = np.ascontiguousarray(k, np.int64), np.ascontiguousarray(v, np.int64)
    out = np.zeros(k.size, np.int64)
    assert lib.nf_ghash_fp_i64(_p(k), _p(v), k.size, _p(out), pmask, nthreads) == 0
    return out


def test_fp_hand_values(lib):
    cases = [((1, 0), C1), ((0, 1), C2), ((0, 1 << 33), ((1 << 33) ^ 1) * C2),
             ((-1, 0), -C1), ((0, -1), (M64 ^ (M64 >> 33)) * C2), ((2, 3), 2 * C1 + 3 * C2),
             ((0, (1 << 40) + 5), (((1 << 40) + 5) ^ (1 << 7)) * C2)]
    k = np.array([c[0][0] for c in cases], np.int64)
    v = np.array([c[0][1] for c in cases], np.int64)
    got = fp_of(lib, k, v)
    assert got.tolist() == [_s64(c[1]) for c in cases]


def test_fp_numpy_uint64_and_mask(lib):
    rng = np.random.default_rng(1)
    k = rng.integers(-2**63, 2**63 - 1, 5000, dtype=np.int64)
    v = rng.integers(-2**63, 2**63 - 1, 5000, dtype=np.int64)
    uk, uv = k.view(np.uint64), v.view(np.uint64)
    with np.errstate(over="ignore"):
        ref = uk * np.uint64(C1) + (uv ^ (uv >> np.uint64(33))) * np.uint64(C2)
    assert np.array_equal(fp_of(lib, k, v), ref.view(np.int64))
    for P in (16, 64, 256):
        f = fp_of(lib, k, v, P - 1)
        assert np.array_equal(f, (ref & np.uint64(P - 1)).astype(np.int64))
    assert lib.nf_ghash_fp_i64(None, None, 0, None, 15, 1) == 0
    assert lib.nf_ghash_fp_i64(None, _p(v), 3, _p(v), 15, 1) == -1


def _chunks(n, nthreads, t):
    nt = min(max(1, min(nthreads, 64)), max(n, 1), t)
    cs = -(-n // nt)
    return [(min(i * cs, n), min((i + 1) * cs, n)) for i in range(nt)]


@pytest.mark.parametrize("n,P,nthreads,t", [(1, 16, 1, 1), (1000, 16, 3, 16), (12345, 64, 16, 16),
                                            (777, 256, 100, 64), (5000, 32, 7, 4)])
def test_count_is_per_chunk_bincount(lib, n, P, nthreads, t):
    rng = np.random.default_rng(n)
    fp = rng.integers(0, P, n).astype(np.int64)
    tc = rng.integers(0, 5, t * P).astype(np.int64)  # accumulated into
    before = tc.copy()
    assert lib.nf_ghash_count_i64(_p(fp), n, _p(tc), t, P, nthreads) == 0
    exp = before.reshape(t, P).copy()
    for tt, (a, b) in enumerate(_chunks(n, nthreads, t)):
        exp[tt] += np.bincount(fp[a:b], minlength=P)
    assert np.array_equal(tc.reshape(t, P), exp)
    assert (tc - before).sum() == n and np.array_equal((tc - before).reshape(t, P).sum(0),
                                                       np.bincount(fp, minlength=P))


def test_count_refusals(lib):
    fp = np.array([0, 1, 2], np.int64)
    tc = np.zeros(64, np.int64)
    assert lib.nf_ghash_count_i64(_p(fp), 3, _p(tc), 1, 3, 1) == -3   # p not a power of two
    assert lib.nf_ghash_count_i64(_p(fp), 3, _p(tc), 0, 4, 1) == -3   # t == 0
    assert lib.nf_ghash_count_i64(_p(fp), 3, _p(tc), 65, 4, 1) == -3  # t > 64
    assert lib.nf_ghash_count_i64(_p(fp), 3, _p(tc), 1, 512, 1) == -3
    bad = np.array([0, 4, 1], np.int64)
    assert lib.nf_ghash_count_i64(_p(bad), 3, _p(tc), 1, 4, 1) == -2  # id outside [0, p)
    assert lib.nf_ghash_count_i64(None, 0, None, 1, 3, 1) == 0


def driver_prefix(fp, n, P, T, nthreads):
    """grouped_native_hash.py:296-307: counts per chunk -> starts, tpos."""
    tc = np.zeros((T, P), np.int64)
    for tt, (a, b) in enumerate(_chunks(n, nthreads, T)):
        tc[tt] = np.bincount(fp[a:b], minlength=P)
    col = np.cumsum(tc, axis=0)
    starts = np.zeros(P + 1, np.int64)
    starts[1:] = np.cumsum(col[-1])
    tpos = np.empty((T, P), np.int64)
    tpos[0] = starts[:-1]
    if T > 1:
        tpos[1:] = starts[:-1] + col[:-1]
    return starts, np.ascontiguousarray(tpos.reshape(-1))


@pytest.mark.parametrize("n,P,T", [(1, 16, 1), (3000, 16, 4), (20000, 64, 16), (999, 256, 64)])
def test_scatter_with_driver_prefix_is_stable_partition(lib, n, P, T):
    rng = np.random.default_rng(n + P)
    k = rng.integers(-10**12, 10**12, n, dtype=np.int64)
    v = rng.integers(-10**12, 10**12, n, dtype=np.int64)
    fp = fp_of(lib, k, v, P - 1)
    _, tpos = driver_prefix(fp, n, P, T, T)
    pk, pv = np.zeros(n, np.int64), np.zeros(n, np.int64)
    assert lib.nf_ghash_scatter_i64(_p(k), _p(v), _p(fp), n, _p(tpos), T, P, _p(pk), _p(pv),
                                    T) == 0
    order = np.argsort(fp, kind="stable")
    assert np.array_equal(pk, k[order]) and np.array_equal(pv, v[order])


def test_scatter_other_prefix_and_refusals(lib):
    """A prefix that puts later chunks first: each chunk still writes its
    rows in row order from its own cursors."""
    rng = np.random.default_rng(3)
    n, P, T = 4000, 16, 4
    k = np.arange(n, dtype=np.int64)
    v = rng.integers(0, 50, n).astype(np.int64)
    fp = fp_of(lib, k, v, P - 1)
    ch = _chunks(n, T, T)
    cnt = np.array([np.bincount(fp[a:b], minlength=P) for a, b in ch])
    tpos = np.zeros((T, P), np.int64)
    base = 0
    for q in range(P):
        for tt in reversed(range(T)):
            tpos[tt, q] = base
            base += cnt[tt, q]
    pk, pv = np.zeros(n, np.int64), np.zeros(n, np.int64)
    tp = np.ascontiguousarray(tpos.reshape(-1))
    assert lib.nf_ghash_scatter_i64(_p(k), _p(v), _p(fp), n, _p(tp), T, P, _p(pk), _p(pv), T) == 0
    for tt, (a, b) in enumerate(ch):
        for q in range(P):
            rows = np.flatnonzero(fp[a:b] == q) + a
            assert np.array_equal(pk[tpos[tt, q]:tpos[tt, q] + rows.size], k[rows])
    big = tp.copy()
    big[0] = n  # a reservation past the end
    assert lib.nf_ghash_scatter_i64(_p(k), _p(v), _p(fp), n, _p(big), T, P, _p(pk), _p(pv),
                                    T) == -2


def pins_tables(lib, pk, pv, starts, slots):
    P = starts.size - 1
    toff = np.zeros(P + 1, np.int64)
    toff[1:] = np.cumsum(slots)
    m = int(toff[-1])
    tk = np.full(m, 7, np.int64)
    tv = np.full(m, 7, np.int64)
    used = np.zeros(m, np.uint8)
    tm = np.ascontiguousarray(np.asarray(slots, np.int64) - 1)
    rc = lib.nf_ghash_pins_i64(_p(pk), _p(pv), _p(starts), _p(toff), _p(tm), _p(tk), _p(tv),
                               _p(used), P, 4)
    return rc, toff, tk, tv, used


def check_linear_probing_layout(pk, pv, starts, toff, tk, tv, used):
    """Occupied pairs == distinct pairs of the partition, and every slot
    from a pair's home to its slot is occupied by an earlier-first pair."""
    P = starts.size - 1
    for q in range(P):
        a, b = int(starts[q]), int(starts[q + 1])
        s0, s1 = int(toff[q]), int(toff[q + 1])
        mask = s1 - s0 - 1
        first = {}
        for j in range(a, b):
            first.setdefault((int(pk[j]), int(pv[j])), j)
        occ = [(int(tk[s]), int(tv[s])) for s in range(s0, s1) if used[s]]
        assert len(occ) == len(set(occ)) and set(occ) == set(first)
        for s in range(s0, s1):
            if not used[s]:
                continue
            pair = (int(tk[s]), int(tv[s]))
            h = int(ghash_fp(*pair)) & mask
            x = h
            while x != s - s0:
                t = s0 + x
                assert used[t], "hole between home and slot"
                assert first[(int(tk[t]), int(tv[t]))] < first[pair], "later pair ahead"
                x = (x + 1) & mask


def ghash_fp(k, v):
    uk, uv = k & M64, v & M64
    return (uk * C1 + (uv ^ (uv >> 33)) * C2) & M64


@pytest.mark.parametrize("dups,load", [(1, 0.3), (5, 0.6), (30, 0.95)])
def test_pins_layout_is_sequential_linear_probing(lib, dups, load):
    rng = np.random.default_rng(dups)
    P = 16
    sizes = rng.integers(0, 300, P)
    sizes[3] = 0
    starts = np.zeros(P + 1, np.int64)
    starts[1:] = np.cumsum(sizes)
    n = int(starts[-1])
    pk = rng.integers(-1000, 1000, n).astype(np.int64) // dups
    pv = rng.integers(0, 3, n).astype(np.int64)
    slots = []
    for q in range(P):
        d = len(set(zip(pk[starts[q]:starts[q + 1]].tolist(), pv[starts[q]:starts[q + 1]].tolist())))
        s = 1
        while s < max(1, d / load):
            s <<= 1
        slots.append(s)
    rc, toff, tk, tv, used = pins_tables(lib, pk, pv, starts, slots)
    assert rc == 0
    check_linear_probing_layout(pk, pv, starts, toff, tk, tv, used)
    # unused slots keep the caller's contents
    assert (tk[used == 0] == 7).all() and (tv[used == 0] == 7).all()


def test_pins_full_and_refusals(lib):
    pk = np.array([1, 2, 3, 1, 2], np.int64)
    pv = np.zeros(5, np.int64)
    starts = np.array([0, 5], np.int64)
    rc, toff, tk, tv, used = pins_tables(lib, pk, pv, starts, [4])  # 3 distinct in 4 slots
    assert rc == 0 and used.sum() == 3
    rc, *_ = pins_tables(lib, pk, pv, starts, [2])  # 3 distinct, 2 slots: would probe forever
    assert rc == -3
    pk3 = np.array([5, 6, 7, 5], np.int64)
    rc, toff, tk, tv, used = pins_tables(lib, pk3, np.zeros(4, np.int64), np.array([0, 4], np.int64),
                                         [4])
    assert rc == 0 and used.sum() == 3  # exactly full minus one, duplicate found
    rc, *_ = pins_tables(lib, pk, pv, starts, [3])  # not a power of two
    assert rc == -3


@pytest.mark.parametrize("nthreads", [1, 3, 16, 64])
def test_occupancy_count_and_fill(lib, nthreads):
    rng = np.random.default_rng(nthreads)
    P = 32
    slots = rng.choice([64, 128, 256], P)
    toff = np.zeros(P + 1, np.int64)
    toff[1:] = np.cumsum(slots)
    m = int(toff[-1])
    used = (rng.random(m) < 0.4).astype(np.uint8)
    tk = rng.integers(-10**9, 10**9, m, dtype=np.int64)
    cnt = np.zeros(P, np.int64)
    assert lib.nf_ghash_occ_count(_p(used), _p(toff), P, _p(cnt), nthreads) == 0
    exp_cnt = np.array([used[toff[q]:toff[q + 1]].sum() for q in range(P)])
    assert np.array_equal(cnt, exp_cnt)
    ostart = np.zeros(P + 1, np.int64)
    ostart[1:] = np.cumsum(cnt)
    nocc = int(ostart[-1])
    ok = np.zeros(nocc, np.int64)
    assert lib.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(ostart), _p(ok), P, nthreads) == 0
    assert np.array_equal(ok, tk[used.astype(bool)])
    # ostart shifted by a constant: pieces are carved from ok[0]
    sh = ostart + 5
    ok2 = np.full(nocc + 5, -1, np.int64)
    assert lib.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(sh), _p(ok2), P, nthreads) == 0
    assert np.array_equal(ok2[:nocc], ok) and (ok2[nocc:] == -1).all()
    short = ostart.copy()
    short[P] -= 1  # the last piece is one entry short
    assert lib.nf_ghash_occ_fill(_p(tk), _p(used), _p(toff), _p(short), _p(ok), P, nthreads) == -2


def distinct_ref(k, v):
    pairs = np.unique(np.stack([np.asarray(k, np.int64), np.asarray(v, np.int64)], 1), axis=0)
    uk, cnt = np.unique(pairs[:, 0], return_counts=True)
    return uk, cnt


@pytest.mark.parametrize("n,kinds", [(1, "tiny"), (5000, "dups"), (200_000, "wide"),
                                     ((1 << 22) + 999, "dups")])
def test_grouped_distinct_end_to_end(lib, n, kinds):
    rng = np.random.default_rng(n)
    if kinds == "wide":
        k = rng.integers(-2**40, 2**40, n, dtype=np.int64)
        v = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    else:
        k = rng.integers(-50, 500, n).astype(np.int32)
        v = rng.integers(0, 40, n).astype(np.int16)
    for nth in (1, 5):
        uk, nu, _, _ = ghash.grouped_distinct_hash_native(k, v, nthreads=nth, lib=lib)
        rk, rc = distinct_ref(k, v)
        assert np.array_equal(uk, rk) and np.array_equal(nu, rc)


def test_layout_check_rejects_other_insertion_orders(lib):
    """The layout pin must fail on what a plausible mistake produces:
    reverse-order insertion (swap of two pairs inside one probe cluster)."""
    rng = np.random.default_rng(12)
    n = 600
    pool = rng.integers(-2**40, 2**40, 400, dtype=np.int64)
    pk = rng.choice(pool, n)
    pv = rng.integers(0, 2, n).astype(np.int64)
    starts = np.array([0, n], np.int64)
    rc, toff, tk, tv, used = pins_tables(lib, pk, pv, starts, [512])
    assert rc == 0
    check_linear_probing_layout(pk, pv, starts, toff, tk, tv, used)
    # the same pairs inserted last row first
    rc2, _, tk2, tv2, used2 = pins_tables(lib, pk[::-1].copy(), pv[::-1].copy(), starts, [512])
    assert rc2 == 0 and np.array_equal(used2, used)
    assert not np.array_equal(tk2, tk)
    with pytest.raises(AssertionError):
        check_linear_probing_layout(pk, pv, starts, toff, tk2, tv2, used2)
