"""GPU: group-by strategies agree with the oracle, the selector's features
and classifier gradients match their CPU restatement, and its device
training learns."""

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox
from oracle import selector as osel

pytestmark = pytest.mark.gpu


def _check(got, ref):
    assert got["ukeys"].tolist() == ref["ukeys"].tolist()
    for k in ("counts", "sums", "mins", "maxs"):
        assert got[k].tobytes() == ref[k].astype(got[k].dtype).tobytes(), k
    assert got["means"].tobytes() == ref["means"].tobytes()


@pytest.mark.parametrize("kind", ["uniform", "zipf", "runs", "wide", "spread", "spread_zipf",
                                  "one", "big_values"])
def test_every_strategy_gives_the_same_groups(hip, kind):
    S = native.lib("selector")
    rng = np.random.default_rng(3)
    n = 1 << 18
    vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
    if kind == "uniform":
        keys = rng.integers(0, 50000, n).astype(np.int32)
    elif kind == "zipf":
        keys = (rng.zipf(1.3, n) % 100000).astype(np.int32)
    elif kind == "runs":
        keys = np.repeat(rng.integers(-5000, 5000, n // 100 + 1), 100)[:n].astype(np.int32)
    elif kind == "wide":
        keys = rng.integers(-(2 ** 31), 2 ** 31 - 1, n, dtype=np.int64).astype(np.int32)
    elif kind == "spread":  # 3000 distinct keys over the whole int32 range
        ids = rng.integers(-(2 ** 31), 2 ** 31 - 1, 3000, dtype=np.int64).astype(np.int32)
        keys = ids[rng.integers(0, 3000, n)]
    elif kind == "spread_zipf":  # hot keys in LDS, the tail in the hash tables
        ids = rng.integers(-(2 ** 31), 2 ** 31 - 1, 200000, dtype=np.int64).astype(np.int32)
        keys = ids[(rng.zipf(1.2, n) - 1) % 200000]
    elif kind == "one":
        keys = np.full(n, -7, np.int32)
    else:  # values outside +-2^15 and int32 extremes
        keys = rng.integers(0, 2000, n).astype(np.int32) * 1000003
        vals = rng.integers(-(2 ** 31), 2 ** 31 - 1, n, dtype=np.int64).astype(np.int32)
    valid = rng.random(n) > 0.05
    ref = ox.group_agg(keys, vals, valid)
    try:
        for st in (1, 2, 3, 4, 5):
            S.nfh_set_groupby_strategy(st)
            _check(native.group_agg(keys, vals, valid), ref)
            ran = S.nfh_groupby_last_strategy()
            # a choice that cannot run on the column falls back to the rule
            assert ran == st or (st in (1, 2, 5) and ran in (1, 2, 3))
    finally:
        S.nfh_set_groupby_strategy(0)


@pytest.mark.parametrize("span", [1, 4096, 4097, 1 << 19, (1 << 19) + 4097, (1 << 24) + 5,
                                  1 << 25])
def test_partition_counts_agree(hip, span):
    """The partitioned path with 2048- and 4096-key buckets (one or two
    partition passes, a last partial bucket, spans only the wider buckets
    take) against the oracle."""
    S = native.lib("selector")
    rng = np.random.default_rng(span % 1000)
    n = 1 << 20
    base = -(1 << 30) + 17
    keys = (base + rng.integers(0, span, n)).astype(np.int32)
    keys[:3] = [base, base + span - 1, base]  # both ends of the span present
    vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
    vals[::97] = rng.integers(-(2 ** 31), 2 ** 31 - 1, vals[::97].size, dtype=np.int64)
    valid = rng.random(n) > 0.05
    valid[:3] = True
    ref = ox.group_agg(keys, vals, valid)
    try:
        for st in (2, 5):
            S.nfh_set_groupby_strategy(st)
            _check(native.group_agg(keys, vals, valid), ref)
            ran = S.nfh_groupby_last_strategy()
            if st == 5 or span <= (1 << 24):
                assert ran == st
    finally:
        S.nfh_set_groupby_strategy(0)


def test_partitioned_bucket_over_2p24_records(hip):
    """One 2048-key bucket receives more than 2^24 records (2^25 rows over
    2048 keys, a quarter of them hot): the bucket pass must not use its
    packed 24-bit counts there."""
    S = native.lib("selector")
    rng = np.random.default_rng(12)
    n = (1 << 25) + 5
    keys = (1000 + rng.integers(0, 2048, n)).astype(np.int32)
    vals = rng.integers(-100, 100, n).astype(np.int32)
    valid = rng.random(n) > 0.01
    ref = ox.group_agg(keys, vals, valid)
    try:
        S.nfh_set_groupby_strategy(2)
        _check(native.group_agg(keys, vals, valid), ref)
        assert S.nfh_groupby_last_strategy() == 2
    finally:
        S.nfh_set_groupby_strategy(0)


def zipf_keys(rng, n, ids, s):
    """Bounded Zipf(s) over ranks 1..ids (inverse CDF), through a seeded
    permutation of [0, ids)."""
    cdf = np.cumsum(np.arange(1, ids + 1, dtype=np.float64) ** -s)
    r = np.searchsorted(cdf / cdf[-1], rng.random(n), side="right")
    return rng.permutation(ids).astype(np.int32)[np.minimum(r, ids - 1)]


def _hist_bytes(L, fn):
    """Runs fn() with kernel timing on; returns (fn's result, the algorithmic
    bytes the partitioned path's histogram pass recorded)."""
    import ctypes
    L.nfh_profile_enable(1)
    L.nfh_profile_reset()
    try:
        res = fn()
        ms, cnt, by = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
        assert L.nfh_profile_get(b"gp_hist", ctypes.byref(ms), ctypes.byref(cnt)) == 0
        assert L.nfh_profile_bytes(b"gp_hist", ctypes.byref(by)) == 0
        return res, by.value
    finally:
        L.nfh_profile_enable(0)
        L.nfh_profile_reset()


@pytest.mark.parametrize("case", ["zipf_tail_row", "zipf_nulls", "all_hot", "wide_values",
                                  "uniform"])
def test_partitioned_record_mode(hip, case):
    """The histogram pass writes the records itself when the sampled hot
    keys carry at least 40% of the rows (one region of rows per workgroup,
    the first partition pass reading the records): a row count whose last
    regions are empty and whose last region ends off a quad, 40% NULL rows,
    every row hot (no records), values outside 16 bits (unpacked records
    after the record mode), and uniform keys (no record mode), for both
    partition counts."""
    S = native.lib("selector")
    L = native.lib("hip")
    rng = np.random.default_rng(21)
    n = {"zipf_tail_row": (1 << 20) + 1, "zipf_nulls": (1 << 21) + 6}.get(case, 1 << 20)
    vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
    valid = rng.random(n) > 0.05
    if case in ("zipf_tail_row", "wide_values"):
        keys = (zipf_keys(rng, n, 1 << 16, 1.1) * 7).astype(np.int32)
    elif case == "zipf_nulls":
        keys = zipf_keys(rng, n, 1 << 20, 1.1) - (1 << 19)
        valid = rng.random(n) > 0.4
    elif case == "all_hot":  # 200 keys spread over a 2^20 span, all sampled hot
        ids = rng.choice(1 << 20, 200, replace=False).astype(np.int32)
        keys = ids[rng.integers(0, 200, n)]
    else:
        keys = rng.integers(0, 1 << 20, n).astype(np.int32)
    if case == "wide_values":
        vals[::31] = rng.integers(-(2 ** 31), 2 ** 31 - 1, vals[::31].size, dtype=np.int64)
    ref = ox.group_agg(keys, vals, valid)
    try:
        for st in (2, 5):
            S.nfh_set_groupby_strategy(st)
            got, by = _hist_bytes(L, lambda: native.group_agg(keys, vals, valid))
            _check(got, ref)
            assert S.nfh_groupby_last_strategy() == st
            # 9 B per row, plus 8 B per record written in record mode
            if case == "uniform" or case == "all_hot":
                assert by == 9.0 * n
            else:
                assert by > 9.0 * n
            # without a validity mask (every row valid: 8 B per row)
            got, by = _hist_bytes(L, lambda: native.group_agg(keys, vals))
            _check(got, ox.group_agg(keys, vals))
            assert by >= 8.0 * n
    finally:
        S.nfh_set_groupby_strategy(0)


def test_hash_strategy_overflow_falls_back(hip):
    """Distinct keys beyond the hash table's room: the overflow is detected
    and the sort-based path gives the result."""
    S = native.lib("selector")
    rng = np.random.default_rng(8)
    n = 1 << 22  # the most rows a forced hash-scatter accepts
    keys = rng.permutation(np.arange(n, dtype=np.int64) * 977 - (1 << 31)).astype(np.int32)
    vals = rng.integers(-100, 100, n).astype(np.int32)
    valid = rng.random(n) > 0.01
    ref = ox.group_agg(keys, vals, valid)
    try:
        S.nfh_set_groupby_strategy(4)
        _check(native.group_agg(keys, vals, valid), ref)
        assert S.nfh_groupby_last_strategy() in (3, 4)
    finally:
        S.nfh_set_groupby_strategy(0)


def test_forced_hash_refused_above_2p22_rows(hip):
    """A forced hash-scatter on more than 2^22 rows returns -3 without
    running; the built-in rule on the same input still runs."""
    S = native.lib("selector")
    L = native.lib("hip")
    n = (1 << 22) + 1
    keys = np.zeros(n, np.int32)
    vals = np.ones(n, np.int32)
    out = [np.zeros(8, dt) for dt in (np.int64, np.int64, np.int64, np.int32, np.int32,
                                      np.float64)]
    try:
        S.nfh_set_groupby_strategy(4)
        rc = L.nfh_group_agg_i32(keys.ctypes.data, vals.ctypes.data, None, None, n,
                                 *[o.ctypes.data for o in out], 8)
        assert rc == -3
        S.nfh_set_groupby_strategy(0)
        rc = L.nfh_group_agg_i32(keys.ctypes.data, vals.ctypes.data, None, None, n,
                                 *[o.ctypes.data for o in out], 8)
        assert rc == 1 and out[1][0] == n and out[2][0] == n
    finally:
        S.nfh_set_groupby_strategy(0)


@pytest.mark.parametrize("kind", ["ramp", "equal", "random", "small"])
def test_features_match_restatement(hip, kind):
    rng = np.random.default_rng(4)
    keys = {"ramp": np.arange(300000, dtype=np.int32),
            "equal": np.full(70000, -3, np.int32),
            "random": rng.integers(-(2 ** 31), 2 ** 31 - 1, 1 << 20, dtype=np.int64).astype(
                np.int32),
            "small": rng.integers(0, 9, 37).astype(np.int32)}[kind]
    got = native.groupby_features(keys)
    assert np.allclose(got, osel.features(keys), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("N", [1, 63, 1000, 5000])
def test_selector_gradient_matches_oracle(hip, N):
    rng = np.random.default_rng(N)
    p = rng.normal(0, 0.3, osel.P).astype(np.float32)
    X = rng.random((N, 64)).astype(np.float32)
    y = rng.integers(0, osel.C, N).astype(np.int32)
    g, loss = native.selector_grad(p, X, y)
    ref = osel.grad(p, X, y)
    scale = max(1e-12, np.abs(ref).max())
    assert np.max(np.abs(g - ref)) <= 2e-4 * scale
    assert abs(loss - osel.loss(p, X, y)) <= 1e-4 * max(1.0, osel.loss(p, X, y))


@pytest.mark.parametrize("N", [1, 700, 5000])
def test_selector_cost_gradient_matches_oracle(hip, N):
    rng = np.random.default_rng(N + 11)
    p = rng.normal(0, 0.3, osel.P).astype(np.float32)
    X = rng.random((N, 64)).astype(np.float32)
    T = rng.normal(0, 1, (N, osel.C)).astype(np.float32)
    g, loss = native.selector_grad_cost(p, X, T)
    ref = osel.grad_cost(p, X, T)
    scale = max(1e-12, np.abs(ref).max())
    assert np.max(np.abs(g - ref)) <= 2e-4 * scale
    assert abs(loss - osel.loss_cost(p, X, T)) <= 1e-4 * max(1.0, osel.loss_cost(p, X, T))


def test_selector_cost_model_learns_costs():
    """Costs from a rule on two features; the trained cost model picks the
    cheapest strategy on held-out samples."""
    rng = np.random.default_rng(6)
    X = rng.random((4000, 64)).astype(np.float32)
    T = np.stack([X[:, 0], 1 - X[:, 0], 0.5 + 0 * X[:, 1], X[:, 1] + 0.3, 0.8 - X[:, 2]],
                 1).astype(np.float32)
    p, _ = native.selector_train_cost(X[:3000], T[:3000], steps=600, lr=1e-2)
    pick = native.selector_predict_cost(p, X[3000:])
    regret = T[3000:][np.arange(1000), pick] - T[3000:].min(1)
    print("mean regret %.4f" % regret.mean())
    assert regret.mean() < 0.02


def test_selector_training_learns_a_rule():
    """Labels from a rule on two features; device training fits it."""
    rng = np.random.default_rng(5)
    X = rng.random((4000, 64)).astype(np.float32)
    y = np.where(X[:, 0] > 0.6, 2, np.where(X[:, 1] > 0.5, 1, np.where(X[:, 2] > 0.8, 3, 0)))
    y = y.astype(np.int32)
    p0 = native.selector_init(1)
    _, l0 = native.selector_grad(p0, X, y)
    p, _ = native.selector_train(X[:3000], y[:3000], steps=600, lr=1e-2, params=p0)
    _, l1 = native.selector_grad(p, X, y)
    acc = np.mean(native.selector_predict(p, X[3000:]) == y[3000:])
    print("loss %.3f -> %.3f, held-out accuracy %.3f" % (l0, l1, acc))
    assert l1 < 0.5 * l0 and acc > 0.85


def test_selector_end_to_end_small():
    import torch
    from numfast_amd import selector
    X, T, ran, meta = selector.label_columns(torch, native, 60, lo=14, hi=17, seed=2)
    assert np.all(T > 0) and X.shape == (60, 64)
    rep, _ = selector.evaluate(native, X, T, ran, steps=100)
    print(rep)
    assert rep["held_out"] == 12 and rep["speedup_vs_fixed_plan"] > 0
