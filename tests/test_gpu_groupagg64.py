"""GPU parity of the float64 / int64 group-by with min / max / mean
(nfh_group_agg_f64 / nfh_group_agg_i64) against the oracle's restatement of
the executor's min/max branch (cpu.py:3507-3563, :2570-2601), including the
CPU reference configuration: 1,048,576 rows, int32 key uniform in
[0, 1000), int32 value uniform in [-1000, 1000], float64 value uniform in
[0, 1000); group(key, {v: (sum, count, min, max)}) per value column, then
sort by key; seed 0."""

import json
from pathlib import Path

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu
F64_RTOL = 1e-12  # specs-rebuilt/conformance-profile.toml:19-22
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _f64_close(got, ref):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    both_nan = np.isnan(got) & np.isnan(ref)
    ok = both_nan | (got == ref) | (np.abs(got - ref) <= F64_RTOL * np.abs(ref))
    return bool(ok.all())


def _check(got, ref, exact_sums=False):
    assert got["ukeys"].tolist() == ref["ukeys"].tolist()
    assert got["counts"].tobytes() == ref["counts"].tobytes()
    dt = got["mins"].dtype
    for k in ("mins", "maxs"):
        assert got[k].tobytes() == ref[k].astype(dt).tobytes(), k
    if dt == np.float64 and not exact_sums:
        assert _f64_close(got["sums"], ref["sums"])
        assert _f64_close(got["means"], ref["means"])
    else:
        assert got["sums"].tobytes() == ref["sums"].astype(dt).tobytes()
        assert got["means"].tobytes() == ref["means"].tobytes()


def test_cpu_reference_configuration():
    """Config 1's exact query, both value columns, then sort by key (the
    stable radix argsort of the int64 group keys through the C ABI)."""
    rng = np.random.default_rng(0)
    n = 1_048_576
    key = rng.integers(0, 1000, n, dtype=np.int32)
    v_i32 = rng.integers(-1000, 1001, n, dtype=np.int32)
    v_f64 = rng.uniform(0, 1000, n)
    gi = native.group_agg(key, v_i32)
    gf = native.group_agg64(key, v_f64)
    ri, rf = ox.group_agg(key, v_i32), ox.group_agg(key, v_f64)
    for k in ("ukeys", "counts", "sums", "mins", "maxs", "means"):
        assert gi[k].tobytes() == ri[k].astype(gi[k].dtype).tobytes(), k
    _check(gf, rf)
    # sort("key"): the group keys through nf_sort_perm_i64, stable ascending
    perm = native.sort_perm_raw(gf["ukeys"], False)
    assert np.array_equal(perm, ox.sort_step(rf["ukeys"], False))
    assert np.array_equal(gf["ukeys"][perm], np.arange(1000))
    # error against exact rational sums on the first 50 groups (pure Python)
    few = key < 50
    exact = ox.exact_sums_f64(key[few], v_f64[few], 50)
    err = ox.rel_err_vs_exact(gf["sums"][:50], exact)
    ref_err = ox.rel_err_vs_exact(rf["sums"][:50], exact)
    print("config 1 float64 sums, max rel err vs exact: hip %.3g, row order %.3g"
          % (err, ref_err))
    assert err <= 1e-12


@pytest.mark.parametrize("n,span,kind", [
    (1, 1, "plain"), (7, 3, "plain"), (4099, 50, "specials"), (100_003, 1000, "specials"),
    (1 << 20, 1 << 16, "plain"), (1 << 18, 1 << 30, "wide"), (1 << 16, 1 << 31, "extremes"),
    (300_001, 5, "zeros"),
])
@pytest.mark.parametrize("dt", [np.float64, np.int64])
def test_group_agg64_parity(n, span, kind, dt):
    rng = np.random.default_rng(n % 97 + span % 89)
    if kind == "extremes":
        keys = rng.integers(-(2 ** 31), 2 ** 31, n, dtype=np.int64).astype(np.int32)
        keys[:2] = [-(2 ** 31), 2 ** 31 - 1]
    else:
        base = -(1 << 29) if kind == "wide" else -17
        keys = (base + rng.integers(0, span, n)).astype(np.int32)
    if dt == np.float64:
        vals = rng.uniform(-1e6, 1e6, n) * 10.0 ** rng.integers(-8, 8, n)
        if kind == "specials":
            pick = rng.integers(0, 8, n)
            vals[pick == 0] = np.nan
            vals[pick == 1] = 0.0
            vals[pick == 2] = -0.0
            vals[pick == 3] = np.inf
            vals[pick == 4] = -np.inf
        if kind == "zeros":
            vals = rng.choice([0.0, -0.0, 1.0, -1.0], n)
    else:
        vals = rng.integers(-(2 ** 62), 2 ** 62, n, dtype=np.int64)
        if kind in ("specials", "extremes"):
            vals[rng.integers(0, n, max(1, n // 10))] = np.iinfo(np.int64).min
            vals[rng.integers(0, n, max(1, n // 10))] = np.iinfo(np.int64).max
    valid = rng.random(n) > 0.05
    kvalid = rng.random(n) > 0.05
    got = native.group_agg64(keys, vals, valid, kvalid)
    ref = ox.group_agg(keys, vals, valid, kvalid)
    # int64 sums wrap like np.add.at on int64 (the executor admits such a
    # column only under its 2**53 gate, where both are exact)
    _check(got, ref)


def test_group_agg64_signed_zero_and_nan_order():
    """np.minimum/maximum.at keep the first NaN of a group and give a zero
    extreme the sign of the group's last zero."""
    keys = np.array([0, 0, 1, 1, 2, 2, 2, 3, 3, 3], np.int32)
    vals = np.array([0.0, -0.0, -0.0, 0.0, 5.0, np.nan, -1.0, -0.0, 2.0, -0.0])
    got = native.group_agg64(keys, vals)
    ref = ox.group_agg(keys, vals)
    _check(got, ref, exact_sums=True)
    assert np.signbit(got["mins"]).tolist() == [True, False, False, True]
    assert np.isnan(got["mins"][2]) and np.isnan(got["maxs"][2])


def test_group_agg64_null_groups_dropped_and_empty():
    keys = np.array([4, 4, 9, 9, 1], np.int32)
    vals = np.array([1.5, 2.5, 3.0, 4.0, 7.0])
    valid = np.array([1, 1, 0, 0, 1], np.uint8)
    got = native.group_agg64(keys, vals, valid)
    assert got["ukeys"].tolist() == [1, 4]
    assert got["counts"].tolist() == [1, 2]
    assert got["sums"].tolist() == [7.0, 4.0]
    none = native.group_agg64(keys, vals, np.zeros(5, np.uint8))
    assert none["ukeys"].size == 0


def test_group_agg64_capacity():
    """cap below the group count returns -7; exactly ng lanes succeed."""
    L = native.lib("hip")
    keys = np.arange(100, dtype=np.int32) % 10
    vals = np.arange(100, dtype=np.float64)
    outs = [np.zeros(10, dt) for dt in (np.int64, np.int64, np.float64, np.float64,
                                        np.float64, np.float64)]
    p = [o.ctypes.data for o in outs]
    assert L.nfh_group_agg_f64(keys.ctypes.data, vals.ctypes.data, None, None, 100, *p, 9) == -7
    assert L.nfh_group_agg_f64(keys.ctypes.data, vals.ctypes.data, None, None, 100, *p, 10) == 10
    assert outs[2].tolist() == [float(sum(range(k, 100, 10))) for k in range(10)]


def test_group_golden_float_cases():
    """The reference's float group-by vectors (tests/golden) through the f64
    entry point."""
    seen = 0
    for case in GOLD["group"]:
        vals = np.array(case["vals"])
        if vals.dtype.kind != "f":
            continue
        got = native.group_agg64(np.array(case["keys"], np.int32), vals, case.get("valid"))
        assert got["ukeys"].tolist() == case["ukeys"], case["src"]
        if "sums" in case:
            assert _f64_close(got["sums"], case["sums"]), case["src"]
        if "counts" in case:
            assert got["counts"].tolist() == case["counts"], case["src"]
        if "means" in case:
            assert _f64_close(got["means"], case["means"]), case["src"]
        seen += 1
    assert seen > 0
