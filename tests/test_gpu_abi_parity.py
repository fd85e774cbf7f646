"""GPU parity of every reference C-ABI symbol: libnumfast_hip vs the oracle.

Both libraries are driven through the same ctypes signatures (the
reference loader's argtypes, numfast_amd/native.py).  Integer, count,
index and unique outputs must match bit for bit; float64 sums within
1e-12 relative (specs-rebuilt/conformance-profile.toml:19-22) of the
row-order oracle, and within 1 ulp-ish of the exact-rational sum.
"""

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu
I32MIN, I32MAX = -(2 ** 31), 2 ** 31 - 1
F64_RTOL = 1e-12


def _sort(lib, k, desc):
    k = np.ascontiguousarray(k)
    n = k.size
    tu = np.uint32 if k.dtype == np.int32 else np.uint64
    perm = np.full(n, -7, np.int32)
    t0, t1, tp = np.empty(n, tu), np.empty(n, tu), np.empty(n, np.int32)
    fn = lib.nf_sort_perm_i32 if k.dtype == np.int32 else lib.nf_sort_perm_i64
    rc = fn(k.ctypes.data, n, 1 if desc else 0, perm.ctypes.data, t0.ctypes.data,
            t1.ctypes.data, tp.ctypes.data)
    return rc, perm


def _unique(lib, k):
    k = np.ascontiguousarray(k)
    n = k.size
    tu = np.uint32 if k.dtype == np.int32 else np.uint64
    u, inv, perm = np.zeros(n, k.dtype), np.zeros(n, np.int32), np.empty(n, np.int32)
    t0, t1, tp = np.empty(n, tu), np.empty(n, tu), np.empty(n, np.int32)
    fn = lib.nf_unique_inverse_i32 if k.dtype == np.int32 else lib.nf_unique_inverse_i64
    ng = fn(k.ctypes.data, n, u.ctypes.data, inv.ctypes.data, perm.ctypes.data,
            t0.ctypes.data, t1.ctypes.data, tp.ctypes.data)
    return ng, u[:max(ng, 0)], inv


def _keys(dist, n, dt, rng):
    info = np.iinfo(dt)
    if dist == "uniform":
        return rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    if dist == "small":
        return rng.integers(-3, 4, n).astype(dt)
    if dist == "const":
        return np.full(n, 11, dt)
    if dist == "sorted_swaps":
        k = np.sort(rng.integers(info.min, info.max, n, dtype=dt, endpoint=True))
        if n < 2:
            return k
        idx = rng.choice(n, size=max(2, n // 10), replace=False)
        k[idx] = k[rng.permutation(idx)]
        return k
    if dist == "distinct1024":
        pool = rng.integers(info.min, info.max, 1024, dtype=dt, endpoint=True)
        return pool[rng.integers(0, 1024, n)]
    if dist == "extremes":
        return rng.choice(np.array([info.min, info.max, 0, -1, 1], dtype=dt), n)
    raise ValueError(dist)


SIZES = [1, 2, 63, 64, 65, 4095, 4096, 4097, 100_003, 1 << 20]
DISTS = ["uniform", "small", "const", "sorted_swaps", "distinct1024", "extremes"]


@pytest.mark.parametrize("dt", [np.int32, np.int64])
@pytest.mark.parametrize("dist", DISTS)
def test_sort_perm_parity(hip, c_oracle, dt, dist):
    rng = np.random.default_rng(0)
    for n in SIZES:
        if dt == np.int64 and n > 100_003:
            continue
        k = _keys(dist, n, dt, rng)
        for desc in (False, True):
            rc, got = _sort(hip, k, desc)
            rc0, ref = _sort(c_oracle, k, desc)
            assert rc == rc0 == 0
            assert np.array_equal(got, ref), (dist, n, desc)


def test_sort_perm_edge_rc(hip):
    k = np.zeros(4, np.int32)
    t = np.zeros(4, np.uint32)
    p = np.zeros(4, np.int32)
    assert hip.nf_sort_perm_i32(k.ctypes.data, 0, 0, None, None, None, None) == 0
    assert hip.nf_sort_perm_i32(k.ctypes.data, 4, 0, p.ctypes.data, None, t.ctypes.data,
                                p.ctypes.data) == -1
    assert hip.nf_sort_perm_i32(k.ctypes.data, 2 ** 31, 0, p.ctypes.data, t.ctypes.data,
                                t.ctypes.data, p.ctypes.data) == -3


@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_unique_parity(hip, c_oracle, dt):
    rng = np.random.default_rng(42)
    for n in (1, 5, 4096, 4097, 100_000, 1 << 20):
        for card in (1, 2, 1000, n):
            k = rng.integers(-card, card, n).astype(dt)
            ng, u, inv = _unique(hip, k)
            ng0, u0, inv0 = _unique(c_oracle, k)
            assert ng == ng0
            assert np.array_equal(u, u0) and np.array_equal(inv, inv0)
            assert np.array_equal(u[inv], k)
    k = np.zeros(5, dt)
    assert hip.nf_unique_inverse_i32(k.ctypes.data, 0, *([None] * 6)) == 0
    assert hip.nf_unique_inverse_i32(k.ctypes.data, 5, None, *([k.ctypes.data] * 5)) == -2


def _variant(lib, sym, keys, vals, variant, g):
    keys = np.ascontiguousarray(keys, np.int32)
    dt = {"nf_group_variant_i64": np.int64, "nf_group_variant_i32": np.int32,
          "nf_group_variant_f64": np.float64}[sym]
    vals = np.ascontiguousarray(vals, dt)
    sums, counts = np.full(g, 99, dt), np.full(g, 99, np.int64)
    rc = getattr(lib, sym)(keys.ctypes.data, vals.ctypes.data, keys.size, variant,
                           sums.ctypes.data, counts.ctypes.data, g)
    return rc, sums, counts


def test_variant_golden_and_grid(hip, c_oracle):
    """The panic-safety parity grid (tests/fast/test_ops_panic_safety.py:113-178):
    5 D x 5 n x 6 patterns, V_DENSE == V_CHECKED == oracle."""
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
    for case in gold["abi"]:
        rc, s, c = _variant(hip, case["sym"], case["keys"], case["vals"], case["variant"],
                            case["g"])
        assert rc == case["rc"], case["src"]
        if "sums" in case:
            assert s.tolist() == case["sums"] and c.tolist() == case["counts"]
    rng = np.random.default_rng(42)
    pats = ("zeros", "ramp", "uniform", "reversed", "max_key", "single_hot")
    cells = 0
    for d in (1, 2, 8, 1000, 65536):
        for n in (1, 2, 7, 1000, 100_000):
            for pat in pats:
                if pat == "zeros":
                    keys = np.zeros(n, np.int32)
                elif pat == "ramp":
                    keys = (np.arange(n) % d).astype(np.int32)
                elif pat == "uniform":
                    keys = rng.integers(0, d, n).astype(np.int32)
                elif pat == "reversed":
                    keys = ((n - 1 - np.arange(n)) % d).astype(np.int32)
                elif pat == "max_key":
                    keys = np.full(n, d - 1, np.int32)
                else:
                    keys = np.zeros(n, np.int32)
                    keys[n - 1] = d - 1
                data = rng.integers(-(2 ** 20), 2 ** 20, n).astype(np.int64)
                ref = _variant(c_oracle, "nf_group_variant_i64", keys, data, 1, d)
                for v in (0, 1):
                    got = _variant(hip, "nf_group_variant_i64", keys, data, v, d)
                    assert got[0] == ref[0] == 0
                    assert got[1].tobytes() == ref[1].tobytes()
                    assert got[2].tobytes() == ref[2].tobytes()
                cells += 1
    assert cells == 150


PRECEDENCE = [
    ("nf_group_variant_i32", [0, 0, 0], [I32MAX, 1, -1], 0, 1),
    ("nf_group_variant_i32", [0, 0, 0], [I32MAX, -1, 1], 0, 1),
    ("nf_group_variant_i32", [0, 0, 5], [I32MAX, 1, 0], 0, 2),
    ("nf_group_variant_i32", [0, 5, 0], [I32MAX, 0, 1], 0, 2),
    ("nf_group_variant_i32", [1, 1], [I32MIN, -1], 0, 2),
    ("nf_group_variant_i32", [0, 1, 0], [1, 1, 1], 2, 2),
    ("nf_group_variant_i32", [1, 1, 0], [I32MAX, 1, 1], 2, 2),
    ("nf_group_variant_i32", [1, 0, 0], [I32MAX, 1, 1], 2, 2),
    ("nf_group_variant_i32", [0, 0], [1, 2], 1, 1),
    ("nf_group_variant_i64", [0, 0], [2 ** 62, 2 ** 62], 1, 2),
    ("nf_group_variant_i64", [0, 0, 0], [2 ** 62, 2 ** 62, -2 ** 62], 0, 1),
    ("nf_group_variant_i64", [0, 1, 0, 1], [2 ** 62, 2 ** 62, 2 ** 62 - 1, -1], 1, 2),
    ("nf_group_variant_i64", [0, 0], [-(2 ** 63), -1], 1, 1),
    ("nf_group_variant_i64", [0, 0, 1], [1, 2, 3], 2, 2),
    ("nf_group_variant_i64", [0], [1], 3, 1),
    ("nf_group_variant_f64", [0], [1.0], 7, 1),
    ("nf_group_variant_f64", [0, 3], [1.0, 2.0], 1, 2),
]


@pytest.mark.parametrize("case", PRECEDENCE, ids=lambda c: "%s-%r-v%d" % (c[0], c[1], c[3]))
def test_variant_rc_precedence(hip, c_oracle, case):
    sym, keys, vals, variant, g = case
    got = _variant(hip, sym, keys, vals, variant, g)
    ref = _variant(c_oracle, sym, keys, vals, variant, g)
    assert got[0] == ref[0]
    if ref[0] == 0:
        assert got[1].tobytes() == ref[1].tobytes() and np.array_equal(got[2], ref[2])


def test_variant_overflow_random_prefixes(hip, c_oracle):
    """Large values in few groups: many groups exceed the |v| bound, so the
    exact row-order check runs; rc and outputs must match the oracle."""
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(2, 3000))
        g = int(rng.integers(1, 6))
        keys = rng.integers(0, g, n).astype(np.int32)
        vals = rng.integers(-(2 ** 29), 2 ** 29, n).astype(np.int32)
        if trial % 3 == 0:
            keys[rng.integers(0, n)] = g + 3
        for v in (0, 2):
            k = np.sort(keys) if v == 2 and trial % 2 else keys
            got = _variant(hip, "nf_group_variant_i32", k, vals, v, g)
            ref = _variant(c_oracle, "nf_group_variant_i32", k, vals, v, g)
            assert got[0] == ref[0], (trial, v)
            if ref[0] == 0:
                assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
        v64 = (vals.astype(np.int64) << 33)
        got = _variant(hip, "nf_group_variant_i64", keys, v64, 1, g)
        ref = _variant(c_oracle, "nf_group_variant_i64", keys, v64, 1, g)
        assert got[0] == ref[0], trial
        if ref[0] == 0:
            assert np.array_equal(got[1], ref[1])


def _f64_close(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    both_nan = np.isnan(got) & np.isnan(ref)
    ok = both_nan | (got == ref) | (np.abs(got - ref) <= np.maximum(1e-12, F64_RTOL * np.abs(ref)))
    return bool(np.all(ok))


@pytest.mark.parametrize("g", [1, 7, 1000, 100_000])
def test_f64_sum_count_parity(hip, c_oracle, g):
    rng = np.random.default_rng(0)
    n = 1 << 20
    keys = rng.integers(0, g, n).astype(np.int32)
    vals = rng.uniform(0, 1000, n)
    s, c = np.zeros(g), np.zeros(g, np.int64)
    s0, c0 = np.zeros(g), np.zeros(g, np.int64)
    assert hip.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, n, s.ctypes.data,
                                  c.ctypes.data, g) == 0
    assert c_oracle.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, n, s0.ctypes.data,
                                       c0.ctypes.data, g) == 0
    assert np.array_equal(c, c0)
    assert _f64_close(s, s0)
    for variant in (0, 1):
        rc, s1, c1 = _variant(hip, "nf_group_variant_f64", keys, vals, variant, g)
        assert rc == 0 and np.array_equal(s1.view(np.int64), s.view(np.int64))


@pytest.mark.parametrize("scale", ["mixed_magnitudes", "cancelling", "positive"])
def test_f64_exactness_vs_rational(hip, c_oracle, scale):
    """Each group's sum is either the correctly rounded exact sum or, when the
    reference's worst-case row-order error could exceed the tolerance, the
    reference's own row-order sum bit for bit; either way within 1e-12 of
    the reference.  The max error of both against the exact rational sum is
    reported."""
    rng = np.random.default_rng(5)
    n, g = 20000, 13
    keys = rng.integers(0, g, n).astype(np.int32)
    if scale == "mixed_magnitudes":
        vals = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 9, n)
    elif scale == "cancelling":
        vals = rng.uniform(-1e3, 1e3, n)
    else:
        vals = rng.uniform(0, 1e3, n)
    s, c = np.zeros(g), np.zeros(g, np.int64)
    assert hip.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, n, s.ctypes.data,
                                  c.ctypes.data, g) == 0
    r, rc = np.zeros(g), np.zeros(g, np.int64)
    c_oracle.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, n, r.ctypes.data,
                                rc.ctypes.data, g)
    exact = ox.exact_sums_f64(keys, vals, g)
    ex = np.array([float(e) for e in exact])
    either = (s.view(np.int64) == ex.view(np.int64)) | (s.view(np.int64) == r.view(np.int64))
    assert bool(np.all(either))
    assert _f64_close(s, r)
    print("max rel err vs exact: hip %.3g, reference row order %.3g"
          % (ox.rel_err_vs_exact(s, exact), ox.rel_err_vs_exact(r, exact)))


def test_f64_specials(hip, c_oracle):
    keys = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4], np.int32)
    vals = np.array([np.inf, 1.0, np.nan, 2.0, np.inf, -np.inf, -0.0, -0.0, -np.inf])
    s, c = np.zeros(5), np.zeros(5, np.int64)
    s0, c0 = np.zeros(5), np.zeros(5, np.int64)
    assert hip.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, 9, s.ctypes.data,
                                  c.ctypes.data, 5) == 0
    c_oracle.nf_group_sum_count(keys.ctypes.data, vals.ctypes.data, 9, s0.ctypes.data,
                                c0.ctypes.data, 5)
    assert np.array_equal(s.view(np.int64)[[0, 3, 4]], s0.view(np.int64)[[0, 3, 4]])
    assert np.isnan(s[1]) and np.isnan(s[2]) and np.isnan(s0[1]) and np.isnan(s0[2])


def test_multi_and_mixed_parity(hip, c_oracle):
    rng = np.random.default_rng(1)
    for g in (3, 500, 70000):
        n = 200_000
        keys = rng.integers(0, g, n).astype(np.int32)
        for ncols in (1, 3, 6):
            vals = rng.uniform(-1e3, 1e3, ncols * n)
            out = []
            for lib in (hip, c_oracle):
                s, c = np.zeros(ncols * g), np.zeros(g, np.int64)
                rc = lib.nf_group_multi_sum_count(keys.ctypes.data, vals.ctypes.data, n, ncols,
                                                  s.ctypes.data, c.ctypes.data, g)
                out.append((rc, s, c))
            assert out[0][0] == out[1][0] == 0
            assert np.array_equal(out[0][2], out[1][2]) and _f64_close(out[0][1], out[1][1])
        for ni, nf in ((2, 1), (1, 1), (0, 2), (5, 0)):
            iv = rng.integers(-100, 100, ni * n).astype(np.int32)
            fv = rng.uniform(0, 1, nf * n)
            out = []
            for lib in (hip, c_oracle):
                si, sf = np.zeros(ni * g, np.int32), np.zeros(nf * g)
                rc = lib.nf_group_mixed_sum_only(keys.ctypes.data, iv.ctypes.data if ni else None,
                                                 fv.ctypes.data if nf else None, n, ni, nf,
                                                 si.ctypes.data if ni else None,
                                                 sf.ctypes.data if nf else None, g)
                out.append((rc, si, sf))
            assert out[0][0] == out[1][0] == 0
            assert np.array_equal(out[0][1], out[1][1]) and _f64_close(out[0][2], out[1][2])
        c1, c2 = np.zeros(g, np.int64), np.zeros(g, np.int64)
        assert hip.nf_group_count_only(keys.ctypes.data, n, c1.ctypes.data, g) == 0
        c_oracle.nf_group_count_only(keys.ctypes.data, n, c2.ctypes.data, g)
        assert np.array_equal(c1, c2)
    # bad key and overflow precedence on the mixed lane
    keys = np.array([0, 0, 9], np.int32)
    iv = np.array([I32MAX, 1, 0], np.int32)
    for lib, want in ((hip, -4), (c_oracle, -4)):
        si = np.zeros(2, np.int32)
        assert lib.nf_group_mixed_sum_only(keys.ctypes.data, iv.ctypes.data, None, 3, 1, 0,
                                           si.ctypes.data, None, 2) == want


def test_owner_parity(hip, c_oracle):
    rng = np.random.default_rng(2)
    n, g = 50_000, 3000
    keys = rng.integers(0, g, n).astype(np.int32)
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    c = rng.uniform(0, 1, n)
    outs = []
    for lib in (hip, c_oracle):
        s1, s2 = np.full(g, 5, np.int32), np.full(g, 6, np.int32)
        s3, cc = np.full(g, 7.0), np.full(g, 8, np.int64)
        for lo, hi in ((0, 1000), (1000, 2500), (2500, 3000)):
            assert lib.nf_group_owner_2i32_1f64(keys.ctypes.data, a.ctypes.data, b.ctypes.data,
                                                c.ctypes.data, n, lo, hi, s1.ctypes.data,
                                                s2.ctypes.data, s3.ctypes.data, cc.ctypes.data,
                                                g) == 0
        outs.append((s1, s2, s3, cc))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][3], outs[1][3]) and _f64_close(outs[0][2], outs[1][2])
    # partial range keeps the other lanes
    s1, s2 = np.full(g, 5, np.int32), np.full(g, 6, np.int32)
    s3, cc = np.full(g, 7.0), np.full(g, 8, np.int64)
    assert hip.nf_group_owner_2i32_1f64(keys.ctypes.data, a.ctypes.data, b.ctypes.data,
                                        c.ctypes.data, n, 10, 20, s1.ctypes.data, s2.ctypes.data,
                                        s3.ctypes.data, cc.ctypes.data, g) == 0
    assert np.all(s1[:10] == 5) and np.all(s1[20:] == 5) and np.all(cc[20:] == 8)
    assert hip.nf_group_owner_2i32_1f64(keys.ctypes.data, a.ctypes.data, b.ctypes.data,
                                        c.ctypes.data, n, 20, 10, s1.ctypes.data, s2.ctypes.data,
                                        s3.ctypes.data, cc.ctypes.data, g) == -1


def test_sorted_run_and_carry_parity(hip, c_oracle):
    rng = np.random.default_rng(9)
    for n in (1, 5, 5000, 300_000):
        keys = np.sort(rng.integers(-500, 500, n)).astype(np.int32)
        vi = rng.integers(-2 ** 40, 2 ** 40, n).astype(np.int64)
        vf = rng.uniform(-1, 1, n)
        for vals, fn in ((vi, "nf_sorted_run_i64"), (vf, "nf_sorted_run_f64")):
            res = []
            for lib in (hip, c_oracle):
                uk, c = np.zeros(n, np.int64), np.zeros(n, np.int64)
                s = np.zeros(n, vals.dtype)
                ng = getattr(lib, fn)(keys.ctypes.data, vals.ctypes.data, n, uk.ctypes.data,
                                      s.ctypes.data, c.ctypes.data)
                res.append((ng, uk[:ng], s[:ng], c[:ng]))
            assert res[0][0] == res[1][0]
            assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][3], res[1][3])
            if vals.dtype == np.int64:
                assert np.array_equal(res[0][2], res[1][2])
            else:
                assert _f64_close(res[0][2], res[1][2])
    keys = np.array([1, 2, 1], np.int32)
    vals = np.zeros(3, np.int64)
    z = np.zeros(3, np.int64)
    assert hip.nf_sorted_run_i64(keys.ctypes.data, vals.ctypes.data, 3, z.ctypes.data,
                                 z.ctypes.data, z.ctypes.data) == -1
    for m in (1, 4096, 4097, 1 << 20):
        cm = rng.integers(0, 3, m).astype(np.int64)
        for sm in (rng.integers(-9, 9, m).astype(np.int64), rng.uniform(0, 1, m)):
            fn = "nf_carry_build_i64" if sm.dtype == np.int64 else "nf_carry_build_f64"
            res = []
            for lib in (hip, c_oracle):
                uk, cc, ss = np.zeros(m, np.int64), np.zeros(m, np.int64), np.zeros(m, sm.dtype)
                ng = getattr(lib, fn)(cm.ctypes.data, sm.ctypes.data, m, uk.ctypes.data,
                                      cc.ctypes.data, ss.ctypes.data)
                res.append((ng, uk[:ng], cc[:ng], ss[:ng]))
            assert res[0][0] == res[1][0]
            for a, b in zip(res[0][1:], res[1][1:]):
                assert a.tobytes() == b.tobytes()


def test_pack_and_fused_parity(hip, c_oracle):
    rng = np.random.default_rng(4)
    n = 100_000
    k1 = rng.integers(0, 300, n).astype(np.int32)
    k2 = rng.integers(0, 100, n).astype(np.int32)
    o1, o2 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    assert hip.nf_pack_i32_direct(k1.ctypes.data, k2.ctypes.data, 100, n, o1.ctypes.data) == 0
    assert c_oracle.nf_pack_i32_direct(k1.ctypes.data, k2.ctypes.data, 100, n, o2.ctypes.data) == 0
    assert np.array_equal(o1, o2)
    bad = k1.copy()
    bad[77] = -1
    assert hip.nf_pack_i32_direct(bad.ctypes.data, k2.ctypes.data, 100, n, o1.ctypes.data) == -2
    big = np.array([2 ** 30], np.int32)
    assert hip.nf_pack_i32_direct(big.ctypes.data, big.ctypes.data, 4, 1, o1.ctypes.data) == -2
    g = 30000
    vi = rng.integers(-1000, 1000, n).astype(np.int32)
    vf = rng.uniform(0, 1, n)
    for vals, fn in ((vi, "nf_pack_sum_count_i32"), (vf, "nf_pack_sum_count_f64")):
        res = []
        for lib in (hip, c_oracle):
            s, c = np.zeros(g, vals.dtype), np.zeros(g, np.int64)
            rc = getattr(lib, fn)(k1.ctypes.data, k2.ctypes.data, 100, vals.ctypes.data, n,
                                  s.ctypes.data, c.ctypes.data, g)
            res.append((rc, s, c))
        assert res[0][0] == res[1][0] == 0
        assert np.array_equal(res[0][2], res[1][2])
        if vals.dtype == np.int32:
            assert np.array_equal(res[0][1], res[1][1])
        else:
            assert _f64_close(res[0][1], res[1][1])
        # packed key beyond g -> -2 on both
        res = []
        for lib in (hip, c_oracle):
            s, c = np.zeros(100, vals.dtype), np.zeros(100, np.int64)
            res.append(getattr(lib, fn)(k1.ctypes.data, k2.ctypes.data, 100, vals.ctypes.data,
                                        n, s.ctypes.data, c.ctypes.data, 100))
        assert res[0] == res[1] == -2


def _pattern(lib, rows, prefix, offs=None):
    data, o = native.build_pattern_buffers(rows)
    if offs is not None:
        o = np.asarray(offs, np.int32)
    n = o.size - 1
    dp = np.frombuffer(data + b"\0", np.uint8)
    pp = np.frombuffer(prefix.encode() + b"\0", np.uint8)
    codes, valid = np.full(n, 77, np.int32), np.full(n, 3, np.uint8)
    w, e = np.zeros(1, np.int32), np.zeros(1, np.int32)
    rc = lib.nf_pattern_encode(dp.ctypes.data, len(data), o.ctypes.data, n, pp.ctypes.data,
                               len(prefix), codes.ctypes.data, valid.ctypes.data, w.ctypes.data,
                               e.ctypes.data)
    return rc, codes, valid, int(w[0]), int(e[0])


def test_pattern_parity(hip, c_oracle):
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
    for case in gold["pattern"]:
        rc, codes, valid, w, _ = _pattern(hip, case["rows"], case["prefix"])
        assert rc == 0 and codes.tolist() == case["codes"] and valid.tolist() == case["valid"]
        assert w == case["width"]
    rows = ["ID12", "ID-5", "X1", "ID", "ID-", "ID0042", "IDab", "ID-2147483648",
            "ID2147483647", "ID00000000000000000000000007", "ID99999999999999999999x"]
    cases = [(rows, "ID", None),
             (rows + ["ID2147483648", "ID5"], "ID", None),
             (["a1", "a2", "a3"], "a", [0, 2, 1, 6]),
             (["a1", "a2"], "a", [0, 2, 9]),
             ([], "a", None)]
    for rws, pf, offs in cases:
        got = _pattern(hip, rws, pf, offs)
        ref = _pattern(c_oracle, rws, pf, offs)
        assert got[0] == ref[0], (rws, got[0], ref[0])
        assert got[1].tolist() == ref[1].tolist() and got[2].tolist() == ref[2].tolist()
        assert got[3:] == ref[3:]
