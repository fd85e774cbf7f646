"""Reentrancy of the C ABI under the reference's thread pool: NumFast calls
the symbols concurrently on disjoint row slices from a ThreadPoolExecutor
with the GIL released by ctypes (src/Drivers/CPU/_lib/cpu.py:521-636,
numfast-native/abi/CENSUS.md:805).  Host NumPy buffers, so every call
stages through its thread's own device scratch."""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("t", [2, 8, 16])
def test_native_mt_slices_match_single_thread(hip, c_oracle, t):
    rng = np.random.default_rng(t)
    n, m = 1 << 20, 1000
    keys = rng.integers(0, m, n).astype(np.int32)
    i32 = rng.integers(-1000, 1001, n).astype(np.int32)
    f64 = rng.uniform(0, 1000, n)
    got = native.native_mt(keys, [i32, f64], m, t)
    assert got is not None
    uk, cnt, (si, sf) = got
    ref_i, ref_f = ox.group_agg(keys, i32), ox.group_agg(keys, f64)
    assert uk.tolist() == ref_i["ukeys"].tolist()
    assert cnt.tolist() == ref_i["counts"].tolist()
    assert np.array_equal(si, ref_i["sums"])
    # the reference merges row-order slice sums in slice order
    bounds = np.linspace(0, n, t + 1).astype(np.int64)
    merged = np.zeros(m)
    for w in range(t):
        a, b = bounds[w], bounds[w + 1]
        s_w, c_w = np.zeros(m), np.zeros(m, np.int64)
        kk, vv = np.ascontiguousarray(keys[a:b]), np.ascontiguousarray(f64[a:b])
        assert c_oracle.nf_group_sum_count(kk.ctypes.data, vv.ctypes.data, kk.size,
                                           s_w.ctypes.data, c_w.ctypes.data, m) == 0
        merged += s_w
    assert np.allclose(sf, merged[uk], rtol=1e-12, atol=0)


def test_mixed_entry_points_from_many_threads(hip):
    """Sorts, group-bys and uniques of different sizes at once from 12
    threads, each checked against the oracle."""
    rng = np.random.default_rng(77)
    jobs = []
    for j in range(24):
        n = int(rng.integers(1, 300_000))
        keys = rng.integers(-(2 ** 20), 2 ** 20, n).astype(np.int32)
        vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
        valid = rng.random(n) > 0.05
        jobs.append((j % 3, keys, vals, valid))

    def run(job):
        kind, keys, vals, valid = job
        if kind == 0:
            return np.array_equal(native.sort_perm_raw(keys, bool(keys[0] & 1)),
                                  ox.sort_step(keys, bool(keys[0] & 1)))
        if kind == 1:
            got, ref = native.group_agg(keys, vals, valid), ox.group_agg(keys, vals, valid)
            return all(got[k].tobytes() == ref[k].astype(got[k].dtype).tobytes()
                       for k in ("ukeys", "counts", "sums", "mins", "maxs", "means"))
        u, inv = native.unique_valid(keys, valid)
        ru, ri = ox.unique_inverse(keys, valid)
        return np.array_equal(u, ru) and np.array_equal(inv, ri)

    with ThreadPoolExecutor(max_workers=12) as ex:
        results = list(ex.map(run, jobs))
    assert all(results)
