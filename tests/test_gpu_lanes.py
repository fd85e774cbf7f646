"""The host-side lanes of the boundary on the GPU (numfast_amd/native.py,
the mirror of src/Drivers/CPU/_lib/native_cpu.py and the executor pieces
of cpu.py that call it):

* the multi-key, NULL-last stable sort composed of HIP argsort steps
  (cpu.py:1174-1225) on the reference's sort vectors
  (tests/fast/test_ops_sort.py:40-197) and its random lexicographic case;
* sort_perm_step's monotone shortcuts "ident" / "rev" (native_cpu.py:
  1602-1652), which answer without a library call exactly where the
  reference's do;
* return code -> exception mapping of the int lanes (native_cpu.py:394-445)
  and mixed_sum_count degrading to None (native_cpu.py:448-492,
  tests/fast/test_ops_panic_safety.py:99-110);
* the executor's single-thread native route (cpu.py:461-518) with its
  exactness gates, against the oracle's group results.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


@pytest.mark.parametrize("case", GOLD["sort"], ids=lambda c: c["src"])
def test_multikey_sort_golden(hip, case):
    dt = np.int64 if case.get("dtype") == "int64" else np.int32
    cols = [np.array(c, dtype=dt) for c in case["keys"]]
    perm = native.sort_perm(cols, case["desc"], case.get("valid"))
    assert perm.tolist() == case["perm"]


def test_multikey_sort_random_lexicographic(hip):
    """tests/fast/test_ops_sort.py:73-87: n = 10k, seed 42, vs np.lexsort."""
    rng = np.random.default_rng(42)
    n = 10_000
    k1 = rng.integers(0, 50, size=n).astype(np.int32)
    k2 = rng.integers(0, 50, size=n).astype(np.int32)
    perm = native.sort_perm([k1, k2], [False, False])
    assert perm.tolist() == np.lexsort((k2, k1)).tolist()


@pytest.mark.parametrize("n", [1000, 200_003])
def test_multikey_sort_mixed_directions_nulls(hip, n):
    rng = np.random.default_rng(n)
    cols = [rng.integers(-5, 5, n).astype(np.int32),
            rng.integers(-(2 ** 40), 2 ** 40, n).astype(np.int64),
            rng.integers(0, 3, n).astype(np.int32)]
    desc = [True, False, True]
    valid = rng.random(n) > 0.1
    assert np.array_equal(native.sort_perm(cols, desc, valid), ox.sort_perm(cols, desc, valid))


def test_sort_invalid_rows_last_both_directions(hip):
    """tests/fast/test_ops_sort.py:90-100."""
    vals = np.array([3, 1, 2, 0], np.int32)
    validity = [1, 0, 1, 0]
    for desc in (False, True):
        got = native.sort_perm([vals], [desc], validity).tolist()
        assert got[2:] == [1, 3]
        assert [int(vals[i]) for i in got[:2]] == ([2, 3] if not desc else [3, 2])


@pytest.mark.parametrize("desc", [False, True])
def test_sort_step_shortcuts(hip, desc):
    """Each monotone shape gets the reference's backend, and every step is
    the stable permutation of its direction."""
    n = 1000
    up = np.arange(n, dtype=np.int32)
    cases = {
        "up": (up, "ident" if not desc else "rev"),
        "down": (up[::-1].copy(), "rev" if not desc else "ident"),
        "equal": (np.full(n, 7, np.int32), "ident"),
        # non-strict cross-monotone: the reference hands it to NumPy, this
        # mirror to the radix kernel; the permutation is the same
        "down_dups": (np.repeat(np.arange(n // 2, dtype=np.int32)[::-1], 2),
                      "native" if not desc else "ident"),
        "up_dups": (np.repeat(np.arange(n // 2, dtype=np.int32), 2),
                    "ident" if not desc else "native"),
        "random": (np.random.default_rng(1).integers(0, 50, n).astype(np.int32), "native"),
        "one": (np.array([5], np.int32), "ident"),
    }
    for name, (k, backend) in cases.items():
        step, got_backend = native.sort_perm_step(k, desc)
        assert got_backend == backend, name
        assert np.array_equal(step, ox.sort_step(k, desc)), name


def test_mixed_sum_count_degrades_and_serves(hip):
    """tests/fast/test_ops_panic_safety.py:99-110 (golden 'mixed' cases)."""
    for case in GOLD["mixed"]:
        keys = np.array(case["keys"], np.int32)
        col = np.array(case["col"], np.int64)
        got = native.mixed_sum_count(keys, [col], case["g"])
        if "result" in case and case["result"] is None:
            assert got is None, case["src"]
            continue
        assert got is not None, case["src"]
        sums, counts = got
        assert counts.tolist() == case["counts"], case["src"]
        assert int(sums[0][7]) == case["sum7"], case["src"]


def test_rc_maps_to_exceptions(hip):
    """native_cpu.py:394-445: -4 -> OverflowError, -2 -> RuntimeError."""
    k = np.zeros(2, np.int32)
    with pytest.raises(OverflowError):
        native.sum_count_i32(k, np.array([2 ** 31 - 1, 1], np.int32), 1)
    with pytest.raises(OverflowError):
        native.sum_count_i64(k, np.array([2 ** 63 - 1, 1], np.int64), 1)
    with pytest.raises(RuntimeError, match="dense key outside"):
        native.sum_count_i64(np.array([0, 3], np.int32), np.array([1, 1], np.int64), 2)
    # in range: exact sums
    s, c = native.sum_count_i32(np.array([0, 1, 1], np.int32), np.array([5, -2, 9], np.int32), 2)
    assert s.tolist() == [5, 7] and c.tolist() == [1, 2]


def test_native_st_route(hip):
    """cpu.py:461-518 over int32 codes: integer, float and mixed columns
    against the oracle's groups; None exactly where the executor falls back
    (a column beyond int64 through the lanes and beyond 2**53 in float64)."""
    rng = np.random.default_rng(9)
    n, m = 50_000, 300
    keys = rng.integers(0, m, n).astype(np.int32)
    keys[keys == 17] = 18  # an absent code
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    i64 = rng.integers(-(2 ** 40), 2 ** 40, n).astype(np.int64)
    f64 = rng.uniform(-1, 1, n)
    got = native.native_st(keys, [i32, i64, f64], m)
    assert got is not None
    uk, cnt, sums = got
    for col, s in zip((i32, i64, f64), sums):
        ref = ox.group_agg(keys, col)
        assert uk.tolist() == ref["ukeys"].tolist() and cnt.tolist() == ref["counts"].tolist()
        if col.dtype.kind == "f":
            assert np.allclose(s, ref["sums"], rtol=1e-12, atol=0)
        else:
            assert np.array_equal(s.astype(np.int64), ref["sums"])
    # int64 sums that overflow the checked lane and fail the 2**53 gate
    big = np.full(n, 2 ** 62, np.int64)
    assert not ox.int_exact_ok([big])
    assert native.native_st(keys, [big], m) is None
