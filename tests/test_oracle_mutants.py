"""Planted mistakes: each check that pins the oracle must reject a
plausible wrong restatement (CPU only).

Every mutant here is a copy of an oracle function with one mistake
planted (a dropped term, a wrong order, a wrong sign or rounding).  The
test runs the same acceptance check that pins the real oracle
(tests/test_oracle_golden.py) and asserts that the real function passes
and the mutant fails, so a passing oracle test means something.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import executor as ox

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
I32MIN, I32MAX = -(2 ** 31), 2 ** 31 - 1


# ------------------------------------------------------------------ sort
def sort_golden_ok(sort_perm):
    for case in GOLD["sort"]:
        dt = np.int64 if case.get("dtype") == "int64" else np.int32
        cols = [np.array(c, dtype=dt) for c in case["keys"]]
        if sort_perm(cols, case["desc"], case.get("valid")).tolist() != case["perm"]:
            return False
    return True


def mut_desc_reversed(cols, descending, valid=None):
    """Descending as the plain reverse of the stable ascending order: equal
    keys come out in reverse input order."""
    orig = ox.stable_desc_idx
    try:
        ox.stable_desc_idx = lambda k: np.argsort(k, kind="stable")[::-1]
        return ox.sort_perm(cols, descending, valid)
    finally:
        ox.stable_desc_idx = orig


def mut_null_first(cols, descending, valid=None):
    """NULL rows placed first instead of last."""
    p = ox.sort_perm(cols, descending, valid)
    if valid is None:
        return p
    nv = int(np.count_nonzero(~np.asarray(valid, bool)))
    return np.concatenate([p[p.size - nv:], p[:p.size - nv]])


def mut_first_key_first(cols, descending, valid=None):
    """Stable passes from the first key instead of the last (the last pass
    then dominates: wrong lexicographic priority)."""
    return ox.sort_perm(cols[::-1], list(descending)[::-1], valid)


@pytest.mark.parametrize("mutant", [mut_desc_reversed, mut_null_first, mut_first_key_first])
def test_sort_checks_reject_mutants(mutant):
    assert sort_golden_ok(ox.sort_perm)
    assert not sort_golden_ok(mutant)


# ------------------------------------------------------------------ unique
def unique_ok(unique_fn):
    rng = np.random.default_rng(42)
    for k in (rng.integers(0, 50, 2000).astype(np.int32),
              rng.integers(-2 ** 40, 2 ** 40, 500).astype(np.int64)):
        u, inv = unique_fn(k)
        ru, ri = np.unique(k, return_inverse=True)
        if not (np.array_equal(u, ru) and np.array_equal(inv, ri) and np.array_equal(u[inv], k)):
            return False
    return True


def mut_unique_first_seen(k):
    """Groups in first-occurrence order (hash-table order), not ascending."""
    _, first, inv = np.unique(k, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return k[np.sort(first)], rank[inv].astype(np.int32)


def mut_unique_offby1(k):
    u, inv = ox.unique_inverse(k)
    return u, np.minimum(inv + 1, u.size - 1).astype(np.int32)


@pytest.mark.parametrize("mutant", [mut_unique_first_seen, mut_unique_offby1])
def test_unique_checks_reject_mutants(mutant):
    assert unique_ok(ox.unique_inverse)
    assert not unique_ok(mutant)


# ------------------------------------------------------------------ group-by
def group_golden_ok(group_fn):
    for case in GOLD["group"]:
        vals = np.array(case["vals"])
        vals = vals.astype(np.float64) if vals.dtype.kind == "f" else vals.astype(np.int32)
        r = group_fn(np.array(case["keys"], np.int32), vals, case.get("valid"))
        if r["ukeys"].tolist() != case["ukeys"]:
            return False
        if "sums" in case and r["sums"].tolist() != case["sums"]:
            return False
        if "counts" in case and r["counts"].tolist() != case["counts"]:
            return False
        if "means" in case and not np.allclose(r["means"], case["means"], rtol=1e-12, atol=0):
            return False
    return True


def mut_ignore_validity(keys, vals, valid=None, key_valid=None):
    return ox.group_agg(keys, vals, None, key_valid)


def mut_int_mean(keys, vals, valid=None, key_valid=None):
    r = ox.group_agg(keys, vals, valid, key_valid)
    if r["sums"].dtype.kind == "i":
        r["means"] = (r["sums"] // np.maximum(r["counts"], 1)).astype(np.float64)
    return r


def mut_int32_sums(keys, vals, valid=None, key_valid=None):
    """Integer sums accumulated in int32 (wrapping) instead of int64."""
    r = ox.group_agg(keys, vals, valid, key_valid)
    if r["sums"].dtype.kind == "i":
        r["sums"] = r["sums"].astype(np.int32).astype(np.int64)
    return r


@pytest.mark.parametrize("mutant", [mut_ignore_validity, mut_int_mean, mut_int32_sums])
def test_group_checks_reject_mutants(mutant):
    assert group_golden_ok(ox.group_agg)
    assert not group_golden_ok(mutant)


def test_minmax_dict_check_rejects_swapped():
    """The dict brute force (test_group_agg_minmax_vs_python_dict) rejects
    min and max exchanged."""
    rng = np.random.default_rng(42)
    keys = rng.integers(0, 200, size=3000).astype(np.int32)
    vals = rng.integers(-50, 50, size=3000).astype(np.int32)
    r = ox.group_agg(keys, vals)
    lo = {}
    for k, v in zip(keys.tolist(), vals.tolist()):
        lo[k] = min(lo.get(k, v), v)
    assert r["mins"].tolist() == [lo[k] for k in sorted(lo)]
    assert r["maxs"].tolist() != [lo[k] for k in sorted(lo)]


def test_f64_bincount_check_rejects_pairwise():
    """Row-order f64 sums (dense.rs:5-7, np.bincount) differ in the last
    bits from a pairwise / sorted-order summation on the data the C oracle
    test uses, so that bit-exact check pins the order."""
    rng = np.random.default_rng(0)
    n, g = 1 << 16, 1000
    keys = rng.integers(0, g, n).astype(np.int32)
    vals = rng.uniform(-1000, 1000, n)
    row_order = np.bincount(keys, weights=vals, minlength=g)
    perm = np.argsort(keys, kind="stable")
    sk = keys[perm]
    starts = np.flatnonzero(np.r_[True, sk[1:] != sk[:-1]])
    pairwise = np.add.reduceat(vals[perm], starts)
    assert not np.array_equal(row_order.view(np.int64), pairwise.view(np.int64))


# ------------------------------------------------------------------ return codes
def checked_i32(keys, vals, g, final_only=False, last_wins=False):
    """Python model of the checked int32 dense scatter (scaled.rs:82-105):
    -2 for a key outside [0, g), -4 when a row-order prefix leaves int32;
    the first failing row decides.  Mutants: check only the final sums,
    or let the last failing row decide."""
    sums = [0] * g
    codes = []
    for i, (k, v) in enumerate(zip(keys, vals)):
        if not 0 <= k < g:
            codes.append(-2)
            continue
        sums[k] += v
        if not final_only and not I32MIN <= sums[k] <= I32MAX:
            codes.append(-4)
            sums[k] -= v
    if final_only and any(not I32MIN <= s <= I32MAX for s in sums):
        codes.append(-4)
    if not codes:
        return 0
    return codes[-1] if last_wins else codes[0]


RC_CASES = [  # the hand-derived cases of test_c_rc_prefix_and_precedence
    (([0, 0, 0], [I32MAX, 1, -1], 1), -4),
    (([0, 0, 0], [I32MAX, -1, 1], 1), 0),
    (([0, 0, 5], [I32MAX, 1, 0], 2), -4),
    (([0, 5, 0], [I32MAX, 0, 1], 2), -2),
    (([1, 1], [I32MIN, -1], 2), -4),
]


@pytest.mark.parametrize("flags", [dict(final_only=True), dict(last_wins=True)])
def test_rc_cases_reject_mutants(flags):
    assert all(checked_i32(*a) == want for a, want in RC_CASES)
    assert any(checked_i32(*a, **flags) != want for a, want in RC_CASES)


# ------------------------------------------------------------------ pipeline
def test_pipeline_check_rejects_saturation():
    """derive(v*3) wraps like NumPy int32 (cpu.py:3034-3035, :3074); a
    saturating multiply would give INT32_MAX for 2^30*3."""
    v = np.array([5, -1, 3, 0, 2 ** 30, 7, 3], np.int32)
    w, _ = ox.filter_derive_sort_limit(v, limit=4)
    sat = np.clip(v[v > 0].astype(np.int64) * 3, I32MIN, I32MAX)
    assert w.tolist() == [-(2 ** 30), 9, 9, 15]
    assert np.sort(sat)[:4].tolist() != w.tolist()
