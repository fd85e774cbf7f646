"""The executor's exactness gates, restated in oracle/executor.py and
mirrored on the product's host side (numfast_amd/native.py), pinned on
their boundaries (CPU only):

* i32_gate_ok  n * max|v| <  2**31 - 1   (native_cpu.py:360-379)
* int_exact_ok n * max|v| <  2**53       (cpu.py:424-446)
* intlane_ok   n * max|v| <  2**63 - 1   (cpu.py:353-378)

Each boundary case sits exactly on or next to its bound, so an off-by-one
(<= for <), an int32 abs that wraps at INT32_MIN, or a float column that
counts would change an answer; the planted mistakes below show it.
"""

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

I32MIN = -(2 ** 31)


def cases():
    # (name, columns, n override for the i32 gate, expected i32 / exact / lane)
    return [
        ("i32 below", [np.array([2 ** 30 - 1, 0], np.int32)], None, (True, True, True)),
        ("i32 at bound", [np.array([2 ** 30, 1], np.int32)], None, (False, True, True)),
        ("i32 INT32_MIN", [np.array([I32MIN], np.int32)], None, (False, True, True)),
        ("i32 n override", [np.array([1000], np.int32)], 2 ** 21, (True, True, True)),
        ("i32 n override over", [np.array([1024], np.int32)], 2 ** 21, (False, True, True)),
        ("exact below", [np.full(2 ** 10, 2 ** 43 - 1, np.int64)], None, (False, True, True)),
        ("exact at bound", [np.full(2 ** 10, 2 ** 43, np.int64)], None, (False, False, True)),
        ("exact negative", [np.full(2 ** 10, -(2 ** 43), np.int64)], None, (False, False, True)),
        ("lane below", [np.array([2 ** 62 - 1, 0], np.int64)], None, (False, False, True)),
        ("lane at bound", [np.array([2 ** 63 - 1], np.int64)], None, (False, False, False)),
        ("lane over", [np.array([2 ** 62, -(2 ** 62)], np.int64)], None, (False, False, False)),
        ("lane INT64_MIN", [np.array([-(2 ** 63)], np.int64)], None, (False, False, False)),
        ("uint64 big", [np.array([2 ** 63], np.uint64)], None, (False, False, False)),
        ("float ignored", [np.array([1e300, -1e300])], None, (None, True, True)),
        ("empty ignored", [np.zeros(0, np.int64)], None, (True, True, True)),
        ("two columns", [np.array([1, 2], np.int32), np.full(4, 2 ** 51, np.int64)], None,
         (None, False, True)),
    ]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_gates_on_their_bounds(case):
    _, cols, n, (i32, exact, lane) = case
    if i32 is not None:
        assert ox.i32_gate_ok(cols[0], n) == i32
        assert native.i32_gate_ok(cols[0], n) == i32
    assert ox.int_exact_ok(cols) == exact
    assert ox.intlane_ok(cols) == lane
    assert native.int_exact_ok(cols) == exact
    assert native.intlane_ok(cols) == lane


def test_mirror_equals_oracle_on_random_columns():
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(0, 50))
        b = int(2 ** rng.integers(0, 63))
        col = rng.integers(-b, b, n, dtype=np.int64) if b > 1 else np.zeros(n, np.int64)
        assert native.int_exact_ok([col]) == ox.int_exact_ok([col])
        assert native.intlane_ok([col]) == ox.intlane_ok([col])
        c32 = (col % (2 ** 31)).astype(np.int32)
        assert native.i32_gate_ok(c32) == ox.i32_gate_ok(c32)


def _mut_le(vcols):  # <= instead of <
    return all(not (np.asarray(v).dtype.kind in "iu" and np.asarray(v).size) or
               np.asarray(v).size * max(abs(int(np.asarray(v).min())),
                                        abs(int(np.asarray(v).max()))) <= 2 ** 53
               for v in vcols)


def _mut_wrap_abs(vals, n=None):  # max|v| taken in int32 (INT32_MIN stays negative)
    vals = np.asarray(vals)
    if vals.size == 0:
        return True
    nn = vals.size if n is None else int(n)
    return nn * int(np.abs(vals).max()) < 2 ** 31 - 1


def _mut_floats_count(vcols):  # float columns counted like integers
    for v in vcols:
        v = np.asarray(v)
        if v.size and v.size * max(abs(float(v.min())), abs(float(v.max()))) >= 2 ** 53:
            return False
    return True


def test_boundary_cases_reject_mutants():
    by = {c[0]: c for c in cases()}
    assert _mut_le(by["exact at bound"][1]) != by["exact at bound"][3][1]
    assert _mut_wrap_abs(by["i32 INT32_MIN"][1][0]) != by["i32 INT32_MIN"][3][0]
    assert _mut_floats_count(by["float ignored"][1]) != by["float ignored"][3][1]
