"""The IR gather (cpu.py:1110-1138 _gather_ref, :3170-3176) on the GPU:
positional take in index order for 1-, 4- and 8-byte values with int32 /
int64 positions, validity taken alike, the contiguous-range case, and the
out-of-range refusal -- against NumPy's take."""

import numpy as np
import pytest

from numfast_amd import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.uint8, np.int32, np.float32, np.int64, np.float64])
@pytest.mark.parametrize("it", [np.int32, np.int64])
@pytest.mark.parametrize("n,k", [(1, 1), (1000, 7), (100_003, 1 << 20), (1 << 22, 1000)])
def test_gather_matches_take(hip, dt, it, n, k):
    rng = np.random.default_rng(n + k)
    vals = rng.integers(0, 250, n).astype(dt) if dt == np.uint8 else \
        (rng.standard_normal(n) * 1e6).astype(dt)
    idx = rng.integers(0, n, k).astype(it)
    valid = rng.random(n) > 0.2
    out, ov = native.gather(vals, idx, valid)
    assert out.tobytes() == vals[idx].tobytes()
    assert np.array_equal(ov, valid[idx])


def test_gather_contiguous_range_and_sort_perm(hip):
    """The reference's fast path (a contiguous range) gives the same values;
    gathering by a sort permutation gives the sorted column."""
    vals = np.arange(100, 200, dtype=np.int64)
    out, _ = native.gather(vals, np.arange(10, 30, dtype=np.int32))
    assert out.tolist() == list(range(110, 130))
    k = np.random.default_rng(2).integers(-50, 50, 5000).astype(np.int32)
    perm = native.sort_perm_raw(k, False)
    out, _ = native.gather(k, perm)
    assert np.array_equal(out, np.sort(k, kind="stable"))


def test_gather_refusals(hip):
    vals = np.arange(10, dtype=np.int32)
    for bad in ([0, 10], [-1], [3, 2, 100]):
        with pytest.raises(IndexError):
            native.gather(vals, np.array(bad, np.int64))
    with pytest.raises(TypeError):
        native.gather(vals, np.array([0.0, 1.0]))
    out, _ = native.gather(vals, np.zeros(0, np.int32))
    assert out.size == 0
