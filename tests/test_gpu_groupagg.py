"""GPU parity of the NULL-aware group-by and unique (additive entry points)
against the oracle's restatement of the reference executor."""

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox

pytestmark = pytest.mark.gpu
I32MIN, I32MAX = -(2 ** 31), 2 ** 31 - 1


def _check(got, ref):
    assert got["ukeys"].tolist() == ref["ukeys"].tolist()
    for k in ("counts", "sums", "mins", "maxs"):
        assert got[k].tobytes() == ref[k].astype(got[k].dtype).tobytes(), k
    assert got["means"].tobytes() == ref["means"].tobytes()


def zipf_keys(rng, n, ids, s):
    """Bounded Zipf(s) over ranks 1..ids by inverse CDF, mapped through a
    seeded permutation of [0, ids)."""
    ranks = np.arange(1, ids + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -s)
    cdf /= cdf[-1]
    r = np.searchsorted(cdf, rng.random(n), side="right")
    perm = rng.permutation(ids).astype(np.int32)
    return perm[np.minimum(r, ids - 1)]


def test_group_agg_golden():
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
    for case in gold["group"]:
        vals = np.array(case["vals"])
        if vals.dtype.kind == "f":
            continue  # float values: covered by the f64 ABI tests
        got = native.group_agg(np.array(case["keys"], np.int32), vals.astype(np.int32),
                               case.get("valid"))
        assert got["ukeys"].tolist() == case["ukeys"], case["src"]
        if "sums" in case:
            assert got["sums"].tolist() == case["sums"], case["src"]
        if "counts" in case:
            assert got["counts"].tolist() == case["counts"], case["src"]
        if "means" in case:
            assert np.allclose(got["means"], case["means"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("n,ids,kind", [
    (1, 1, "uniform"), (7, 3, "uniform"), (100_000, 10, "uniform"),
    (1 << 20, 1 << 16, "zipf"), (1 << 20, 1 << 20, "uniform"),
    (1 << 22, 1 << 24, "zipf"), (1 << 22, 1 << 24, "uniform"), (1 << 21, 1 << 11, "zipf"),
    (1 << 18, 1 << 30, "wide"), (1 << 18, 1 << 31, "extremes"),
])
def test_group_agg_parity(n, ids, kind):
    rng = np.random.default_rng(0)
    if kind == "zipf":
        keys = zipf_keys(rng, n, ids, 1.1)
    elif kind == "wide":
        keys = rng.integers(-ids, ids, n).astype(np.int32)
    elif kind == "extremes":
        keys = rng.choice(np.array([I32MIN, I32MAX, 0, -1, 5], np.int32), n)
    else:
        keys = rng.integers(0, ids, n).astype(np.int32)
    vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
    valid = rng.random(n) > 0.05
    _check(native.group_agg(keys, vals, valid), ox.group_agg(keys, vals, valid))
    _check(native.group_agg(keys, vals), ox.group_agg(keys, vals))
    kv = rng.random(n) > 0.3
    _check(native.group_agg(keys, vals, valid, kv), ox.group_agg(keys, vals, valid, kv))


@pytest.mark.parametrize("n,ids", [(1 << 20, 1 << 16), (1 << 22, 1 << 24)])
def test_group_agg_mixed_value_ranges(n, ids):
    """Hot keys see values both inside and outside [-2^15, 2^15) (the two
    accumulator forms of the hot path), including the exact boundaries."""
    rng = np.random.default_rng(7)
    keys = zipf_keys(rng, n, ids, 1.1)
    vals = rng.integers(-(2 ** 15), 2 ** 15, n).astype(np.int32)
    wide = rng.random(n) < 0.2
    vals[wide] = rng.integers(I32MIN, I32MAX, int(wide.sum()), dtype=np.int64,
                              endpoint=True).astype(np.int32)
    edge = rng.random(n) < 0.02
    vals[edge] = rng.choice(np.array([-32769, -32768, 32767, 32768, I32MIN, I32MAX], np.int32),
                            int(edge.sum()))
    valid = rng.random(n) > 0.05
    _check(native.group_agg(keys, vals, valid), ox.group_agg(keys, vals, valid))


def test_group_agg_extreme_values():
    keys = np.array([3, 3, 3, 4, 4, 9], np.int32)
    vals = np.array([I32MAX, I32MAX, I32MAX, I32MIN, I32MIN, 0], np.int32)
    _check(native.group_agg(keys, vals), ox.group_agg(keys, vals))
    none = native.group_agg(keys, vals, np.zeros(6, bool))
    assert none["ukeys"].size == 0


@pytest.mark.parametrize("n,card,wide", [(1, 1, False), (5000, 37, False),
                                         (1 << 20, 1 << 18, False), (1 << 18, 0, True)])
def test_unique_valid_parity(n, card, wide):
    rng = np.random.default_rng(1)
    if wide:
        keys = rng.integers(I32MIN, I32MAX, n, dtype=np.int64, endpoint=True).astype(np.int32)
    else:
        keys = rng.integers(-card, card + 1, n).astype(np.int32)
    for valid in (None, rng.random(n) > 0.05, np.zeros(n, bool)):
        u, inv = native.unique_valid(keys, valid)
        ru, ri = ox.unique_inverse(keys, valid)
        assert np.array_equal(u, ru) and np.array_equal(inv, ri)


@pytest.mark.parametrize("case", ["top_edge", "bottom_edge", "full_words", "zipf", "unaligned"])
def test_unique_valid_bitmap_edges(case):
    """Dense-span unique: spans that end mid-word, keys at the int32 limits,
    fully set bitmap words, skewed keys, and the unaligned (row-at-a-time)
    path."""
    rng = np.random.default_rng(7)
    n = 100_003
    if case == "top_edge":
        keys = (I32MAX - rng.integers(0, 1001, n)).astype(np.int32)
    elif case == "bottom_edge":
        keys = (I32MIN + rng.integers(0, 77, n)).astype(np.int32)
    elif case == "full_words":
        keys = rng.permutation(np.repeat(np.arange(-4096, 4096, dtype=np.int32), 13))[:n]
    elif case == "zipf":
        keys = ((rng.zipf(1.1, n) % (1 << 20)) * 7919 % (1 << 20) - 5).astype(np.int32)
    else:
        keys = rng.integers(-50_000, 50_000, n).astype(np.int32)
    for valid in (None, rng.random(keys.size) > 0.3):
        ru, ri = ox.unique_inverse(keys, valid)
        if case == "unaligned":
            u, inv = _unique_device_offset(keys, valid)
        else:
            u, inv = native.unique_valid(keys, valid)
        assert np.array_equal(u, ru) and np.array_equal(inv, ri), case


@pytest.mark.parametrize("vec,outside", [(True, False), (False, False), (True, True)])
def test_unique_valid_hot_keys(vec, outside):
    """Skewed keys at a size where the hot keys are resolved in LDS: one hot
    key has only NULL rows (in the hot table, absent from the result; with
    `outside` it also lies below every valid key), one has a single valid
    row, and the rest mix NULL and valid rows."""
    rng = np.random.default_rng(11)
    n = (1 << 22) + 5
    keys = zipf_keys(rng, n, 1 << 20, 1.2) - 3
    hot = np.bincount(keys + 3).argsort()[::-1][:3] - 3
    if outside:
        keys[keys == hot[0]] = keys.min() - 1000
        hot[0] = keys.min()
    valid = rng.random(n) > 0.1
    valid[keys == hot[0]] = False
    rows1 = np.flatnonzero(keys == hot[1])
    valid[rows1] = False
    valid[rows1[-1]] = True
    ru, ri = ox.unique_inverse(keys, valid)
    if vec:
        u, inv = native.unique_valid(keys, valid)
    else:
        u, inv = _unique_device_offset(keys, valid)
    assert hot[0] not in u and hot[1] in u
    assert np.array_equal(u, ru) and np.array_equal(inv, ri)


def _unique_device_offset(keys, valid):
    """nfh_unique_valid_i32 on device buffers that start one element past a
    16-byte boundary, so the row-at-a-time kernels run."""
    import torch
    n = keys.size
    k = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    k[1:] = torch.from_numpy(keys).cuda()
    m = None
    if valid is not None:
        m = torch.empty(n + 1, dtype=torch.uint8, device="cuda")
        m[1:] = torch.from_numpy(valid.astype(np.uint8)).cuda()
    uniq = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    inv = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    ng = native.lib("hip").nfh_unique_valid_i32(
        k[1:].data_ptr(), None if m is None else m[1:].data_ptr(), n, uniq[1:].data_ptr(),
        inv[1:].data_ptr())
    torch.cuda.synchronize()
    assert ng >= 0
    return uniq[1:1 + ng].cpu().numpy(), inv[1:].cpu().numpy()


@pytest.mark.parametrize("dist,masked,vec", [("uniform", False, True), ("uniform", True, True),
                                              ("zipf", True, True), ("edges", True, True),
                                              ("uniform", True, False), ("zipf", False, False),
                                              ("span25", True, True)])
def test_unique_valid_large_span(dist, masked, vec):
    """Dense-span unique with flags larger than an XCD's L2 share (spans
    over 2^21 ids at 2^22+ rows): rows not a multiple of 4, keys at both
    ends of the span, hot keys, NULL rows, and the row-at-a-time path.
    From 2^22 rows the aligned path marks presence by key range (2^20
    offsets per range): spans of 2^24 and 2^25 give 16 and 32 ranges,
    "edges" a third range holding one offset."""
    rng = np.random.default_rng(21)
    n = (1 << 22) + 7
    if dist == "uniform":
        keys = rng.integers(-(1 << 23), 1 << 23, n).astype(np.int32)
    elif dist == "span25":
        keys = rng.integers(-(1 << 24), 1 << 24, n).astype(np.int32)
        keys[:2] = [-(1 << 24), (1 << 24) - 1]
    elif dist == "zipf":
        keys = zipf_keys(rng, n, 1 << 24, 1.1) - 12345
    else:  # span exactly 2^21 + 1, both ends present
        keys = rng.integers(0, (1 << 21) + 1, n).astype(np.int32)
        keys[:2] = [0, 1 << 21]
        keys[-1] = 1 << 21
    valid = rng.random(n) > 0.1 if masked else None
    if valid is not None:
        valid[:2] = True
        valid[-1] = True
    ru, ri = ox.unique_inverse(keys, valid)
    u, inv = native.unique_valid(keys, valid) if vec else _unique_device_offset(keys, valid)
    assert np.array_equal(u, ru) and np.array_equal(inv, ri)
