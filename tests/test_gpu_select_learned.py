"""GPU parity: selection / masks (nf_select_*, nf_mask_*), the filter ->
derive -> sort -> limit pipeline, the learned-CDF sort and the training
kernels' gradients, against the oracle."""

import json
import zlib
from pathlib import Path

import numpy as np
import pytest

from numfast_amd import native
from oracle import executor as ox
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
I32MIN, I32MAX = -(2 ** 31), 2 ** 31 - 1


# ------------------------------------------------------------------ select
def test_select_golden(hip):
    for case in GOLD["filter"]:
        v = np.array(case["vals"], np.int32)
        m = np.array(case["mask"], np.uint8)
        assert native.select_scatter(v, m).tolist() == case["out"], case["src"]
        assert native.select_count(m) == len(case["out"])
    g = GOLD["masks"]
    assert native.mask_op("and", g["hi"], g["lo"]).tolist() == g["and"]
    assert native.mask_op("or", g["hi"], g["lo"]).tolist() == g["or"]
    assert native.mask_op("not", g["hi"]).tolist() == g["not_hi"]


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.float32, np.float64, np.uint8])
@pytest.mark.parametrize("n", [1, 3, 4, 5, 4095, 4096, 4097, 100003, (1 << 20) + 7])
def test_select_parity(hip, c_oracle, dt, n):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 250, n).astype(dt)
    for m in (rng.integers(0, 3, n).astype(np.uint8), np.zeros(n, np.uint8),
              np.full(n, 7, np.uint8)):
        got = native.select_scatter(v, m)
        assert np.array_equal(got, v[m != 0])
        assert native.select_count(m) == int(np.count_nonzero(m))


def test_select_unaligned_device_pointers(hip):
    """Device views at odd offsets take the scalar path."""
    import torch
    n = 100_001
    rng = np.random.default_rng(5)
    v = torch.from_numpy(rng.integers(-9, 9, n + 3).astype(np.int32)).cuda()
    m = torch.from_numpy(rng.integers(0, 2, n + 3).astype(np.uint8)).cuda()
    out = torch.zeros(n, dtype=torch.int32, device="cuda")
    L = native.lib("select")
    r = L.nf_select_scatter_i32(v[1:].data_ptr(), m[3:].data_ptr(), n, out.data_ptr())
    ref = v[1:n + 1].cpu().numpy()[m[3:n + 3].cpu().numpy() != 0]
    assert r == ref.size and np.array_equal(out[:r].cpu().numpy(), ref)


def test_masks_random(hip, c_oracle):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 3, 70001).astype(np.uint8)
    b = rng.integers(0, 3, 70001).astype(np.uint8)
    assert np.array_equal(native.mask_op("and", a, b), ((a != 0) & (b != 0)).astype(np.uint8))
    assert np.array_equal(native.mask_op("or", a, b), ((a != 0) | (b != 0)).astype(np.uint8))
    assert np.array_equal(native.mask_op("not", a), (a == 0).astype(np.uint8))


def test_select_null_pointer_codes(hip):
    L = native.lib("select")
    assert L.nf_select_count(None, 5) == -2
    assert L.nf_select_scatter_i32(None, None, 5, None) == -2
    assert L.nf_select_count(None, 0) == 0
    assert L.nf_mask_and(None, None, 3, None) == -1


# ------------------------------------------------------------------ learned sort
def _dist(kind, n, rng):
    if kind == "uniform":
        return rng.integers(I32MIN, I32MAX, n, dtype=np.int64, endpoint=True).astype(np.int32)
    if kind == "swaps":
        a = np.sort(rng.integers(I32MIN, I32MAX, n, dtype=np.int64, endpoint=True)).astype(
            np.int32)
        pos = rng.choice(n, n // 10, replace=False)
        a[pos] = a[rng.permutation(pos)]
        return a
    if kind == "distinct1024":
        vals = rng.integers(I32MIN, I32MAX, 1024, dtype=np.int64, endpoint=True).astype(np.int32)
        return vals[rng.integers(0, 1024, n)]
    if kind in ("clustered", "narrow"):
        c = rng.integers(-(10 ** 9), 10 ** 9, 8)
        sd = 1e7 if kind == "clustered" else 2000.0
        return np.clip(c[rng.integers(0, 8, n)] + rng.normal(0, sd, n), I32MIN, I32MAX).astype(
            np.int32)
    if kind == "bumps":
        c = rng.integers(-(10 ** 9), 10 ** 9, 64)
        wgt = rng.random(64)
        pick = rng.choice(64, n, p=wgt / wgt.sum())
        return np.clip(c[pick] + rng.normal(0, 3e7, n), I32MIN, I32MAX).astype(np.int32)
    if kind == "exponential":
        return np.minimum(rng.exponential(5e7, n), I32MAX).astype(np.int32)
    if kind == "lognormal":
        return np.minimum(rng.lognormal(3, 2.5, n), I32MAX).astype(np.int32)
    if kind == "extremes":
        return rng.choice(np.array([I32MIN, I32MAX, 0, -1, 1], np.int32), n)
    if kind == "equal":
        return np.full(n, 42, np.int32)
    raise ValueError(kind)


KINDS = ["uniform", "swaps", "distinct1024", "clustered", "narrow", "bumps", "exponential",
         "lognormal", "extremes", "equal"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [1000, 1 << 16, (1 << 20) + 13])
def test_learned_sort_equals_stable_argsort(hip, kind, n):
    rng = np.random.default_rng(zlib.crc32(("%s/%d" % (kind, n)).encode()))
    k = _dist(kind, n, rng)
    model, _ = native.cdf_train(k, steps=50)
    perm, stats = native.sort_perm_learned(k, model)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))


@pytest.mark.parametrize("kind", ["uniform", "swaps", "distinct1024", "equal"])
def test_learned_sort_hint_changes_speed_only(hip, kind):
    """The model's sortedness hint (set by training from its row-order
    sample) picks the code passes' run ranking; both settings give the
    stable argsort on the same keys, and training sets it only for nearly
    sorted columns."""
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    n = (1 << 20) + 77
    k = _dist(kind, n, rng)
    model, _ = native.cdf_train(k, steps=10)
    assert model.hints == (1 if kind in ("swaps", "equal") else 0)
    ref = np.argsort(k, kind="stable")
    for h in (0, 1):
        model.hints = h
        perm, _ = native.sort_perm_learned(k, model)
        assert np.array_equal(perm, ref), h


@pytest.mark.parametrize("kind", ["narrow", "lognormal", "exponential"])
def test_learned_sort_mismatched_model_uses_fixup(hip, kind):
    """A model trained on other data puts many keys in few buckets: the
    over-capacity buckets go through the copy / plain-sort fix-up and the
    result is still the stable argsort."""
    rng = np.random.default_rng(9)
    n = 1 << 20
    model, _ = native.cdf_train(_dist("uniform", n, rng), steps=20)
    k = _dist(kind, n, rng)
    perm, stats = native.sort_perm_learned(k, model)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))
    assert stats[2] > 0


def test_learned_sort_large_bucket_paths(hip):
    """Over-capacity buckets of both sizes: one of about 10^4 unequal keys
    (sorted by one workgroup in global memory), one of 3 * 10^5 unequal keys
    (the plain-sort fix-up) and one of equal keys (copied)."""
    rng = np.random.default_rng(5)
    base = _dist("uniform", 1 << 20, rng)
    model, _ = native.cdf_train(base, steps=20)
    k = np.concatenate([base, (np.arange(300_000) + 12_345).astype(np.int32),
                        rng.integers(1 << 30, (1 << 30) + 1_000_000, 8000).astype(np.int32),
                        np.full(20_000, -(1 << 29), np.int32)])
    k = k[rng.permutation(k.size)]
    perm, stats = native.sort_perm_learned(k, model)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))
    assert stats[3] >= 1 and stats[2] >= stats[3] + 2, stats


def test_learned_sort_garbage_model_is_exact(hip):
    rng = np.random.default_rng(2)
    k = _dist("uniform", 1 << 18, rng)
    model, _ = native.cdf_train(k, steps=10)
    for bad in (np.nan, np.inf, -1e30):
        for i in range(native.MLP_P):
            model.p[i] = bad
        perm, _ = native.sort_perm_learned(k, model)
        assert np.array_equal(perm, np.argsort(k, kind="stable"))


def _linear_model(k):
    """The method without the trained network: a CDF linear in the key word
    between the sample min and max (f(x) = x)."""
    model, _ = native.cdf_train(k, steps=0, linear_start=True)
    return model


@pytest.mark.parametrize("kind", ["clustered", "exponential"])
def test_trained_network_gain_over_linear_cdf(hip, kind):
    """What the network adds: on skewed keys a linear CDF (the method
    without it) overfills buckets; the trained network fills them more
    evenly: a smaller largest bucket and no more over-capacity buckets (each
    of which costs a fix-up)."""
    rng = np.random.default_rng(4)
    n = 1 << 21
    k = _dist(kind, n, rng)
    trained, loss = native.cdf_train(k, steps=200)
    p1, s1 = native.sort_perm_learned(k, trained)
    p0, s0 = native.sort_perm_learned(k, _linear_model(k))
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(p1, ref) and np.array_equal(p0, ref)
    print(kind, "trained: buckets %d largest %d over %d fixed %d loss %.2e" % (
        *[int(v) for v in s1], loss), "| linear: buckets %d largest %d over %d fixed %d" % tuple(
        int(v) for v in s0))
    assert s1[1] < s0[1]
    assert s1[2] <= s0[2]


@pytest.mark.parametrize("kind", ["clustered", "exponential", "bumps"])
def test_gradient_training_gain_on_buckets(hip, kind):
    """What the gradient training itself adds, on the sort it serves: from
    the linear-CDF start, 300 device steps against none, compared on the
    training loss (the mean squared CDF error it minimises) and on the
    learned sort's buckets; both sorts stay exact.  Only the loss is
    asserted: the squared CDF error it minimises does not bound the largest
    bucket or the count of buckets over the LDS capacity, which move either
    way (the figures are printed)."""
    rng = np.random.default_rng(8)
    n = 1 << 21
    k = _dist(kind, n, rng)
    m0, l0 = native.cdf_train(k, steps=0, linear_start=True)
    m1, l1 = native.cdf_train(k, steps=300, linear_start=True)
    p0, s0 = native.sort_perm_learned(k, m0)
    p1, s1 = native.sort_perm_learned(k, m1)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(p0, ref) and np.array_equal(p1, ref)
    print(kind, "linear start, 0 steps: largest %d over %d loss %.2e | 300 steps: largest %d "
          "over %d loss %.2e" % (s0[1], s0[2], l0, s1[1], s1[2], l1))
    assert l1 < 0.7 * l0


@pytest.mark.parametrize("kind", ["clustered", "exponential", "bumps", "uniform"])
def test_trained_vs_greedy_start(hip, kind):
    """The sort's own model: the greedy 32-piece fit (0 steps) against the
    same start after 200 device steps.  Backtracking never accepts a worse
    fit; the buckets it gives are reported (on these columns the greedy fit
    is already at fp32 resolution, so the steps change little)."""
    rng = np.random.default_rng(9)
    n = 1 << 21
    k = _dist(kind, n, rng)
    m0, l0 = native.cdf_train(k, steps=0)
    m1, l1 = native.cdf_train(k, steps=200)
    _, s0 = native.sort_perm_learned(k, m0)
    p1, s1 = native.sort_perm_learned(k, m1)
    assert np.array_equal(p1, np.argsort(k, kind="stable"))
    print(kind, "greedy start, 0 steps: largest %d over %d loss %.3e | 200 steps: largest %d "
          "over %d loss %.3e" % (s0[1], s0[2], l0, s1[1], s1[2], l1))
    assert l1 <= l0
    assert s1[2] <= s0[2] + 1


def test_training_reduces_loss(hip):
    """The device training lowers the loss: from the linear-CDF start it has
    to place the slope changes itself (strict decrease); from the greedy
    piecewise-linear start used by the sort it never returns a worse model
    (backtracking), even where no fp32 step improves on it."""
    rng = np.random.default_rng(6)
    for kind in ("bumps", "clustered"):
        k = _dist(kind, 1 << 20, rng)
        _, u0 = native.cdf_train(k, steps=0, linear_start=True)
        _, u1 = native.cdf_train(k, steps=300, linear_start=True)
        _, g0 = native.cdf_train(k, steps=0)
        _, g1 = native.cdf_train(k, steps=300)
        print(kind, "linear start %.3e -> %.3e, greedy start %.3e -> %.3e" % (u0, u1, g0, g1))
        assert u1 < 0.7 * u0
        assert g1 <= g0


# ------------------------------------------------------------------ gradients
@pytest.mark.parametrize("S", [1, 127, 1000, 70001])
def test_mlp_gradient_matches_oracle(hip, S):
    rng = np.random.default_rng(S)
    p = rng.normal(0, 0.5, omlp.P).astype(np.float32)
    x = rng.random(S).astype(np.float32)
    y = np.sort(rng.random(S)).astype(np.float32)
    g, loss, pred = native.mlp_grad(p, x, y)
    ref = omlp.grad(p, x, y)
    scale = max(1e-12, np.abs(ref).max())
    assert np.max(np.abs(g - ref)) <= 2e-4 * scale, np.max(np.abs(g - ref)) / scale
    assert abs(loss - omlp.loss(p, x, y)) <= 1e-4 * max(1e-12, omlp.loss(p, x, y))
    assert np.allclose(pred, omlp.forward(p, x), rtol=1e-4, atol=1e-5)


# ------------------------------------------------------------------ pipeline
@pytest.mark.parametrize("kind", ["uniform", "swaps", "distinct1024"])
@pytest.mark.parametrize("n,limit", [(1000, 1000), (1 << 16, 1000), ((1 << 20) + 3, 1000),
                                     (5000, 1 << 20)])
def test_pipeline_parity(hip, kind, n, limit):
    rng = np.random.default_rng(n + limit)
    v = _dist(kind, n, rng)
    rw, rr = ox.filter_derive_sort_limit(v, limit=limit, mul=3)
    w, rows = native.filter_derive_sort_limit(v, limit=limit)
    assert np.array_equal(w, rw) and np.array_equal(rows, rr)
    derived = (v[v > 0].astype(np.int64) * 3).astype(np.int32)
    model, _ = native.cdf_train(derived if derived.size else v, steps=30)
    w2, rows2 = native.filter_derive_sort_limit(v, limit=limit, model=model)
    assert np.array_equal(w2, rw) and np.array_equal(rows2, rr)


def test_pipeline_edges(hip):
    assert native.filter_derive_sort_limit(np.array([-1, 0, -5], np.int32))[0].size == 0
    w, r = native.filter_derive_sort_limit(np.array([2 ** 30, 1], np.int32), limit=5)
    assert w.tolist() == [-(2 ** 30), 3] and r.tolist() == [0, 1]
