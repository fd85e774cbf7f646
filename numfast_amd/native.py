"""ctypes binding of libnumfast_hip.so -- the host mirror of NumFast's
native boundary (src/Drivers/CPU/_lib/native_cpu.py).

Same loader contract as the reference:

* the library path comes from ``NUMFAST_NATIVE_DLL`` (default: the in-tree
  ``numfast_amd/libnumfast_hip.so``), ``NUMFAST_NATIVE_DISABLE=1`` disables
  it (native_cpu.py:93-143);
* the 8 eager symbols get their argtypes at load time and all must exist
  (native_cpu.py:125-140); each optional group (variant, mixed, owner,
  fused, unique, sort) is bound on first use and fails independently
  (native_cpu.py:324-352, :503-532, :548-574, :621-651, :1399-1431,
  :1546-1576);
* return codes map to the same exceptions (-4 OverflowError, -2
  RuntimeError "dense key outside", anything else ``assert rc == 0``).

One deliberate difference: there is NO NumPy fallback.  When the library
is missing or disabled the wrappers raise :class:`NativeUnavailable`; the
HIP path is the only compute path of this package.

The wrappers accept NumPy arrays (host memory: the library stages them
through device memory) and, for the ``*_device`` helpers, torch tensors
already resident on the GPU (pointers are passed through untouched).
"""

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
DEFAULT_LIB = str(_HERE / "libnumfast_hip.so")

_vp, _sz, _i32, _i64, _u8, _u32 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32,
                                   ctypes.c_int64, ctypes.c_uint8, ctypes.c_uint32)

# (name, argtypes, restype) -- native_cpu.py:125-140 (the eager 8).
EAGER = [
    ("nf_group_sum_count", [_vp] * 2 + [_sz] + [_vp] * 2 + [_sz], _i32),
    ("nf_group_multi_sum_count", [_vp] * 2 + [_sz] * 2 + [_vp] * 2 + [_sz], _i32),
    ("nf_pack_i32_direct", [_vp] * 2 + [_i32, _sz, _vp], _i32),
    ("nf_pattern_encode", [_vp, _sz] * 3 + [_vp] * 4, _i32),
    ("nf_sorted_run_i64", [_vp] * 2 + [_sz] + [_vp] * 3, _i64),
    ("nf_sorted_run_f64", [_vp] * 2 + [_sz] + [_vp] * 3, _i64),
    ("nf_carry_build_i64", [_vp] * 2 + [_sz] + [_vp] * 3, _i64),
    ("nf_carry_build_f64", [_vp] * 2 + [_sz] + [_vp] * 3, _i64),
]

# Optional groups, each all-or-nothing (native_cpu.py, lines in the doc).
GROUPS = {
    "variant": [
        ("nf_group_variant_i64", [_vp, _vp, _sz, _u32, _vp, _vp, _sz], _i32),
        ("nf_group_variant_i32", [_vp, _vp, _sz, _u32, _vp, _vp, _sz], _i32),
        # lib.rs:539 -- exported by the reference, not bound by its Python
        ("nf_group_variant_f64", [_vp, _vp, _sz, _u32, _vp, _vp, _sz], _i32),
    ],
    "mixed": [
        ("nf_group_mixed_sum_only", [_vp, _vp, _vp, _sz, _sz, _sz, _vp, _vp, _sz], _i32),
        ("nf_group_count_only", [_vp, _sz, _vp, _sz], _i32),
    ],
    "owner": [
        ("nf_group_owner_2i32_1f64", [_vp] * 4 + [_sz] * 3 + [_vp] * 4 + [_sz], _i32),
    ],
    "fused": [
        ("nf_pack_sum_count_i32", [_vp, _vp, _i32, _vp, _sz, _vp, _vp, _sz], _i32),
        ("nf_pack_sum_count_f64", [_vp, _vp, _i32, _vp, _sz, _vp, _vp, _sz], _i32),
    ],
    "unique": [
        ("nf_unique_inverse_i32", [_vp, _sz] + [_vp] * 6, _i64),
        ("nf_unique_inverse_i64", [_vp, _sz] + [_vp] * 6, _i64),
    ],
    "sort": [
        ("nf_sort_perm_i32", [_vp, _sz, _u8] + [_vp] * 4, _i32),
        ("nf_sort_perm_i64", [_vp, _sz, _u8] + [_vp] * 4, _i32),
    ],
    # native_cpu.py:164-200 (_req_select).  The reference declares the
    # scatter restype as c_int32 although the symbols return i64
    # (lib.rs:654-659); the binding keeps that quirk (exact for m < 2**31).
    "select": [
        ("nf_select_count", [_vp, _sz], _i64),
        ("nf_select_scatter_i32", [_vp, _vp, _sz, _vp], _i32),
        ("nf_select_scatter_f32", [_vp, _vp, _sz, _vp], _i32),
        ("nf_select_scatter_f64", [_vp, _vp, _sz, _vp], _i32),
        ("nf_select_scatter_i64", [_vp, _vp, _sz, _vp], _i32),
        ("nf_select_scatter_u8", [_vp, _vp, _sz, _vp], _i32),
        ("nf_mask_and", [_vp, _vp, _sz, _vp], _i32),
        ("nf_mask_or", [_vp, _vp, _sz, _vp], _i32),
        ("nf_mask_not", [_vp, _sz, _vp], _i32),
    ],
    # src/Drivers/GroupedHashMT/_lib/grouped_native_hash.py:95-128
    # (partitioned pair-hash for grouped COUNT DISTINCT, lib.rs:2266-2470).
    "ghash": [
        ("nf_ghash_fp_i64", [_vp, _vp, _sz, _vp, _i64, _sz], _i32),
        ("nf_ghash_count_i64", [_vp, _sz, _vp, _sz, _sz, _sz], _i32),
        ("nf_ghash_scatter_i64", [_vp, _vp, _vp, _sz, _vp, _sz, _sz, _vp, _vp, _sz], _i32),
        ("nf_ghash_pins_i64", [_vp] * 8 + [_sz, _sz], _i32),
        ("nf_ghash_occ_count", [_vp, _vp, _sz, _vp, _sz], _i32),
        ("nf_ghash_occ_fill", [_vp] * 5 + [_sz, _sz], _i32),
    ],
    # Additive: learned-CDF sort, its training, the sort pipeline.
    "learned": [
        ("nfh_cdf_train_i32", [_vp, _sz, _u32, _u32, _vp, _vp], _i32),
        ("nfh_sort_perm_learned_i32", [_vp, _sz, _vp, _vp, _vp], _i32),
        ("nfh_mlp_grad", [_vp, _vp, _vp, _sz, _vp, _vp, _vp], _i32),
        ("nfh_filter_derive_sort_limit_i32", [_vp, _sz, _i32, _sz, _vp, _vp, _vp], _i64),
    ],
    # Additive: learned group-by strategy selector.
    "selector": [
        ("nfh_groupby_features_i32", [_vp, _sz, _vp], _i32),
        ("nfh_selector_grad", [_vp, _vp, _vp, _sz, _vp, _vp], _i32),
        ("nfh_selector_train", [_vp, _vp, _sz, _u32, ctypes.c_float, _vp, _vp], _i32),
        ("nfh_selector_predict", [_vp, _vp, _sz, _vp], _i32),
        ("nfh_selector_grad_cost", [_vp, _vp, _vp, _sz, _vp, _vp], _i32),
        ("nfh_selector_train_cost", [_vp, _vp, _sz, _u32, ctypes.c_float, _vp, _vp], _i32),
        ("nfh_selector_predict_cost", [_vp, _vp, _sz, _vp], _i32),
        ("nfh_set_groupby_strategy", [_i32], _i32),
        ("nfh_groupby_last_strategy", [], _i32),
    ],
    # Additive: batched group-by plans (hipGraph).
    "batch": [
        ("nfh_group_batch_plan", [_vp] * 4 + [_sz] + [_vp] * 4 + [_sz, _vp], _vp),
        ("nfh_group_batch_run", [_vp, _i32], _i64),
        ("nfh_group_batch_info", [_vp, _vp], _i32),
        ("nfh_group_batch_info2", [_vp, _vp], _i32),
        ("nfh_group_batch_free", [_vp], None),
    ],
    # Additive surface of this library (no reference counterpart).
    "hip": [
        ("nfh_set_stream", [_vp], _i32),
        ("nfh_device_ok", [], _i32),
        ("nfh_abi_version", [], _i32),
        ("nfh_group_agg_i32", [_vp] * 4 + [_sz] + [_vp] * 6 + [_sz], _i64),
        ("nfh_group_agg_f64", [_vp] * 4 + [_sz] + [_vp] * 6 + [_sz], _i64),
        ("nfh_group_agg_i64", [_vp] * 4 + [_sz] + [_vp] * 6 + [_sz], _i64),
        ("nfh_unique_valid_i32", [_vp, _vp, _sz, _vp, _vp], _i64),
        ("nfh_sort_perm_i32_noalloc", [_vp, _sz, _u8] + [_vp] * 4, _i32),
        ("nfh_gather", [_vp, _sz, _sz, _vp, _sz, _sz, _vp, _vp, _vp], _i32),
        ("nfh_profile_enable", [_i32], _i32),
        ("nfh_profile_reset", [], _i32),
        ("nfh_profile_get", [ctypes.c_char_p, _vp, _vp], _i32),
        ("nfh_profile_bytes", [ctypes.c_char_p, _vp], _i32),
        ("nfh_profile_names", [_vp, _sz], _i64),
    ],
}


class NativeUnavailable(RuntimeError):
    """The HIP library is missing, disabled or lacks a symbol group."""


def bind(lib, groups=None):
    """Set argtypes/restype on a loaded CDLL; returns the set of bound groups.

    Raises AttributeError when an eager symbol is missing (the reference's
    all-or-nothing rule).  Optional groups that are incomplete are skipped.
    """
    for name, args, res in EAGER:
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    ok = set()
    for g, syms in GROUPS.items():
        if groups is not None and g not in groups:
            continue
        try:
            for name, args, res in syms:
                f = getattr(lib, name)
                f.argtypes = args
                f.restype = res
            ok.add(g)
        except AttributeError:
            pass
    return ok


_lib = None
_groups = set()
_why = "unprobed"
_probe_env = None


def _probe():
    global _lib, _groups, _why, _probe_env
    key = (os.environ.get("NUMFAST_NATIVE_DISABLE"),
           os.environ.get("NUMFAST_NATIVE_DLL", DEFAULT_LIB))
    if key == _probe_env and _why != "unprobed":
        return
    _probe_env = key
    if key[0] == "1":
        _lib, _groups, _why = None, set(), "disabled-by-env"
        return
    try:
        # torch ships its own libamdhip64 with the same soname; whichever is
        # loaded first serves the process.  Loading torch's first keeps one
        # HIP runtime for both (the other order leaves torch without a GPU).
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        lib = ctypes.CDLL(key[1])
        _groups = bind(lib)
        _lib, _why = lib, "loaded:" + key[1]
    except (OSError, AttributeError) as e:
        _lib, _groups, _why = None, set(), "unavailable:%s" % e


def available():
    """True when the library is loaded (env-aware, like native_cpu.available)."""
    _probe()
    return _lib is not None


def why():
    _probe()
    return _why


def lib(group=None):
    """The loaded CDLL, or NativeUnavailable (no silent fallback)."""
    _probe()
    if _lib is None:
        raise NativeUnavailable("libnumfast_hip %s" % _why)
    if group is not None and group not in _groups:
        raise NativeUnavailable("libnumfast_hip lacks the %r symbol group" % group)
    return _lib


def loaded_path():
    _probe()
    return _why.split(":", 1)[1] if _why.startswith("loaded:") else None


def _p(a):
    """Data pointer of a NumPy array or a torch tensor (None for None)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------ group-by lanes
def i32_gate_ok(vals, n=None, _b=None):
    """int32-direct gate (native_cpu.py:360-379): n*maxabs < 2**31-1."""
    vals = np.asarray(vals)
    if _b is None:
        if vals.size == 0:
            return True
        mn, mx = int(vals.min()), int(vals.max())
    else:
        mn, mx = _b
    nn = vals.size if n is None else int(n)
    return nn * max(abs(mn), abs(mx)) < 2 ** 31 - 1


def _bounds(v):
    """(min, max) as Python ints, or None for a float / empty column."""
    v = np.asarray(v)
    if v.dtype.kind not in "iu" or v.size == 0:
        return None
    return int(v.min()), int(v.max())


def int_exact_ok(vcols):
    """Executor gate for the float64 lanes (cpu.py:424-446): integer
    columns summed through nf_group_sum_count / nf_group_multi_sum_count
    stay exact while n * max|v| < 2**53 (every partial sum is then a
    float64 integer)."""
    for v in vcols:
        b = _bounds(v)
        if b is not None and int(np.asarray(v).size) * max(abs(b[0]), abs(b[1])) >= 2 ** 53:
            return False
    return True


def intlane_ok(vcols):
    """Executor gate for the integer lanes (cpu.py:353-378): every integer
    column has n * max|v| < 2**63 - 1, so the int32 / checked int64 lanes
    serve it exactly."""
    for v in vcols:
        b = _bounds(v)
        if b is not None and int(np.asarray(v).size) * max(abs(b[0]), abs(b[1])) >= 2 ** 63 - 1:
            return False
    return True


def native_st(keys, vcols, m):
    """The executor's single-thread native dense aggregate (cpu.py:461-518)
    over int32 codes in [0, m): integer-bearing columns take the mixed lanes
    (int32 when the i32 gate holds, checked int64, float64), anything else
    or a lane failure takes the float64 fused lanes when int_exact_ok holds;
    then one carry compaction per column (nf_carry_build_*).  Returns
    (ukeys, counts, sums_list) or None when the executor would fall back."""
    try:
        keys = _c(keys, np.int32)
        cols = [np.ascontiguousarray(v) for v in vcols]
        res = None
        if any(c.dtype.kind in "iu" for c in cols):
            res = mixed_sum_count(keys, cols, m)
        if res is None:
            if not int_exact_ok(cols):
                return None
            if len(cols) == 1:
                s1, counts_m = fused_sum_count(keys, _c(cols[0], np.float64), m)
                res = ([s1], counts_m)
            else:
                res = multi_sum_count(keys, [_c(c, np.float64) for c in cols], m)
        sums_list, counts_m = res
        ukeys, counts, out = None, None, []
        for s_m in sums_list:
            uk, cnt, sm = carry_build(counts_m, s_m)
            if ukeys is None:
                ukeys, counts = uk, cnt
            out.append(sm)
        return ukeys, counts, out
    except (OverflowError, RuntimeError, AssertionError, ValueError):
        return None


def native_mt(keys, vcols, m, t):
    """The executor's host multi-thread dense aggregate (cpu.py:521-636,
    the S-state path): a contiguous row partition into t slices, each
    aggregated by one pool worker through the lanes (mixed lanes when a
    column is an integer, the float64 lanes otherwise; ctypes releases the
    GIL, and each thread has its own library stream and scratch), then a
    merge of the slice states in slice order (integer states widened to
    int64) and one carry compaction per column.  Returns (ukeys, counts,
    sums_list) or None where the executor would fall back."""
    from concurrent.futures import ThreadPoolExecutor
    keys = _c(keys, np.int32)
    cols = [np.ascontiguousarray(v) for v in vcols]
    mix = any(c.dtype.kind in "iu" for c in cols)
    if (not mix and not int_exact_ok(cols)) or (mix and not intlane_ok(cols)):
        return None
    n = int(keys.size)
    t = max(1, min(int(t), n))
    bounds = np.linspace(0, n, t + 1).astype(np.int64)

    def work(w):
        a, b = int(bounds[w]), int(bounds[w + 1])
        kk, vv = keys[a:b], [c[a:b] for c in cols]
        if mix:
            r = mixed_sum_count(kk, vv, m)
            if r is not None:
                return r
        if len(vv) == 1:
            s1, cc = fused_sum_count(kk, _c(vv[0], np.float64), m)
            return [s1], cc
        return multi_sum_count(kk, [_c(x, np.float64) for x in vv], m)

    try:
        with ThreadPoolExecutor(max_workers=t) as ex:
            parts = list(ex.map(work, range(t)))
        sums_m = [np.zeros(m, dtype=np.int64 if s_.dtype.kind in "iu" else s_.dtype)
                  for s_ in parts[0][0]]
        counts_m = np.zeros(m, dtype=np.int64)
        for ss, cc in parts:  # slice order: the sequential merge
            for j, s_ in enumerate(ss):
                sums_m[j] += s_
            counts_m += cc
        ukeys, counts, out = None, None, []
        for s_m in sums_m:
            uk, cnt, sm = carry_build(counts_m, s_m)
            if ukeys is None:
                ukeys, counts = uk, cnt
            out.append(sm)
        return ukeys, counts, out
    except (OverflowError, RuntimeError, AssertionError, ValueError):
        return None


def fused_sum_count(keys, vals, g):
    """native_cpu.py:791-802 -> nf_group_sum_count."""
    keys, vals = _c(keys, np.int32), _c(vals, np.float64)
    L = lib()
    sums = np.zeros(g)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_group_sum_count(_p(keys), _p(vals), keys.size, _p(sums), _p(counts), g)
    assert rc == 0, rc
    return sums, counts


def multi_sum_count(keys, cols, g):
    """native_cpu.py:805-818 -> nf_group_multi_sum_count (SoA columns)."""
    keys = _c(keys, np.int32)
    cols = [_c(c, np.float64) for c in cols]
    ncols = len(cols)
    L = lib()
    vflat = np.ascontiguousarray(np.stack(cols).reshape(-1)) if ncols else np.zeros(0)
    sums = np.zeros(ncols * g)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_group_multi_sum_count(_p(keys), _p(vflat), keys.size, ncols, _p(sums),
                                    _p(counts), g)
    assert rc == 0, rc
    return [sums[i * g:(i + 1) * g] for i in range(ncols)], counts


def sum_count_i32(keys, vals, g):
    """native_cpu.py:394-413 -> nf_group_variant_i32(variant 0)."""
    keys, vals = _c(keys, np.int32), _c(vals, np.int32)
    L = lib("variant")
    sums = np.zeros(g, dtype=np.int32)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_group_variant_i32(_p(keys), _p(vals), keys.size, 0, _p(sums), _p(counts), g)
    if rc == -4:
        raise OverflowError("sum_count_i32: int32 accumulator overflow")
    assert rc == 0, rc
    return sums, counts


def sum_count_i64(keys, ticks, g):
    """native_cpu.py:416-445 -> nf_group_variant_i64(variant 1, V_CHECKED)."""
    keys, ticks = _c(keys, np.int32), _c(ticks, np.int64)
    L = lib("variant")
    sums = np.zeros(g, dtype=np.int64)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_group_variant_i64(_p(keys), _p(ticks), keys.size, 1, _p(sums), _p(counts), g)
    if rc == -4:
        raise OverflowError("sum_count_i64: int64 accumulator overflow")
    if rc == -2:
        raise RuntimeError("sum_count_i64: dense key outside [0, %d)" % g)
    assert rc == 0, rc
    return sums, counts


def mixed_sum_count(keys, cols, g, _bounds=None):
    """native_cpu.py:448-492: per-column dtype routing; None on any lane failure."""
    try:
        keys = _c(keys, np.int32)
        raw = [np.ascontiguousarray(c) for c in cols]
        n = int(keys.size)
        outs, counts = [], None
        for i, c in enumerate(raw):
            b = None if _bounds is None else _bounds[i]
            if c.dtype == np.dtype(np.int32):
                if i32_gate_ok(c, n, _b=b):
                    s, cc = sum_count_i32(keys, c, g)
                else:
                    s, cc = sum_count_i64(keys, _c(c, np.int64), g)
            elif c.dtype.kind in "iu":
                s, cc = sum_count_i64(keys, _c(c, np.int64), g)
            elif c.dtype.kind == "f":
                s, cc = fused_sum_count(keys, _c(c, np.float64), g)
            else:
                return None
            outs.append(s)
            if counts is None:
                counts = cc
        return outs, counts
    except (OverflowError, RuntimeError, AssertionError, MemoryError, ValueError):
        return None


def count_only(keys, g):
    """native_cpu.py:719-729 -> nf_group_count_only."""
    keys = _c(keys, np.int32)
    L = lib("mixed")
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_group_count_only(_p(keys), keys.size, _p(counts), g)
    assert rc == 0, rc
    return counts


def mixed_sum_only_count(keys, cols, g, _bounds=None):
    """native_cpu.py:732-788: one sum-only pass over i32+f64 lanes + counts."""
    try:
        keys = _c(keys, np.int32)
        raw = [np.ascontiguousarray(c) for c in cols]
        L = lib("mixed")
        n = int(keys.size)
        i32c, f64c, order = [], [], []
        for i, c in enumerate(raw):
            b = None if _bounds is None else _bounds[i]
            if c.dtype == np.dtype(np.int32):
                if not i32_gate_ok(c, n, _b=b):
                    return None
                order.append(("i", len(i32c)))
                i32c.append(c)
            elif c.dtype.kind == "f":
                order.append(("f", len(f64c)))
                f64c.append(_c(c, np.float64))
            else:
                return None
        ni, nf = len(i32c), len(f64c)
        i32a = np.ascontiguousarray(np.stack(i32c).reshape(-1)) if ni else None
        f64a = np.ascontiguousarray(np.stack(f64c).reshape(-1)) if nf else None
        s32 = np.zeros(ni * g, dtype=np.int32)
        s64 = np.zeros(nf * g, dtype=np.float64)
        rc = L.nf_group_mixed_sum_only(_p(keys), _p(i32a), _p(f64a), n, ni, nf,
                                       _p(s32) if ni else None, _p(s64) if nf else None, g)
        if rc != 0:
            return None
        counts = count_only(keys, g)
        outs = [None] * len(raw)
        for j, (k, ci) in enumerate(order):
            outs[j] = s32[ci * g:(ci + 1) * g] if k == "i" else s64[ci * g:(ci + 1) * g]
        return outs, counts
    except (OverflowError, RuntimeError, AssertionError, MemoryError, ValueError):
        return None


def owner_shard_into(keys, a, b, c, lo, hi, m, s1, s2, s3, cc):
    """native_cpu.py:582-615 -> nf_group_owner_2i32_1f64; True on rc == 0."""
    try:
        L = lib("owner")
        keys, av, bv, cv = _c(keys, np.int32), _c(a, np.int32), _c(b, np.int32), _c(c, np.float64)
        rc = L.nf_group_owner_2i32_1f64(_p(keys), _p(av), _p(bv), _p(cv), keys.size, int(lo),
                                        int(hi), _p(s1), _p(s2), _p(s3), _p(cc), int(m))
        return rc == 0
    except (OverflowError, RuntimeError, AssertionError, MemoryError, ValueError):
        return False


def pack_sum_count_i32(k1, k2, m2, vals, g):
    """native_cpu.py:674-695 -> nf_pack_sum_count_i32."""
    k1, k2, vals = _c(k1, np.int32), _c(k2, np.int32), _c(vals, np.int32)
    L = lib("fused")
    sums = np.zeros(g, dtype=np.int32)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_pack_sum_count_i32(_p(k1), _p(k2), int(m2), _p(vals), k1.size, _p(sums),
                                 _p(counts), g)
    if rc == -4:
        raise OverflowError("pack_sum_count_i32: int32 accumulator overflow")
    assert rc == 0, rc
    return sums, counts


def pack_sum_count_f64(k1, k2, m2, vals, g):
    """native_cpu.py:698-716 -> nf_pack_sum_count_f64."""
    k1, k2, vals = _c(k1, np.int32), _c(k2, np.int32), _c(vals, np.float64)
    L = lib("fused")
    sums = np.zeros(g, dtype=np.float64)
    counts = np.zeros(g, dtype=np.int64)
    rc = L.nf_pack_sum_count_f64(_p(k1), _p(k2), int(m2), _p(vals), k1.size, _p(sums),
                                 _p(counts), g)
    assert rc == 0, rc
    return sums, counts


def pack_i32_direct(k1, k2, m2):
    """native_cpu.py:1288-1297 -> nf_pack_i32_direct."""
    k1, k2 = _c(k1, np.int32), _c(k2, np.int32)
    L = lib()
    out = np.zeros(k1.size, dtype=np.int32)
    rc = L.nf_pack_i32_direct(_p(k1), _p(k2), int(m2), k1.size, _p(out))
    assert rc == 0, rc
    return out


def build_pattern_buffers(strs):
    """strings -> (data bytes, offs int32) (native_cpu.py:1300-1305)."""
    enc = [s.encode("utf-8") if isinstance(s, str) else bytes(s) for s in strs]
    offs = np.zeros(len(enc) + 1, dtype=np.int32)
    offs[1:] = np.cumsum([len(b) for b in enc]).astype(np.int64)
    return b"".join(enc), offs


def pattern_encode_buffers(data, offs, prefix):
    """native_cpu.py:1308-1325 -> nf_pattern_encode."""
    n = offs.size - 1
    L = lib()
    p = prefix.encode() if isinstance(prefix, str) else bytes(prefix)
    dp = np.frombuffer(data + b"\0", dtype=np.uint8)
    pp = np.frombuffer(p + b"\0", dtype=np.uint8)
    offs = _c(offs, np.int32)
    codes = np.zeros(n, dtype=np.int32)
    valid = np.zeros(n, dtype=np.uint8)
    w = np.zeros(1, dtype=np.int32)
    e = np.zeros(1, dtype=np.int32)
    rc = L.nf_pattern_encode(_p(dp), len(data), _p(offs), n, _p(pp), len(p), _p(codes),
                             _p(valid), _p(w), _p(e))
    if rc == -2:
        raise OverflowError("pattern_encode: int32 overflow at row %d" % int(e[0]))
    assert rc == 0, (rc, int(e[0]))
    return codes, valid, int(w[0])


def sorted_run(keys, vals):
    """native_cpu.py:1328-1353 -> nf_sorted_run_i64/_f64."""
    keys = _c(keys, np.int32)
    n = keys.size
    L = lib()
    if np.asarray(vals).dtype.kind in "iu":
        vals = _c(vals, np.int64)
        uk, s, c = (np.zeros(n, dtype=np.int64) for _ in range(3))
        ng = int(L.nf_sorted_run_i64(_p(keys), _p(vals), n, _p(uk), _p(s), _p(c)))
    else:
        vals = _c(vals, np.float64)
        uk, c = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
        s = np.zeros(n, dtype=np.float64)
        ng = int(L.nf_sorted_run_f64(_p(keys), _p(vals), n, _p(uk), _p(s), _p(c)))
    assert ng >= 0, ng
    return uk[:ng], s[:ng], c[:ng]


def carry_build(counts_m, sums_m):
    """native_cpu.py:1356-1381 -> nf_carry_build_i64/_f64."""
    counts_m = _c(counts_m, np.int64)
    m = counts_m.size
    L = lib()
    uk, c = np.zeros(m, dtype=np.int64), np.zeros(m, dtype=np.int64)
    if np.asarray(sums_m).dtype.kind in "iu":
        sums_m = _c(sums_m, np.int64)
        s = np.zeros(m, dtype=np.int64)
        ng = int(L.nf_carry_build_i64(_p(counts_m), _p(sums_m), m, _p(uk), _p(c), _p(s)))
    else:
        sums_m = _c(sums_m, np.float64)
        s = np.zeros(m, dtype=np.float64)
        ng = int(L.nf_carry_build_f64(_p(counts_m), _p(sums_m), m, _p(uk), _p(c), _p(s)))
    assert ng >= 0, ng
    return uk[:ng], c[:ng], s[:ng]


# ------------------------------------------------------------ unique / sort
def unique_inverse(keys):
    """native_cpu.py:1465-1496: (uniq, inv int64, backend)."""
    keys = np.ascontiguousarray(np.asarray(keys))
    dt = keys.dtype
    if dt not in (np.dtype(np.int32), np.dtype(np.int64)):
        raise TypeError("unique_inverse: int32/int64 keys only, got %s" % dt)
    n = keys.size
    if n == 0:
        return keys[:0], np.zeros(0, dtype=np.int64), "native"
    L = lib("unique")
    tu = np.uint32 if dt == np.dtype(np.int32) else np.uint64
    uniq = np.empty(n, dtype=dt)
    inv = np.empty(n, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    t0, t1 = np.empty(n, dtype=tu), np.empty(n, dtype=tu)
    tp = np.empty(n, dtype=np.int32)
    fn = L.nf_unique_inverse_i32 if dt == np.dtype(np.int32) else L.nf_unique_inverse_i64
    ng = int(fn(_p(keys), n, _p(uniq), _p(inv), _p(perm), _p(t0), _p(t1), _p(tp)))
    if ng < 0:
        raise RuntimeError("nf_unique_inverse rc=%d" % ng)
    return uniq[:ng].copy(), inv.astype(np.int64), "native"


def _g_sort_asc(keys):
    """Monotone probe (native_cpu.py:1512-1522): (nondecreasing, strict)."""
    if keys.size < 2:
        return True, True
    if not bool(np.all(keys[1:] >= keys[:-1])):
        return False, False
    return True, bool(np.all(keys[1:] != keys[:-1]))


def _g_sort_desc(keys):
    """native_cpu.py:1525-1533."""
    if keys.size < 2:
        return True, True
    if not bool(np.all(keys[1:] <= keys[:-1])):
        return False, False
    return True, bool(np.all(keys[1:] != keys[:-1]))


def sort_perm_raw(keys, descending):
    """One nf_sort_perm_* call (no monotone shortcuts): int32 permutation."""
    k = np.ascontiguousarray(np.asarray(keys))
    dt = k.dtype
    n = k.size
    if dt not in (np.dtype(np.int32), np.dtype(np.int64)):
        raise TypeError("sort_perm: int32/int64 keys only, got %s" % dt)
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    L = lib("sort")
    tu = np.uint32 if dt == np.dtype(np.int32) else np.uint64
    perm = np.empty(n, dtype=np.int32)
    t0, t1 = np.empty(n, dtype=tu), np.empty(n, dtype=tu)
    tp = np.empty(n, dtype=np.int32)
    fn = L.nf_sort_perm_i32 if dt == np.dtype(np.int32) else L.nf_sort_perm_i64
    rc = int(fn(_p(k), n, 1 if descending else 0, _p(perm), _p(t0), _p(t1), _p(tp)))
    if rc != 0:
        raise RuntimeError("nf_sort_perm rc=%d" % rc)
    return perm


def sort_perm_step(keys, descending):
    """native_cpu.py:1602-1652: (step, backend) with the monotone guards.

    "ident"/"rev" answer without a library call exactly where the reference
    does; a non-strict cross-monotone column goes to the radix kernel (the
    reference hands it to NumPy, which produces the same permutation).
    """
    k = np.ascontiguousarray(np.asarray(keys))
    n = k.size
    if k.dtype not in (np.dtype(np.int32), np.dtype(np.int64)) or n > 2 ** 31 - 1:
        raise TypeError("sort_perm_step: int32/int64 keys, n <= 2**31-1")
    if n == 0:
        return np.zeros(0, dtype=np.int32), "ident"
    if descending:
        ok_d, _ = _g_sort_desc(k)
        if ok_d:
            return np.arange(n, dtype=np.int32), "ident"
        ok_a, st_a = _g_sort_asc(k)
        if ok_a and st_a:
            return np.arange(n - 1, -1, -1, dtype=np.int32), "rev"
    else:
        ok_a, _ = _g_sort_asc(k)
        if ok_a:
            return np.arange(n, dtype=np.int32), "ident"
        ok_d, st_d = _g_sort_desc(k)
        if ok_d and st_d:
            return np.arange(n - 1, -1, -1, dtype=np.int32), "rev"
    return sort_perm_raw(k, descending), "native"


def sort_perm(cols, descending, valid=None):
    """Multi-key stable sort permutation (cpu.py:1174-1225, _sort_perm):
    successive stable key steps from the last key to the first, each from
    sort_perm_step (the HIP radix argsort or a monotone shortcut); rows
    with valid False go last in input order.  int32 positions."""
    cols = [np.asarray(c) for c in cols]
    n = cols[0].size
    for c in cols:
        if c.size != n:
            raise ValueError("sort keys size mismatch %d != %d" % (c.size, n))
    if valid is not None:
        msk = np.asarray(valid, dtype=bool).reshape(-1)
        if msk.size != n:
            raise ValueError("sort validity size %d != keys %d" % (msk.size, n))
        vpos, ipos = np.flatnonzero(msk), np.flatnonzero(~msk)
    else:
        vpos, ipos = np.arange(n), np.zeros(0, dtype=np.int64)
    idx = np.arange(vpos.size)
    for c, desc in zip(reversed(cols), reversed(list(descending))):
        k = c[vpos][idx] if vpos.size else c[:0]
        step, backend = sort_perm_step(k, bool(desc))
        if backend != "ident":
            idx = idx[step]
    perm = np.empty(n, dtype=np.int32)
    perm[:vpos.size] = vpos[idx]
    perm[vpos.size:] = ipos
    return perm


def gather(values, indices, valid=None):
    """Positional take (cpu.py:1110-1138 _gather_ref, the IR gather at
    :3170-3176): values[indices] in index order, validity taken alike.
    Raises IndexError for an index outside [0, n) and TypeError for
    non-integer indices, as the reference refuses them."""
    v = np.ascontiguousarray(np.asarray(values))
    idx = np.asarray(indices)
    if idx.dtype.kind not in "iu":
        raise TypeError("gather needs integer indices, got %s" % idx.dtype)
    idx = _c(idx, np.int64 if idx.dtype.itemsize > 4 or idx.dtype == np.uint32 else np.int32)
    if v.dtype.itemsize not in (1, 4, 8):
        raise TypeError("gather: 1-, 4- or 8-byte elements only, got %s" % v.dtype)
    k = idx.size
    out = np.empty(k, dtype=v.dtype)
    vin = None if valid is None else _c(valid, np.uint8)
    vout = None if valid is None else np.empty(k, dtype=np.uint8)
    rc = int(lib("hip").nfh_gather(_p(v), v.size, v.dtype.itemsize, _p(idx), idx.dtype.itemsize,
                                   k, _p(out), _p(vin), _p(vout)))
    if rc == -2:
        raise IndexError("gather index out of range")
    if rc != 0:
        raise RuntimeError("nfh_gather rc=%d" % rc)
    return out, (None if vout is None else vout.astype(bool))


# ------------------------------------------------------------ additive surface
def group_agg(keys, vals, valid=None, key_valid=None):
    """Group-by with sum/count/min/max/mean and NULL masks (nfh_group_agg_i32).

    Returns a dict shaped like ColumnCarry (carry.py:21-62): ukeys int64
    ascending, counts int64, sums int64, mins/maxs int32, means float64.
    """
    keys, vals = _c(keys, np.int32), _c(vals, np.int32)
    n = keys.size
    L = lib("hip")
    vm = None if valid is None else _c(valid, np.uint8)
    km = None if key_valid is None else _c(key_valid, np.uint8)
    out = {"ukeys": np.zeros(n, np.int64), "counts": np.zeros(n, np.int64),
           "sums": np.zeros(n, np.int64), "mins": np.zeros(n, np.int32),
           "maxs": np.zeros(n, np.int32), "means": np.zeros(n, np.float64)}
    ng = int(L.nfh_group_agg_i32(_p(keys), _p(vals), _p(vm), _p(km), n, _p(out["ukeys"]),
                                 _p(out["counts"]), _p(out["sums"]), _p(out["mins"]),
                                 _p(out["maxs"]), _p(out["means"]), n))
    if ng < 0:
        raise RuntimeError("nfh_group_agg_i32 rc=%d" % ng)
    return {k: v[:ng] for k, v in out.items()}


def group_agg64(keys, vals, valid=None, key_valid=None):
    """Group-by of a float64 or int64 value column with sum/count/min/max/
    mean and NULL masks (nfh_group_agg_f64 / nfh_group_agg_i64): the
    executor's min/max branch for 64-bit values (cpu.py:3507-3563,
    :2570-2601).  Returns a ColumnCarry-shaped dict (carry.py:21-62):
    ukeys / counts int64, sums / mins / maxs in the value dtype, means
    float64."""
    vals = np.asarray(vals)
    dt = np.float64 if vals.dtype.kind == "f" else np.int64
    keys, vals = _c(keys, np.int32), _c(vals, dt)
    n = keys.size
    L = lib("hip")
    vm = None if valid is None else _c(valid, np.uint8)
    km = None if key_valid is None else _c(key_valid, np.uint8)
    out = {"ukeys": np.zeros(n, np.int64), "counts": np.zeros(n, np.int64),
           "sums": np.zeros(n, dt), "mins": np.zeros(n, dt), "maxs": np.zeros(n, dt),
           "means": np.zeros(n, np.float64)}
    fn = L.nfh_group_agg_f64 if dt is np.float64 else L.nfh_group_agg_i64
    ng = int(fn(_p(keys), _p(vals), _p(vm), _p(km), n, _p(out["ukeys"]), _p(out["counts"]),
                _p(out["sums"]), _p(out["mins"]), _p(out["maxs"]), _p(out["means"]), n))
    if ng < 0:
        raise RuntimeError("%s rc=%d" % (fn.__name__, ng))
    return {k: v[:ng] for k, v in out.items()}


def unique_valid(keys, valid=None):
    """np.unique over valid rows, inv = -1 on NULL rows (nfh_unique_valid_i32)."""
    keys = _c(keys, np.int32)
    n = keys.size
    L = lib("hip")
    vm = None if valid is None else _c(valid, np.uint8)
    uniq, inv = np.zeros(n, np.int32), np.zeros(n, np.int32)
    ng = int(L.nfh_unique_valid_i32(_p(keys), _p(vm), n, _p(uniq), _p(inv)))
    if ng < 0:
        raise RuntimeError("nfh_unique_valid_i32 rc=%d" % ng)
    return uniq[:ng], inv


# ------------------------------------------------------------ select / mask
_SCATTER = {np.dtype(np.int32): "nf_select_scatter_i32", np.dtype(np.int64): "nf_select_scatter_i64",
            np.dtype(np.float32): "nf_select_scatter_f32",
            np.dtype(np.float64): "nf_select_scatter_f64", np.dtype(np.uint8): "nf_select_scatter_u8"}


def select_count(mask):
    """native_cpu.py select lane -> nf_select_count: rows with mask != 0."""
    m = _c(mask, np.uint8)
    r = int(lib("select").nf_select_count(_p(m), m.size))
    if r < 0:
        raise RuntimeError("nf_select_count rc=%d" % r)
    return r


def select_scatter(src, mask):
    """Order-preserving compaction src[mask != 0] (nf_select_scatter_*)."""
    src = np.ascontiguousarray(src)
    m = _c(mask, np.uint8)
    fn = getattr(lib("select"), _SCATTER[src.dtype])
    out = np.zeros(src.size, src.dtype)
    r = int(fn(_p(src), _p(m), src.size, _p(out)))
    if r < 0:
        raise RuntimeError("select scatter rc=%d" % r)
    return out[:r]


def mask_op(op, a, b=None):
    """nf_mask_and / nf_mask_or / nf_mask_not; output normalised 0/1."""
    a = _c(a, np.uint8)
    out = np.zeros(a.size, np.uint8)
    L = lib("select")
    if op == "not":
        rc = L.nf_mask_not(_p(a), a.size, _p(out))
    else:
        b = _c(b, np.uint8)
        rc = getattr(L, "nf_mask_" + op)(_p(a), _p(b), a.size, _p(out))
    if rc != 0:
        raise RuntimeError("nf_mask_%s rc=%d" % (op, rc))
    return out


# ------------------------------------------------------------ learned sort
MLP_P = 1153


class CdfModel(ctypes.Structure):
    """struct nfh_cdf_model (include/numfast_hip.h)."""
    _fields_ = [("p", ctypes.c_float * MLP_P), ("wmin", ctypes.c_uint32),
                ("span", ctypes.c_uint32), ("hints", ctypes.c_uint32)]

    def params(self):
        return np.ctypeslib.as_array(self.p).copy()


def cdf_train(keys, steps=100, linear_start=False):
    """Train the CDF model on a column (host array or device tensor);
    linear_start: begin from the linear CDF instead of the greedy fit.
    Returns (model, final loss)."""
    L = lib("learned")
    n = keys.numel() if hasattr(keys, "numel") else np.asarray(keys).size
    k = keys if hasattr(keys, "data_ptr") else _c(keys, np.int32)
    model = CdfModel()
    loss = ctypes.c_float()
    rc = L.nfh_cdf_train_i32(_p(k), n, steps, 1 if linear_start else 0, ctypes.byref(model),
                             ctypes.byref(loss))
    if rc != 0:
        raise RuntimeError("nfh_cdf_train_i32 rc=%d" % rc)
    return model, float(loss.value)


def sort_perm_learned(keys, model):
    """Stable ascending argsort via the learned-CDF bucket sort; returns
    (perm, stats) with stats = (buckets, largest, over the LDS capacity,
    re-sorted by the plain-sort fix-up)."""
    keys = _c(keys, np.int32)
    perm = np.zeros(keys.size, np.int32)
    stats = np.zeros(4, np.uint64)
    rc = lib("learned").nfh_sort_perm_learned_i32(_p(keys), keys.size, ctypes.byref(model),
                                                  _p(perm), _p(stats))
    if rc != 0:
        raise RuntimeError("nfh_sort_perm_learned_i32 rc=%d" % rc)
    return perm, stats


def mlp_grad(params, x, y):
    """(grad[1153], loss, pred) of the CDF network's mean squared error,
    from the training kernels."""
    P = _c(params, np.float32)
    x, y = _c(x, np.float32), _c(y, np.float32)
    g = np.zeros(MLP_P, np.float32)
    loss = np.zeros(1, np.float32)
    pred = np.zeros(x.size, np.float32)
    rc = lib("learned").nfh_mlp_grad(_p(P), _p(x), _p(y), x.size, _p(g), _p(loss), _p(pred))
    if rc != 0:
        raise RuntimeError("nfh_mlp_grad rc=%d" % rc)
    return g, float(loss[0]), pred


def filter_derive_sort_limit(v, limit=1000, mul=3, model=None):
    """filter(v > 0) -> derive(w = v * mul, int32 wrap) -> sort(w) -> limit:
    (w, source rows) of the first min(limit, m) rows."""
    v = _c(v, np.int32)
    w = np.zeros(limit, np.int32)
    rows = np.zeros(limit, np.int32)
    r = int(lib("learned").nfh_filter_derive_sort_limit_i32(
        _p(v), v.size, mul, limit, None if model is None else ctypes.byref(model), _p(w),
        _p(rows)))
    if r < 0:
        raise RuntimeError("nfh_filter_derive_sort_limit_i32 rc=%d" % r)
    return w[:r], rows[:r]


# ------------------------------------------------------------ batched group-by
class GroupBatch:
    """A batched group-by plan over device tensors (nfh_group_batch_plan).

    keys / vals: int32 CUDA tensors of all queries' rows, concatenated;
    row_off: int64 (nq + 1); card: int32 (nq).  run(mode) returns
    (ng, ukeys, counts, sums, group_off) views of the device outputs;
    mode 0 replays the captured graph, mode 1 launches query by query.
    """

    def __init__(self, torch, keys, vals, row_off, card, cap=None):
        L = lib("batch")
        self._L, self._torch = L, torch
        nq = int(card.numel() if hasattr(card, "numel") else len(card))
        n = int(keys.numel())
        cap = n if cap is None else int(cap)
        dev = keys.device
        self.keys, self.vals = keys, vals
        self.row_off = torch.as_tensor(row_off, dtype=torch.int64, device=dev)
        self.card = torch.as_tensor(card, dtype=torch.int32, device=dev)
        self.ukeys = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        self.counts = torch.empty_like(self.ukeys)
        self.sums = torch.empty_like(self.ukeys)
        self.group_off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
        rc = ctypes.c_int32(0)
        torch.cuda.synchronize()
        self.plan = L.nfh_group_batch_plan(
            keys.data_ptr(), vals.data_ptr(), self.row_off.data_ptr(), self.card.data_ptr(), nq,
            self.ukeys.data_ptr(), self.counts.data_ptr(), self.sums.data_ptr(),
            self.group_off.data_ptr(), cap, ctypes.byref(rc))
        if not self.plan:
            raise RuntimeError("nfh_group_batch_plan rc=%d" % rc.value)

    def run(self, mode=0):
        ng = int(self._L.nfh_group_batch_run(self.plan, mode))
        if ng < 0:
            return ng, None, None, None, None
        return ng, self.ukeys[:ng], self.counts[:ng], self.sums[:ng], self.group_off

    def info(self):
        out = np.zeros(4, np.uint64)
        self._L.nfh_group_batch_info(self.plan, out.ctypes.data)
        o2 = np.zeros(3, np.uint64)
        self._L.nfh_group_batch_info2(self.plan, o2.ctypes.data)
        return {"rows_per_item": int(out[0]), "workgroups": int(out[1]),
                "work_items": int(out[2]), "slots": int(out[3]),
                "wide_tiles": int(o2[0]), "units": int(o2[1]), "entries": int(o2[2])}

    def close(self):
        if self.plan:
            self._L.nfh_group_batch_free(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ strategy selector
# dense, partitioned, sort-based, hash-scatter, partitioned with 4096-key
# buckets (strategy = class + 1)
SEL_CLASSES = 5
SEL_P = 64 * 64 + 64 + SEL_CLASSES * 64 + SEL_CLASSES


def groupby_features(keys):
    """The selector's 64 features of a key column (host array or tensor)."""
    out = np.zeros(64, np.float32)
    k = keys if hasattr(keys, "data_ptr") else _c(keys, np.int32)
    n = k.numel() if hasattr(k, "numel") else k.size
    rc = lib("selector").nfh_groupby_features_i32(_p(k), n, _p(out))
    if rc != 0:
        raise RuntimeError("nfh_groupby_features_i32 rc=%d" % rc)
    return out


def selector_init(seed=0):
    """He-style random start of the 64 -> 64 -> 5 classifier."""
    rng = np.random.default_rng(seed)
    p = np.zeros(SEL_P, np.float32)
    p[:64 * 64] = rng.normal(0, np.sqrt(2 / 64), 64 * 64)
    w2 = SEL_CLASSES * 64
    p[64 * 64 + 64:64 * 64 + 64 + w2] = rng.normal(0, np.sqrt(1 / 64), w2)
    return p


def selector_grad(params, X, y):
    g = np.zeros(SEL_P, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib("selector").nfh_selector_grad(_p(_c(params, np.float32)), _p(_c(X, np.float32)),
                                           _p(_c(y, np.int32)), len(y), _p(g), _p(loss))
    if rc != 0:
        raise RuntimeError("nfh_selector_grad rc=%d" % rc)
    return g, float(loss[0])


def selector_train(X, y, steps=300, lr=1e-2, params=None):
    p = selector_init() if params is None else _c(params, np.float32).copy()
    loss = np.zeros(1, np.float32)
    rc = lib("selector").nfh_selector_train(_p(_c(X, np.float32)), _p(_c(y, np.int32)), len(y),
                                            steps, lr, _p(p), _p(loss))
    if rc != 0:
        raise RuntimeError("nfh_selector_train rc=%d" % rc)
    return p, float(loss[0])


def selector_predict(params, X):
    X = _c(X, np.float32)
    out = np.zeros(X.shape[0], np.int32)
    rc = lib("selector").nfh_selector_predict(_p(_c(params, np.float32)), _p(X), X.shape[0],
                                              _p(out))
    if rc != 0:
        raise RuntimeError("nfh_selector_predict rc=%d" % rc)
    return out


def selector_grad_cost(params, X, T):
    """Cost-model gradient and loss at params over (X, T[n][4])."""
    g = np.zeros(SEL_P, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib("selector").nfh_selector_grad_cost(_p(_c(params, np.float32)), _p(_c(X, np.float32)),
                                                _p(_c(T, np.float32)), len(T), _p(g), _p(loss))
    if rc != 0:
        raise RuntimeError("nfh_selector_grad_cost rc=%d" % rc)
    return g, float(loss[0])


def selector_train_cost(X, T, steps=300, lr=1e-2, params=None):
    p = selector_init() if params is None else _c(params, np.float32).copy()
    loss = np.zeros(1, np.float32)
    rc = lib("selector").nfh_selector_train_cost(_p(_c(X, np.float32)), _p(_c(T, np.float32)),
                                                 len(T), steps, lr, _p(p), _p(loss))
    if rc != 0:
        raise RuntimeError("nfh_selector_train_cost rc=%d" % rc)
    return p, float(loss[0])


def selector_predict_cost(params, X):
    X = _c(X, np.float32)
    out = np.zeros(X.shape[0], np.int32)
    rc = lib("selector").nfh_selector_predict_cost(_p(_c(params, np.float32)), _p(X), X.shape[0],
                                                   _p(out))
    if rc != 0:
        raise RuntimeError("nfh_selector_predict_cost rc=%d" % rc)
    return out
