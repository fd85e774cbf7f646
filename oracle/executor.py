"""NumPy restatement of the reference executor semantics (TEST ONLY).

Checker for numfast_amd; see oracle/__init__.py for the import rule.
Each function cites the reference file:line it restates.
"""

import ctypes
import os
from fractions import Fraction
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
ORACLE_LIB = ROOT / "libnf_oracle.so"


def build_if_missing():
    """Compile nf_oracle.c when the shared object is absent or stale."""
    src = ROOT / "nf_oracle.c"
    if ORACLE_LIB.exists() and ORACLE_LIB.stat().st_mtime >= src.stat().st_mtime:
        return ORACLE_LIB
    import subprocess
    tmp = str(ORACLE_LIB) + ".tmp"
    subprocess.run(["gcc", "-O3", "-march=x86-64-v3", "-fPIC", "-shared",
                    "-fvisibility=hidden", "-o", tmp, str(src)], check=True)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def load_c_oracle():
    """ctypes handle on the C restatement, argtypes bound like the product."""
    from numfast_amd import native
    lib = ctypes.CDLL(str(build_if_missing()))
    groups = native.bind(lib)
    return lib, groups


# ----------------------------------------------------------------- sort
def stable_desc_idx(k):
    """src/Drivers/CPU/_lib/cpu.py:1154-1171: ascending stable argsort, then
    reverse the order of equal-key groups keeping input order inside."""
    asc = np.argsort(k, kind="stable")
    if asc.size == 0:
        return asc
    sk = k[asc]
    ch = np.empty(sk.shape, dtype=bool)
    ch[0] = True
    if sk.size > 1:
        ch[1:] = sk[1:] != sk[:-1]
    grp = np.cumsum(ch)
    rev = grp.max(initial=0) - grp
    return asc[np.argsort(rev, kind="stable")]


def sort_step(k, desc):
    """One stable key step (np.argsort / _stable_desc_idx), int32 positions."""
    k = np.asarray(k)
    return (stable_desc_idx(k) if desc else np.argsort(k, kind="stable")).astype(np.int32)


def sort_perm(cols, descending, valid=None):
    """src/Drivers/CPU/_lib/cpu.py:1174-1225 (_sort_perm): lexicographic via
    successive stable passes from the last key; invalid rows last, in input
    order."""
    cols = [np.asarray(c) for c in cols]
    n = cols[0].size
    if valid is not None:
        m = np.asarray(valid, dtype=bool).reshape(-1)
        vpos, ipos = np.flatnonzero(m), np.flatnonzero(~m)
    else:
        vpos, ipos = np.arange(n), np.zeros(0, dtype=np.int64)
    idx = np.arange(vpos.size)
    for c, desc in zip(reversed(cols), reversed(list(descending))):
        k = c[vpos][idx] if vpos.size else c[:0]
        idx = idx[sort_step(k, desc)]
    perm = np.empty(n, dtype=np.int32)
    perm[:vpos.size] = vpos[idx]
    perm[vpos.size:] = ipos
    return perm


# ----------------------------------------------------------------- unique
def unique_inverse(keys, valid=None):
    """np.unique(return_inverse=True) with int32 codes; with validity the
    unique set is taken over valid rows and inv = -1 on NULL rows
    (src/Drivers/CPU/_lib/cpu.py:4006-4040)."""
    keys = np.asarray(keys)
    if valid is None:
        u, inv = np.unique(keys, return_inverse=True)
        return u, inv.astype(np.int32)
    m = np.asarray(valid, dtype=bool)
    u, iv = np.unique(keys[m], return_inverse=True)
    inv = np.full(keys.shape, -1, dtype=np.int32)
    inv[m] = iv.astype(np.int32)
    return u, inv


# ----------------------------------------------------------------- group-by
def group_agg(keys, vals, valid=None, key_valid=None):
    """Group-by with sum/count/min/max/mean and explicit NULL masks.

    Restates the groupby_multi min/max branch of
    src/Drivers/CPU/_lib/cpu.py:3507-3563: rows whose key or value is NULL
    are compacted away first (_valid_mask AND, cpu.py:1374-1382,
    :3508-3511), so a key whose every value is NULL has no group
    (KNOWN_LIMITATIONS.md:147-165).  Groups come from the sorted unique of
    the surviving keys (ukeys ascending, groupindex.py:3-8); counts are
    occurrences; integer sums are exact int64 (cpu.py:3381-3383 casts the
    integer lane to int64 -- the reference's float64 bincount is exact under
    its 2**53 gate _int_exact_ok, cpu.py:424-446); min/max keep the value
    dtype (cpu.py:2470-2601); mean = sums.astype(f64) / counts
    (carry.py:55-62).
    """
    keys = np.asarray(keys)
    vals = np.asarray(vals)
    m = np.ones(keys.shape, dtype=bool)
    if valid is not None:
        m &= np.asarray(valid, dtype=bool)
    if key_valid is not None:
        m &= np.asarray(key_valid, dtype=bool)
    kk, vv = keys[m], vals[m]
    u, inv = np.unique(kk, return_inverse=True)
    ng = u.size
    counts = np.bincount(inv, minlength=ng).astype(np.int64)
    if vv.dtype.kind in "iu":
        sums = np.zeros(ng, dtype=np.int64)
        np.add.at(sums, inv, vv.astype(np.int64))
        info = np.iinfo(vv.dtype)
        mins = np.full(ng, info.max, dtype=vv.dtype)
        maxs = np.full(ng, info.min, dtype=vv.dtype)
    else:
        sums = np.bincount(inv, weights=vv.astype(np.float64), minlength=ng)
        mins = np.full(ng, np.inf, dtype=vv.dtype)
        maxs = np.full(ng, -np.inf, dtype=vv.dtype)
    np.minimum.at(mins, inv, vv)
    np.maximum.at(maxs, inv, vv)
    with np.errstate(divide="ignore", invalid="ignore"):
        means = sums.astype(np.float64) / counts
    return {"ukeys": u.astype(np.int64), "counts": counts, "sums": sums,
            "mins": mins, "maxs": maxs, "means": means}


def i32_gate_ok(vals, n=None):
    """src/Drivers/CPU/_lib/native_cpu.py:360-379: the int32-direct lane is
    safe when n * max|v| < 2**31 - 1 (all rows in one group)."""
    vals = np.asarray(vals)
    if vals.size == 0:
        return True
    nn = vals.size if n is None else int(n)
    return nn * max(abs(int(vals.min())), abs(int(vals.max()))) < 2 ** 31 - 1


def int_exact_ok(vcols):
    """src/Drivers/CPU/_lib/cpu.py:424-446 (_int_exact_ok): integer columns
    stay exact through the float64 lanes iff n * max|v| < 2**53; float and
    empty columns do not count."""
    for v in vcols:
        v = np.asarray(v)
        if v.dtype.kind in "iu" and v.size:
            if v.size * max(abs(int(v.min())), abs(int(v.max()))) >= 2 ** 53:
                return False
    return True


def intlane_ok(vcols):
    """src/Drivers/CPU/_lib/cpu.py:353-378 (_intlane_ok): every integer
    column has n * max|v| < 2**63 - 1."""
    for v in vcols:
        v = np.asarray(v)
        if v.dtype.kind in "iu" and v.size:
            if v.size * max(abs(int(v.min())), abs(int(v.max()))) >= 2 ** 63 - 1:
                return False
    return True


def exact_sums_f64(keys, vals, g):
    """Exact-rational per-group sums (Fraction), for error reporting only.
    Pure Python: small inputs."""
    acc = [Fraction(0)] * g
    for k, v in zip(np.asarray(keys).tolist(), np.asarray(vals, dtype=np.float64).tolist()):
        acc[k] += Fraction(v)
    return acc


def rel_err_vs_exact(got, exact):
    """max |got - exact| / max(|exact|, tiny) over groups (float)."""
    worst = 0.0
    for gv, ev in zip(np.asarray(got, dtype=np.float64).tolist(), exact):
        d = abs(Fraction(gv) - ev)
        den = abs(ev) if ev != 0 else Fraction(1)
        worst = max(worst, float(d / den))
    return worst


# ----------------------------------------------------------------- pipeline
def filter_derive_sort_limit(v, limit=1000, mul=3):
    """filter(v>0) -> derive(w = v*mul with int32 wrap) -> sort(w) stable
    ascending -> limit(limit).  The derive multiplies in NumPy's int32
    arithmetic and keeps the int32 result (cpu.py:3034-3035, :3074); sort
    is _sort_perm with one ascending key (cpu.py:1174-1225); limit is the
    row slice (cpu.py:3188-3202).  Returns (w_sorted_limited, source rows)."""
    v = np.asarray(v, dtype=np.int32)
    rows = np.flatnonzero(v > 0)
    w = (v[rows].astype(np.int64) * mul).astype(np.int32)
    perm = np.argsort(w, kind="stable")[:limit]
    return w[perm], rows[perm].astype(np.int64)


# ----------------------------------------------------------------- batched group-by
def group_batch(keys, vals, row_off, card):
    """Independent dense group-bys (the dense lane, groupby/dense.rs:18-37,
    with exact int64 sums of int32 values), compacted per query like the
    carry build (core/compact.rs:16-37).  Returns (ukeys, counts, sums,
    group_off) concatenated query by query, or -2 when a key is outside
    its query's range [0, card)."""
    keys = np.asarray(keys, np.int64)
    vals = np.asarray(vals, np.int64)
    uk, ct, sm, go = [], [], [], [0]
    for q in range(len(card)):
        k = keys[row_off[q]:row_off[q + 1]]
        v = vals[row_off[q]:row_off[q + 1]]
        if k.size and (k.min() < 0 or k.max() >= card[q]):
            return -2
        c = np.bincount(k, minlength=card[q]).astype(np.int64)
        s = np.zeros(card[q], np.int64)
        np.add.at(s, k, v)
        present = np.flatnonzero(c)
        uk.append(present.astype(np.int64))
        ct.append(c[present])
        sm.append(s[present])
        go.append(go[-1] + present.size)
    cat = (lambda a: np.concatenate(a) if a else np.zeros(0, np.int64))
    return cat(uk), cat(ct), cat(sm), np.array(go, np.int64)


def batch_workload(rng, nq, total_rows, card_lo=10, card_hi=10 ** 6, rows_lo=1e3,
                   rows_hi=1e7, s_hi=1.5, vlo=-(2 ** 15), vhi=2 ** 15):
    """Seeded shape of the batched configuration: row counts log-uniform in
    [rows_lo, rows_hi] rescaled to total_rows (floored at rows_lo), key
    cardinality log-uniform in [card_lo, card_hi], keys Zipf(s), s uniform
    in [0, s_hi] (s = 0 uniform), values uniform in [vlo, vhi)."""
    n = np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi), nq))
    n = np.maximum(np.floor(n * (total_rows / n.sum())), rows_lo).astype(np.int64)
    card = np.exp(rng.uniform(np.log(card_lo), np.log(card_hi), nq)).astype(np.int64)
    svals = rng.uniform(0, s_hi, nq)
    ks, vs = [], []
    for q in range(nq):
        c = int(card[q])
        w = np.arange(1, c + 1, dtype=np.float64) ** -svals[q]
        cdf = np.cumsum(w)
        cdf /= cdf[-1]
        r = np.minimum(np.searchsorted(cdf, rng.random(n[q]), side="right"), c - 1)
        ks.append(rng.permutation(c)[r].astype(np.int32))
        vs.append(rng.integers(vlo, vhi, n[q]).astype(np.int32))
    row_off = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    return np.concatenate(ks), np.concatenate(vs), row_off, card.astype(np.int32)
