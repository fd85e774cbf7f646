"""Grouped COUNT DISTINCT through the partitioned pair-hash entry points:
the host side of src/Drivers/GroupedHashMT/_lib/grouped_native_hash.py,
with the same geometry rules and call sequence (fingerprint -> per-chunk
counts -> caller prefix -> reservation scatter -> per-partition linear
probing tables -> occupancy count / fill -> adaptive unique), driving
libnumfast_hip.so (or any library with the same six symbols).

    ukeys, nunique, hash_ms, scan_ms = grouped_distinct_hash_native(k, v)

ukeys: the distinct keys, ascending; nunique[i]: distinct values of ukeys[i].
"""

import math
import os
import time

import numpy as np

_DENSE_CAP = 1 << 24       # grouped_native_hash.py:151
_ROWS_PER_PART = 1 << 18   # :152
_SLOTF = 1.5               # :153


class HashMiss(RuntimeError):
    """The lane cannot cover the input (the reference's _HashMiss)."""


def _next_pow2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


def partitions_for(n):
    """grouped_native_hash.py:199-201: power of two in [16, 256]."""
    return int(min(256, max(16, _next_pow2(max(1, (n + _ROWS_PER_PART - 1) // _ROWS_PER_PART)))))


def threads_for(p):
    """grouped_native_hash.py:204-212."""
    c = os.cpu_count() or 16
    return int(min(max(c, 1), p, 16))


def widen_i64(a):
    """grouped_native_hash.py:215-240: exact widening to int64."""
    a = np.ascontiguousarray(np.asarray(a)).ravel()
    if a.dtype.kind not in ("i", "u"):
        raise HashMiss("non-integer dtype %s" % (a.dtype,))
    if a.dtype == np.dtype(np.int64):
        return a
    if a.dtype.kind == "u" and a.dtype.itemsize == 8 and a.size and int(a.max()) > (1 << 63) - 1:
        raise HashMiss("uint64 past i64 range")
    return np.ascontiguousarray(a.astype(np.int64))


def adaptive_unique(ok):
    """grouped_native_hash.py:243-265: sorted (ukeys, counts)."""
    if ok.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    lo, hi = int(ok.min()), int(ok.max())
    if hi - lo <= _DENSE_CAP:
        bc = np.bincount((ok - lo).astype(np.int64), minlength=(hi - lo) + 1)
        nz = np.flatnonzero(bc)
        return (nz + lo).astype(np.int64), bc[nz].astype(np.int64)
    uk, cn = np.unique(ok, return_counts=True)
    return uk.astype(np.int64), cn.astype(np.int64)


def grouped_distinct_hash_native(kc, vc, nthreads=None, lib=None):
    """grouped_native_hash.py:268-359 (same buffers, same call order)."""
    if lib is None:
        from . import native
        lib = native.lib("ghash")
    K, V = widen_i64(kc), widen_i64(vc)
    n = int(K.size)
    if n != int(V.size):
        raise HashMiss("size mismatch")
    if n == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), 0.0, 0.0
    P = partitions_for(n)
    T = threads_for(P) if nthreads is None else int(max(1, min(int(nthreads), P, 64)))
    t0 = time.perf_counter()
    FP = np.empty(n, np.int64)
    _rc(lib.nf_ghash_fp_i64(K.ctypes.data, V.ctypes.data, n, FP.ctypes.data, P - 1, T), "fp")
    TC = np.zeros(T * P, np.int64)
    _rc(lib.nf_ghash_count_i64(FP.ctypes.data, n, TC.ctypes.data, T, P, T), "count")
    TC = TC.reshape(T, P)
    col = np.cumsum(TC, axis=0)
    counts = col[-1]
    starts = np.zeros(P + 1, np.int64)
    starts[1:] = np.cumsum(counts)
    tpos = np.empty((T, P), np.int64)
    tpos[0] = starts[:-1]
    if T > 1:
        tpos[1:] = starts[:-1] + col[:-1]
    PK, PV = np.empty(n, np.int64), np.empty(n, np.int64)
    _rc(lib.nf_ghash_scatter_i64(K.ctypes.data, V.ctypes.data, FP.ctypes.data, n,
                                 tpos.ctypes.data, T, P, PK.ctypes.data, PV.ctypes.data, T),
        "scatter")
    mp = np.array([_next_pow2(max(64, int(math.ceil(int(c) * _SLOTF)))) for c in counts],
                  np.int64)
    toff = np.zeros(P + 1, np.int64)
    toff[1:] = np.cumsum(mp)
    m = int(toff[-1])
    TK, TV, used = np.empty(m, np.int64), np.empty(m, np.int64), np.zeros(m, np.uint8)
    tm = np.ascontiguousarray(mp - 1)
    _rc(lib.nf_ghash_pins_i64(PK.ctypes.data, PV.ctypes.data, starts.ctypes.data,
                              toff.ctypes.data, tm.ctypes.data, TK.ctypes.data, TV.ctypes.data,
                              used.ctypes.data, P, T), "pins")
    hash_ms = (time.perf_counter() - t0) * 1e3
    t1 = time.perf_counter()
    CNT = np.empty(P, np.int64)
    _rc(lib.nf_ghash_occ_count(used.ctypes.data, toff.ctypes.data, P, CNT.ctypes.data, T),
        "occ_count")
    ostart = np.zeros(P + 1, np.int64)
    ostart[1:] = np.cumsum(CNT)
    nocc = int(ostart[-1])
    OK = np.empty(max(nocc, 1), np.int64)
    _rc(lib.nf_ghash_occ_fill(TK.ctypes.data, used.ctypes.data, toff.ctypes.data,
                              ostart.ctypes.data, OK.ctypes.data, P, T), "occ_fill")
    ukeys, nunique = adaptive_unique(OK[:nocc])
    return ukeys, nunique, hash_ms, (time.perf_counter() - t1) * 1e3


def _rc(rc, what):
    if rc != 0:
        raise HashMiss("%s rc=%d" % (what, rc))
