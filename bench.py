#!/usr/bin/env python3
"""bench.py -- NumFast's relational hot path on one MI355X.

A step is one pass of the hot path over one batch of synthetic input, the
high-cardinality group-by configuration: 2**28 rows of int32 keys drawn
Zipf(s=1.1) over 2**24 ids, int32 values uniform in [-2**15, 2**15), 5%
NULL values; group-by with sum/count/min/max/mean, then unique on the key
column with its own 5% NULL mask.  Inputs are generated on the GPU and are
resident in HBM before the timed region; outputs go to device buffers.

Also measured in the same run (reported under "extra"): the stable radix
argsort of 2**28 int32 keys (uniform full range) and the uniform-2**24
group-by variant.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: this build runs on one
GPU (the multi-GPU sharding is a design note only).
"""

import argparse
import ctypes
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

METRIC = ("rows/s + % HBM peak: group-sum & int32 argsort 2^28; learned/batched vs plain")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows-log2", type=int, default=28)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="only the group-by + unique step")
    ap.add_argument("--legs",
                    default="step,host,argsort,learned,batch,uniform,ghash,selector,cpu",
                    help="comma list of legs to run (for profiling one workload)")
    ap.add_argument("--selector-columns", type=int, default=1000,
                    help="labelled key columns of the strategy-selector leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (reference kernel for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE scaling)")
    return ap.parse_args()


def kernel_avg_ms(L, name):
    ms, cnt = ctypes.c_double(), ctypes.c_uint64()
    if L.nfh_profile_get(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)) != 0 or cnt.value == 0:
        return None, 0
    return ms.value / cnt.value, int(cnt.value)


def kernel_table(L):
    """Per kernel (HIP events on the library's stream): launches, average
    ms, algorithmic bytes per launch and the achieved GB/s."""
    buf = ctypes.create_string_buffer(8192)
    L.nfh_profile_names(buf, 8192)
    out = {}
    for name in buf.value.decode().split():
        ms, cnt, by = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
        if L.nfh_profile_get(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)) != 0:
            continue
        L.nfh_profile_bytes(name.encode(), ctypes.byref(by))
        c = max(1, cnt.value)
        row = {"launches": int(cnt.value), "avg_ms": round(ms.value / c, 4),
               "total_ms": round(ms.value, 4)}
        if by.value > 0 and ms.value > 0:
            row["alg_bytes_per_launch"] = by.value / c
            row["GBps"] = round(by.value / (ms.value * 1e-3) / 1e9, 1)
        out[name] = row
    return out


# HBM traffic per launch of a dominant kernel, from the rocprofv3 counter
# passes summarised in profiles/ (L2->memory read and write requests by
# size): profiled scope -> (summary file, kernel rows, how).  "mean": one
# kernel per launch of the scope, dispatch-weighted over the rows; "sum":
# every row runs once per launch of the scope.
TRAFFIC_ROWS = {
    "unique_part": ("groupby_unique_zipf_2p28.txt", ["k_umark_part"], "mean"),
    "unique_inv_dense": ("groupby_unique_zipf_2p28.txt", ["k_inv<true>"], "mean"),
    "gp_hist": ("groupby_unique_zipf_2p28.txt", ["k_gp_hist"], "mean"),
    # (the step's Zipf keys take the record mode: the first pass reads records)
    "gp_part_first": ("groupby_unique_zipf_2p28.txt",
                      ["k_gp_part<false, true, true, false, true>"], "mean"),
    "radix_onesweep": ("argsort_uniform_2p28.txt",
                       ["k_onesweep<int, true, false, true, true, RadixDigit>",
                        "k_onesweep<int, false, false, true, true, RadixDigit>",
                        "k_onesweep<int, false, true, true, false, RadixDigit>"], "mean"),
    # (the summary's name column holds 58 characters; the uniform keys run
    # the passes without run peeling, LearnedDigitT<false>)
    "learned_pass": ("learned_sort_2p28.txt",
                     ["k_onesweep<int, true, false, true, true, LearnedDigitT<fal",
                      "k_onesweep<int, false, false, true, false, LearnedDigitT<f"], "mean"),
    "ghash_pins": ("ghash_count_distinct_2p26.txt",
                   ["k_gh_whist", "k_gh_wscatter", "k_gh_wsolve", "k_gh_wfix"], "sum"),
}


def traffic_of(scope):
    """Bytes per launch (fetch + write) of the scope's kernels, or None."""
    if scope not in TRAFFIC_ROWS:
        return None
    fname, rows, how = TRAFFIC_ROWS[scope]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", fname)
    tot = calls = 0.0
    per = {}
    try:
        for line in open(path):
            name = line[:58].rstrip()
            if name in rows:
                parts = line[58:].split()
                c = float(parts[0])
                b = (float(parts[3]) + float(parts[4])) * 2 ** 20
                tot += c * b
                calls += c
                per[name] = b
    except (OSError, ValueError, IndexError):
        return None
    if how == "sum":
        return sum(per.values()) if len(per) == len(rows) else None
    return tot / calls if calls else None


def roofline_of(kernels, steps):
    """Roofline object for the kernel with the largest total time that
    carries an algorithmic byte count."""
    cands = [(v["total_ms"], k) for k, v in kernels.items() if "GBps" in v]
    if not cands:
        return None
    _, name = max(cands)
    v = kernels[name]
    tr = traffic_of(name)
    return {"bound": "hbm", "kernel": name, "achieved": v["GBps"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(v["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": None if tr is None else round(tr),
            "alg_bytes_per_launch": v["alg_bytes_per_launch"], "kernel_avg_ms": v["avg_ms"],
            "launches_per_step": round(v["launches"] / max(1, steps), 2)}


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def _zipf_host(rng, n, ids=1 << 24, s=1.1):
    """Bounded Zipf(s) keys over 2^24 ids (the workload generator's rule,
    on the host)."""
    import numpy as np
    ranks = np.arange(1, ids + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -s)
    cdf /= cdf[-1]
    r = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), ids - 1)
    return rng.permutation(ids).astype(np.int32)[r]


def _oracle():
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.executor import load_c_oracle
    lib, _ = load_c_oracle()
    return lib


def _reference_route_step(lib, keys, vals, valid, kvalid):
    """The reference's route for the group-by + unique step (hash lane ->
    the native radix unique, cpu.py:3508-3563, groupindex.py:178-211):
    NULL compaction, the C restatement of nf_unique_inverse_i32 for codes,
    np.bincount sums / counts, np.minimum.at / np.maximum.at, mean = sum /
    count; then unique of the key column over its valid rows."""
    import numpy as np

    def unique_c(k):
        n = k.size
        u, inv, perm = np.empty(n, np.int32), np.empty(n, np.int32), np.empty(n, np.int32)
        t0, t1, tp = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.int32)
        ng = lib.nf_unique_inverse_i32(k.ctypes.data, n, u.ctypes.data, inv.ctypes.data,
                                       perm.ctypes.data, t0.ctypes.data, t1.ctypes.data,
                                       tp.ctypes.data)
        return u[:ng], inv

    kk = np.ascontiguousarray(keys[valid])
    vv = vals[valid]
    u, inv = unique_c(kk)
    ng = u.size
    sums = np.bincount(inv, weights=vv.astype(np.int64), minlength=ng).astype(np.int64)
    counts = np.bincount(inv, minlength=ng).astype(np.int64)
    mins = np.full(ng, np.iinfo(np.int32).max, np.int32)
    maxs = np.full(ng, np.iinfo(np.int32).min, np.int32)
    np.minimum.at(mins, inv, vv)
    np.maximum.at(maxs, inv, vv)
    _ = sums.astype(np.float64) / counts
    del kk, vv, inv
    uk, _ = unique_c(np.ascontiguousarray(keys[kvalid]))
    return ng, uk.size


def cpu_baseline_full(host):
    """The headline workload's reference route at its full size (2^28 rows
    copied from the GPU's seeded inputs), one run, one thread."""
    lib = _oracle()
    t0 = time.perf_counter()
    ng, nu = _reference_route_step(lib, *host)
    t = time.perf_counter() - t0
    n = host[0].size
    return {"value": n / t, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": ("the full step: group-by sum/count/min/max/mean + unique, %d rows "
                       "(the same seeded Zipf(1.1) / 2^24-id / 5%%-NULL inputs, copied to the "
                       "host); C restatement of the reference radix unique (gcc -O3, one "
                       "thread) + NumPy bincount / minimum.at (reference route), one run of "
                       "%.1f s; ng %d / %d" % (n, t, ng, nu))}


def _median_time(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baselines_extra(torch, L, native, threads):
    """Reported CPU baselines beside the headline one (median of 5 unless
    stated): the NumPy executor on a config-2 sample, config 1 (the CPU
    configuration: NumPy min/max branch, and the native sum/count lanes in
    the executor's thread-pool slice-and-merge form), and the stable
    argsort, each beside this GPU's figure for the same call."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from oracle import executor as ox
    lib = _oracle()
    out = {"host_threads": threads}
    rng = np.random.default_rng(0)
    # config-2 sample through the NumPy executor restatement (np.unique +
    # bincount + minimum.at, cpu.py:2373-2601)
    n2 = 1 << 22
    k2 = _zipf_host(rng, n2)
    v2 = rng.integers(-(1 << 15), 1 << 15, n2).astype(np.int32)
    m2 = rng.random(n2) >= 0.05
    t = _median_time(lambda: ox.group_agg(k2, v2, m2), 3)
    out["config2_numpy_executor_2^22"] = {"rows_per_s": n2 / t, "ms": round(t * 1e3, 1),
                                          "cores": 1}
    # config 1: 2^20 rows, key U[0,1000), v_i32 U[-1000,1000], v_f64 U[0,1000)
    rng = np.random.default_rng(0)
    n1 = 1_048_576
    key = rng.integers(0, 1000, n1, dtype=np.int32)
    vi = rng.integers(-1000, 1001, n1, dtype=np.int32)
    vf = rng.uniform(0, 1000, n1)

    def numpy_query():
        ox.group_agg(key, vi)
        ox.group_agg(key, vf)
    t_np = _median_time(numpy_query, 5)

    def mt_lanes():
        # cpu.py:521-636: contiguous slices, one worker each, ordered merge
        bounds = np.linspace(0, n1, threads + 1).astype(np.int64)

        def work(w):
            a, b = int(bounds[w]), int(bounds[w + 1])
            kk, ff = np.ascontiguousarray(key[a:b]), np.ascontiguousarray(vf[a:b])
            s, c = np.zeros(1000), np.zeros(1000, np.int64)
            lib.nf_group_sum_count(kk.ctypes.data, ff.ctypes.data, kk.size, s.ctypes.data,
                                   c.ctypes.data, 1000)
            si, ci = np.zeros(1000, np.int32), np.zeros(1000, np.int64)
            ii = np.ascontiguousarray(vi[a:b])
            lib.nf_group_variant_i32(kk.ctypes.data, ii.ctypes.data, kk.size, 0, si.ctypes.data,
                                     ci.ctypes.data, 1000)
            return s, c, si
        with ThreadPoolExecutor(max_workers=threads) as ex:
            parts = list(ex.map(work, range(threads)))
        sf, cnt, si = np.zeros(1000), np.zeros(1000, np.int64), np.zeros(1000, np.int64)
        for s, c, i in parts:
            sf += s
            cnt += c
            si += i
    t_mt = _median_time(mt_lanes, 5)
    gi = lambda: (native.group_agg(key, vi), native.group_agg64(key, vf))
    t_gh = _median_time(gi, 5)
    kd, vid, vfd = (torch.from_numpy(x).cuda() for x in (key, vi, vf))
    outs = [torch.empty(1000, dtype=dt, device="cuda") for dt in
            (torch.int64, torch.int64, torch.int64, torch.int32, torch.int32, torch.float64)]
    outf = [torch.empty(1000, dtype=dt, device="cuda") for dt in
            (torch.int64, torch.int64, torch.float64, torch.float64, torch.float64,
             torch.float64)]

    def gpu_dev():
        L.nfh_group_agg_i32(kd.data_ptr(), vid.data_ptr(), None, None, n1,
                            *[o.data_ptr() for o in outs], 1000)
        L.nfh_group_agg_f64(kd.data_ptr(), vfd.data_ptr(), None, None, n1,
                            *[o.data_ptr() for o in outf], 1000)
    t_gd = timed(torch, gpu_dev, 20, 3)
    out["config1_2^20"] = {
        "numpy_executor_ms": round(t_np * 1e3, 2),
        "native_lanes_mt_ms": round(t_mt * 1e3, 2),
        "native_lanes_mt_note": "sum/count lanes only (the executor's min/max branch is NumPy)",
        "gpu_device_buffers_ms": round(t_gd * 1e3, 3),
        "gpu_host_buffers_ms": round(t_gh * 1e3, 2),
        "rows_per_s": {"numpy_executor": n1 / t_np, "native_lanes_mt": n1 / t_mt,
                       "gpu_device": n1 / t_gd, "gpu_host_buffers": n1 / t_gh}}
    # stable argsort: C restatement (one thread) and np.argsort(stable)
    na = 1 << 24
    ka = np.random.default_rng(0).integers(-(2 ** 31), 2 ** 31, na).astype(np.int32)
    perm = np.empty(na, np.int32)
    t0b, t1b, tpb = np.empty(na, np.uint32), np.empty(na, np.uint32), np.empty(na, np.int32)
    t_c = _median_time(lambda: lib.nf_sort_perm_i32(ka.ctypes.data, na, 0, perm.ctypes.data,
                                                    t0b.ctypes.data, t1b.ctypes.data,
                                                    tpb.ctypes.data), 3)
    t_a = _median_time(lambda: np.argsort(ka, kind="stable"), 3)
    out["argsort_2^24_uniform"] = {"c_restatement_rows_per_s": na / t_c,
                                   "numpy_stable_rows_per_s": na / t_a,
                                   "c_ms": round(t_c * 1e3, 1), "numpy_ms": round(t_a * 1e3, 1),
                                   "cores": 1}
    return out


def ghash_leg(torch, L, n, steps):
    """Grouped COUNT DISTINCT through the six nf_ghash_* calls on
    device-resident int64 (key, value) pairs: key uniform over 2^22 ids,
    value uniform over 64 (about 15 distinct values per key).  The driver's
    host steps (prefix sums, table geometry) run as torch ops on the device
    between the calls; the table geometry comes back to the host once per
    call, as in the reference driver.  CPU baseline: the oracle's C
    restatement through the same driver on a 2^22-pair sample, one thread."""
    import numpy as np
    from numfast_amd import ghash
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    k = torch.randint(0, 1 << 22, (n,), generator=g, device="cuda", dtype=torch.int64)
    v = torch.randint(0, 64, (n,), generator=g, device="cuda", dtype=torch.int64)
    P, T = ghash.partitions_for(n), 16
    fp = torch.empty(n, dtype=torch.int64, device="cuda")
    pk, pv = torch.empty_like(fp), torch.empty_like(fp)
    state = {}

    def run():
        assert L.nf_ghash_fp_i64(k.data_ptr(), v.data_ptr(), n, fp.data_ptr(), P - 1, T) == 0
        tc = torch.zeros(T * P, dtype=torch.int64, device="cuda")
        assert L.nf_ghash_count_i64(fp.data_ptr(), n, tc.data_ptr(), T, P, T) == 0
        col = tc.view(T, P).cumsum(0)
        counts = col[-1]
        starts = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
        starts[1:] = counts.cumsum(0)
        tpos = torch.empty(T, P, dtype=torch.int64, device="cuda")
        tpos[0] = starts[:-1]
        tpos[1:] = starts[:-1] + col[:-1]
        assert L.nf_ghash_scatter_i64(k.data_ptr(), v.data_ptr(), fp.data_ptr(), n,
                                      tpos.data_ptr(), T, P, pk.data_ptr(), pv.data_ptr(), T) == 0
        hc = counts.cpu().numpy()
        mp = np.array([ghash._next_pow2(max(64, int(np.ceil(int(c) * 1.5)))) for c in hc],
                      np.int64)
        toff_h = np.zeros(P + 1, np.int64)
        toff_h[1:] = np.cumsum(mp)
        m = int(toff_h[-1])
        toff = torch.from_numpy(toff_h).cuda()
        tm = torch.from_numpy(mp - 1).cuda()
        tk = torch.empty(m, dtype=torch.int64, device="cuda")
        tv = torch.empty_like(tk)
        used = torch.zeros(m, dtype=torch.uint8, device="cuda")
        assert L.nf_ghash_pins_i64(pk.data_ptr(), pv.data_ptr(), starts.data_ptr(),
                                   toff.data_ptr(), tm.data_ptr(), tk.data_ptr(), tv.data_ptr(),
                                   used.data_ptr(), P, T) == 0
        cnt = torch.empty(P, dtype=torch.int64, device="cuda")
        assert L.nf_ghash_occ_count(used.data_ptr(), toff.data_ptr(), P, cnt.data_ptr(), T) == 0
        ostart = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
        ostart[1:] = cnt.cumsum(0)
        nocc = int(ostart[-1])
        ok = torch.empty(max(nocc, 1), dtype=torch.int64, device="cuda")
        assert L.nf_ghash_occ_fill(tk.data_ptr(), used.data_ptr(), toff.data_ptr(),
                                   ostart.data_ptr(), ok.data_ptr(), P, T) == 0
        state["nocc"] = nocc
        state["ok"] = ok[:nocc]

    run()  # warm-up (scratch growth) outside the per-kernel averages
    torch.cuda.synchronize()
    L.nfh_profile_reset()
    L.nfh_profile_enable(1)
    t = timed(torch, run, steps, 0)
    L.nfh_profile_enable(0)
    kt = kernel_table(L)
    # exactness at this size: distinct pairs == the occupied entries
    code = torch.unique(k * 64 + v)
    same = int(code.numel()) == state["nocc"]
    uk = torch.unique(state["ok"])
    same = same and torch.equal(uk, torch.unique(code // 64))
    ghk = {a: b for a, b in kt.items() if a.startswith("ghash")}
    # CPU baseline: the oracle through the same driver, 2^22 pairs
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.executor import load_c_oracle
    ora, _ = load_c_oracle()
    ns = 1 << 22
    ks, vs = k[:ns].cpu().numpy(), v[:ns].cpu().numpy()
    t0 = time.perf_counter()
    ghash.grouped_distinct_hash_native(ks, vs, nthreads=1, lib=ora)
    tc_cpu = time.perf_counter() - t0
    return {"pairs": n, "partitions": P, "threads_geometry": T, "distinct_pairs": state["nocc"],
            "ms": round(t * 1e3, 3), "pairs_per_s": n / t, "checked_against_torch_unique": same,
            "kernels": ghk, "roofline": roofline_of(ghk, steps),
            "cpu_baseline": {"value": ns / tc_cpu, "unit": "pairs/s", "cores": 1, "kind": "port",
                             "sample": "first 2^22 pairs, oracle C restatement through the "
                                       "same driver (nthreads=1)"}}


def sort_pipeline_leg(torch, L, W, native, n, steps):
    """Per distribution: plain vs learned argsort (model trained once on the
    column, training time reported apart) and the filter -> derive -> sort
    -> limit(1000) pipeline with each sort; outputs compared for identity."""
    import ctypes
    out = {}
    perm0 = torch.empty(n, dtype=torch.int32, device="cuda")
    perm1 = torch.empty_like(perm0)
    t0b, t1b, tpb = (torch.empty_like(perm0) for _ in range(3))
    lim = 1000
    w0 = torch.empty(lim, dtype=torch.int32, device="cuda")
    r0, w1, r1 = torch.empty_like(w0), torch.empty_like(w0), torch.empty_like(w0)
    for dist in ("uniform", "swaps", "distinct1024"):
        k = W.sort_config(n, dist)
        torch.cuda.synchronize()
        model = native.CdfModel()
        loss = ctypes.c_float()
        t = time.perf_counter()
        rc = L.nfh_cdf_train_i32(k.data_ptr(), n, 100, 0, ctypes.byref(model), ctypes.byref(loss))
        train_s = time.perf_counter() - t
        if rc != 0:
            raise RuntimeError("train rc %d" % rc)
        stats = (ctypes.c_uint64 * 4)()

        def plain():
            if L.nf_sort_perm_i32(k.data_ptr(), n, 0, perm0.data_ptr(), t0b.data_ptr(),
                                  t1b.data_ptr(), tpb.data_ptr()) != 0:
                raise RuntimeError("sort")

        def learned(st=None):
            if L.nfh_sort_perm_learned_i32(k.data_ptr(), n, ctypes.byref(model), perm1.data_ptr(),
                                           st) != 0:
                raise RuntimeError("learned sort")
        # bucket statistics (a host readback of the bucket bounds) from an
        # untimed call; the timed calls ask for none
        learned(stats)
        tp = timed(torch, plain, steps, 1)
        learned()  # warm-up outside the per-kernel averages
        torch.cuda.synchronize()
        L.nfh_profile_reset()
        L.nfh_profile_enable(1)
        tl = timed(torch, learned, steps, 0)
        L.nfh_profile_enable(0)
        lk = kernel_table(L)
        same = bool(torch.equal(perm0, perm1))
        # pipeline: model for the derived column w = v * 3 (int32 wrap)
        wcol = (k[k > 0].to(torch.int64) * 3).to(torch.int32)
        pmodel = native.CdfModel()
        if L.nfh_cdf_train_i32(wcol.data_ptr(), wcol.numel(), 100, 0, ctypes.byref(pmodel),
                               None) != 0:
            raise RuntimeError("train rc")
        del wcol

        def pipe(m, wo, ro):
            def f():
                r = L.nfh_filter_derive_sort_limit_i32(k.data_ptr(), n, 3, lim, m, wo.data_ptr(),
                                                       ro.data_ptr())
                if r < 0:
                    raise RuntimeError("pipeline rc %d" % r)
            return f
        tpp = timed(torch, pipe(None, w0, r0), steps, 1)
        tpl = timed(torch, pipe(ctypes.byref(pmodel), w1, r1), steps, 1)
        out[dist] = {
            "plain_ms": round(tp * 1e3, 3), "learned_ms": round(tl * 1e3, 3),
            "speedup": round(tp / tl, 3), "identical_perm": same,
            "plain_rows_per_s": n / tp, "learned_rows_per_s": n / tl,
            "plain_frac_60B_model": round(60.0 * n / tp / 1e9 / HBM_PEAK_GBS, 4),
            "learned_frac_44B_model": round(44.0 * n / tl / 1e9 / HBM_PEAK_GBS, 4),
            "train_ms": round(train_s * 1e3, 2), "train_loss": float(loss.value),
            "buckets": [int(x) for x in stats], "learned_kernels": lk,
            "pipeline_plain_ms": round(tpp * 1e3, 3), "pipeline_learned_ms": round(tpl * 1e3, 3),
            "pipeline_speedup": round(tpp / tpl, 3),
            "pipeline_identical": bool(torch.equal(w0, w1) and torch.equal(r0, r1))}
        del k
    return out


def batch_leg(torch, W, native, n, steps):
    """4096 independent sum/count group-bys over n rows in total (config
    shape: row counts log-uniform in [1e3, 1e7] rescaled, cardinality
    log-uniform in [10, 1e6], Zipf s in [0, 1.5]): one hipGraph replay of
    the persistent batch vs one launch set per query.  Algorithmic bytes:
    sum_q n_q * 8 + G_q * 24 (G_q present groups)."""
    keys, vals, row_off, card, scale = W.batch_config(nq=4096, total=n)
    N = keys.numel()
    t = time.perf_counter()
    b = native.GroupBatch(torch, keys, vals, row_off, card)
    plan_s = time.perf_counter() - t
    ng0, uk0, ct0, sm0, go0 = b.run(0)
    ref = [x.clone() for x in (uk0, ct0, sm0, go0)]
    ng1, uk1, ct1, sm1, go1 = b.run(1)
    same = (ng0 == ng1 and all(torch.equal(a, c) for a, c in zip(ref, (uk1, ct1, sm1, go1))))
    ng2, uk2, ct2, sm2, go2 = b.run(2)
    same = same and (ng0 == ng2 and all(torch.equal(a, c)
                                        for a, c in zip(ref, (uk2, ct2, sm2, go2))))
    tg = timed(torch, lambda: b.run(0), steps, 1)
    tq = timed(torch, lambda: b.run(1), max(2, steps // 2), 1)
    tl = timed(torch, lambda: b.run(2), max(2, steps // 2), 1)
    alg = 8.0 * N + 24.0 * ng0
    out = {"queries": 4096, "rows": N, "row_scale": scale, "groups": ng0,
           "graph_ms": round(tg * 1e3, 3),
           "plain_library_per_query_ms": round(tl * 1e3, 3),
           "speedup": round(tl / tg, 3),
           "speedup_note": "vs the plain form: each query through the library's own group-by "
                           "(nfh_group_agg_i32 device path, one call per query)",
           "same_kernels_per_query_ms": round(tq * 1e3, 3),
           "speedup_vs_same_kernels_per_query": round(tq / tg, 3), "identical": bool(same),
           "graph_rows_per_s": N / tg, "alg_GBps": round(alg / tg / 1e9, 1),
           "frac_of_peak": round(alg / tg / 1e9 / HBM_PEAK_GBS, 4),
           "plan_ms": round(plan_s * 1e3, 1), "autotuned": b.info()}
    b.close()
    del keys, vals
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; the multi-GPU sharding of the "
              "group-by / sort is a design note only.  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch  # first: the library then shares torch's HIP runtime
    torch.cuda.init()
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from numfast_amd import native
    from numfast_amd import workloads as W
    L = native.lib("hip")
    n = 1 << args.rows_log2
    ids = 1 << 24
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    keys, vals, valid, kvalid = W.groupby_config(n=n, ids=ids)
    cap = ids
    outs = [torch.empty(cap, dtype=dt, device="cuda") for dt in
            (torch.int64, torch.int64, torch.int64, torch.int32, torch.int32, torch.float64)]
    uniq = torch.empty(ids, dtype=torch.int32, device="cuda")
    inv = torch.empty(n, dtype=torch.int32, device="cuda")
    state = {}

    def step_on(k, v, m, km):
        def step():
            ng = L.nfh_group_agg_i32(k.data_ptr(), v.data_ptr(), m.data_ptr(), None, n,
                                     *[o.data_ptr() for o in outs], cap)
            nu = L.nfh_unique_valid_i32(k.data_ptr(), km.data_ptr(), n, uniq.data_ptr(),
                                        inv.data_ptr())
            if ng < 0 or nu < 0:
                raise RuntimeError("library rc %d / %d" % (ng, nu))
            state["ng"], state["nu"] = ng, nu
        return step

    legs = {"step"} if args.no_extra else set(args.legs.split(","))
    extra = {}
    t, roofline = None, None
    torch.cuda.synchronize()
    if "step" in legs:
        L.nfh_profile_enable(0)
        step = step_on(keys, vals, valid, kvalid)
        for _ in range(args.warmup):
            step()
        L.nfh_profile_reset()
        L.nfh_profile_enable(1)
        t = timed(torch, step, args.steps, 0)
        L.nfh_profile_enable(0)
        kernels = kernel_table(L)
        ng, nu = state["ng"], state["nu"]
        roofline = roofline_of(kernels, args.steps)
        # whole-step algorithmic bytes: group-by 9 B/row + 40 B/group,
        # unique 5 B/row read + 4 B/row inverse + 4 B per distinct key
        step_bytes = 9.0 * n + 40.0 * ng + 9.0 * n + 4.0 * nu
        extra.update({"groupby_ng": ng, "unique_ng": nu, "kernels": kernels,
                      "step_alg_GBps": round(step_bytes / t / 1e9, 1),
                      "step_alg_frac": round(step_bytes / t / 1e9 / HBM_PEAK_GBS, 4)})
    host = None
    if not args.no_cpu or "host" in legs:
        host = [keys.cpu().numpy(), vals.cpu().numpy(), valid.cpu().numpy().astype(bool),
                kvalid.cpu().numpy().astype(bool)]
    del keys, vals, valid, kvalid
    if "host" in legs and host is not None:
        # the boundary as NumFast calls it: NumPy host buffers, staged over
        # PCIe by the library (never the reported value)
        import numpy as np
        kh, vh = host[0], host[1]
        mh, kmh = host[2].astype(np.uint8), host[3].astype(np.uint8)
        oh = [np.empty(ids, dt) for dt in (np.int64, np.int64, np.int64, np.int32, np.int32,
                                           np.float64)]
        uh, ih = np.empty(ids, np.int32), np.empty(n, np.int32)

        def host_step():
            a = L.nfh_group_agg_i32(kh.ctypes.data, vh.ctypes.data, mh.ctypes.data, None, n,
                                    *[o.ctypes.data for o in oh], ids)
            b = L.nfh_unique_valid_i32(kh.ctypes.data, kmh.ctypes.data, n, uh.ctypes.data,
                                       ih.ctypes.data)
            if a < 0 or b < 0:
                raise RuntimeError("host-buffer step rc %d / %d" % (a, b))
        th = _median_time(host_step, 3)
        extra["step_host_buffers"] = {"ms": round(th * 1e3, 1), "rows_per_s": n / th,
                                      "note": "PCIe-inclusive: NumPy host inputs and outputs "
                                              "staged by the library"}
        del oh, uh, ih, mh, kmh

    if "argsort" in legs:
        # stable argsort, 2^28 uniform int32 (pass model 60 B/row)
        k = W.sort_config(n, "uniform")
        perm = torch.empty(n, dtype=torch.int32, device="cuda")
        t0b = torch.empty(n, dtype=torch.int32, device="cuda")
        t1b, tpb = torch.empty_like(t0b), torch.empty_like(t0b)

        def sort_step():
            rc = L.nf_sort_perm_i32(k.data_ptr(), n, 0, perm.data_ptr(), t0b.data_ptr(),
                                    t1b.data_ptr(), tpb.data_ptr())
            if rc != 0:
                raise RuntimeError("sort rc %d" % rc)
        sort_step()  # warm-up outside the per-kernel averages
        torch.cuda.synchronize()
        L.nfh_profile_reset()
        L.nfh_profile_enable(1)
        sort_steps = max(3, args.steps // 2)
        ts = timed(torch, sort_step, sort_steps, 0)
        L.nfh_profile_enable(0)
        sk = kernel_table(L)
        extra["argsort_uniform_2^28"] = {
            "rows_per_s": n / ts, "ms": round(ts * 1e3, 3),
            "pass_model_GBps": round(60.0 * n / ts / 1e9, 1),
            "frac_of_peak_60B_model": round(60.0 * n / ts / 1e9 / HBM_PEAK_GBS, 4),
            "compulsory_frac_8B": round(8.0 * n / ts / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": sk, "roofline": roofline_of(sk, sort_steps)}
        del k, perm, t0b, t1b, tpb
    if "learned" in legs:
        # config-3 shapes: plain LSD vs learned-CDF sort, and the
        # filter(v>0) -> derive(v*3) -> sort -> limit(1000) pipeline
        extra["sort_pipeline_2^28"] = sort_pipeline_leg(torch, L, W, native, n,
                                                        max(3, args.steps // 2))
    if "batch" in legs:
        # batched group-by: 4096 queries, one graph-captured persistent launch
        # vs the same kernels launched query by query
        extra["batched_groupby_4096q"] = batch_leg(torch, W, native, n, max(3, args.steps // 2))
    if "selector" in legs:
        # learned strategy selector on a reduced column set (larger runs:
        # python -m numfast_amd.selector)
        from numfast_amd import selector as SEL
        # every plain strategy forced on each seeded column of 2^18-2^22 rows;
        # trained on 80%, reported on the held-out 20%
        t0s = time.perf_counter()
        X, T, ran, _ = SEL.label_columns(torch, native, args.selector_columns, lo=18, hi=22,
                                         seed=7)
        rep, _ = SEL.evaluate(native, X, T, ran, steps=400)
        rep["label_and_train_s"] = round(time.perf_counter() - t0s, 1)
        extra["strategy_selector"] = rep
    if "ghash" in legs:
        # partitioned pair-hash (grouped COUNT DISTINCT), 2^26 int64 pairs
        extra["ghash_count_distinct_2^26"] = ghash_leg(torch, L, 1 << 26, max(3, args.steps // 2))
    if "cpu" in legs and not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        extra["cpu_baselines"] = cpu_baselines_extra(torch, L, native, threads)
    if "uniform" in legs:
        # uniform-2^24 group-by variant
        keys, vals, valid, kvalid = W.groupby_config(n=n, ids=ids, dist="uniform")
        tu = timed(torch, step_on(keys, vals, valid, kvalid), max(3, args.steps // 2), 1)
        extra["groupby_unique_uniform_2^24"] = {"rows_per_s": n / tu, "ms": round(tu * 1e3, 3),
                                                "groupby_ng": state["ng"]}
        del keys, vals, valid, kvalid

    out = {
        "metric": METRIC,
        "value": None if t is None else n / t,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None if t is None else round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": "groupby(sum,count,min,max,mean)+unique, 2^%d rows, int32 keys "
                               "Zipf(1.1) over 2^24 ids, int32 values U[-2^15,2^15), 5%% NULL "
                               "values, 5%% NULL keys for unique" % args.rows_log2,
                   "rows": n, "ids": ids, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu else cpu_baseline_full(host),
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
