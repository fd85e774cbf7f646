/*
 * nf_oracle.c -- CPU restatement of numfast-native's relational hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py may load it; the product
 * library (libnumfast_hip.so) never links, calls or routes through it.
 *
 * It exports the SAME C-ABI symbol names and signatures as the reference
 * crate (numfast-native/src/lib.rs), so one ctypes harness can drive either
 * this oracle or the HIP library and compare results byte for byte.
 * Every function restates the reference algorithm in plain sequential C
 * (row order, single thread) and cites the file:line it follows.
 *
 * Pinning: tests/test_oracle_*.py check every function here against the
 * reference's own golden vectors (tests/golden/*.json), against NumPy
 * library routines (np.argsort(kind="stable"), np.unique, np.bincount,
 * np.add.reduceat, np.minimum/maximum.at) and against brute force on tiny
 * inputs (rc precedence, prefix overflow).
 *
 * Return codes follow numfast-native/src/core/errors.rs:13-38.  Where the
 * reference panics (process abort, Cargo.toml panic="abort") on an
 * out-of-range dense key, this restatement returns -2 instead; tests never
 * feed the panic inputs to the reference-order comparisons.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define NF_OK 0
#define NF_NULL_OR_ABORT (-1)
#define NF_BAD_RANGE (-2)
#define NF_MALFORMED (-3)
#define NF_BAD_VARIANT (-3)
#define NF_OVERFLOW (-4)
#define NF_UNSORTED (-5)

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* checked adds: numfast-native/src/series/reduce.rs:24-45             */
static int add_i32_checked(int32_t *dst, int32_t v) {
    int32_t r;
    if (__builtin_add_overflow(*dst, v, &r)) return 0;
    *dst = r;
    return 1;
}
static int add_i64_checked(int64_t *dst, int64_t v) {
    int64_t r;
    if (__builtin_add_overflow(*dst, v, &r)) return 0;
    *dst = r;
    return 1;
}

/* ------------------------------------------------------------------ */
/* stable LSD radix argsort: numfast-native/src/sort.rs:22-118          */
/* One counting-scatter pass on byte `shift` (sort.rs:22-47).           */
static void lsd_pass_u32(const uint32_t *cw, const int32_t *cp, uint32_t *nw,
                         int32_t *np_, size_t n, unsigned shift) {
    size_t cnt[256];
    memset(cnt, 0, sizeof cnt);
    for (size_t i = 0; i < n; i++) cnt[(cw[i] >> shift) & 0xFFu]++;
    size_t run = 0;
    for (int b = 0; b < 256; b++) {
        size_t c = cnt[b];
        cnt[b] = run;
        run += c;
    }
    for (size_t i = 0; i < n; i++) {
        size_t d = cnt[(cw[i] >> shift) & 0xFFu]++;
        nw[d] = cw[i];
        np_[d] = cp[i];
    }
}
static void lsd_pass_u64(const uint64_t *cw, const int32_t *cp, uint64_t *nw,
                         int32_t *np_, size_t n, unsigned shift) {
    size_t cnt[256];
    memset(cnt, 0, sizeof cnt);
    for (size_t i = 0; i < n; i++) cnt[(cw[i] >> shift) & 0xFFu]++;
    size_t run = 0;
    for (int b = 0; b < 256; b++) {
        size_t c = cnt[b];
        cnt[b] = run;
        run += c;
    }
    for (size_t i = 0; i < n; i++) {
        size_t d = cnt[(cw[i] >> shift) & 0xFFu]++;
        nw[d] = cw[i];
        np_[d] = cp[i];
    }
}

/* sort.rs:55-84: sign-flip word, `!w` for descending, 4 passes, ping-pong
 * A=(tmp0, perm_out) / B=(tmp1, tmp_p) so the even pass count lands the
 * permutation in perm_out. */
static void sort_perm_i32(const int32_t *keys, size_t n, int desc, int32_t *perm,
                          uint32_t *t0, uint32_t *t1, int32_t *tp) {
    for (size_t i = 0; i < n; i++) {
        uint32_t w = ((uint32_t)keys[i]) ^ 0x80000000u;
        t0[i] = desc ? ~w : w;
        perm[i] = (int32_t)i;
    }
    for (unsigned pass = 0; pass < 4; pass++) {
        if ((pass & 1u) == 0)
            lsd_pass_u32(t0, perm, t1, tp, n, pass * 8u);
        else
            lsd_pass_u32(t1, tp, t0, perm, n, pass * 8u);
    }
}
/* sort.rs:89-118: int64 keys, 8 passes. */
static void sort_perm_i64(const int64_t *keys, size_t n, int desc, int32_t *perm,
                          uint64_t *t0, uint64_t *t1, int32_t *tp) {
    for (size_t i = 0; i < n; i++) {
        uint64_t w = ((uint64_t)keys[i]) ^ 0x8000000000000000ull;
        t0[i] = desc ? ~w : w;
        perm[i] = (int32_t)i;
    }
    for (unsigned pass = 0; pass < 8; pass++) {
        if ((pass & 1u) == 0)
            lsd_pass_u64(t0, perm, t1, tp, n, pass * 8u);
        else
            lsd_pass_u64(t1, tp, t0, perm, n, pass * 8u);
    }
}

/* lib.rs:1326-1357 */
EXPORT int32_t nf_sort_perm_i32(const int32_t *keys, size_t n, uint8_t descending,
                                int32_t *perm_out, uint32_t *tmp0, uint32_t *tmp1,
                                int32_t *tmp_p) {
    if (n == 0) return NF_OK;
    if (!keys || !perm_out || !tmp0 || !tmp1 || !tmp_p) return NF_NULL_OR_ABORT;
    if (n > (size_t)INT32_MAX) return NF_MALFORMED;
    sort_perm_i32(keys, n, descending != 0, perm_out, tmp0, tmp1, tmp_p);
    return NF_OK;
}
/* lib.rs:1363-1394 */
EXPORT int32_t nf_sort_perm_i64(const int64_t *keys, size_t n, uint8_t descending,
                                int32_t *perm_out, uint64_t *tmp0, uint64_t *tmp1,
                                int32_t *tmp_p) {
    if (n == 0) return NF_OK;
    if (!keys || !perm_out || !tmp0 || !tmp1 || !tmp_p) return NF_NULL_OR_ABORT;
    if (n > (size_t)INT32_MAX) return NF_MALFORMED;
    sort_perm_i64(keys, n, descending != 0, perm_out, tmp0, tmp1, tmp_p);
    return NF_OK;
}

/* ------------------------------------------------------------------ */
/* run-encode over an ascending permutation: unique.rs:30-58 / :63-91   */
static int64_t run_encode_i32(const int32_t *keys, const int32_t *perm, size_t n,
                              int32_t *uniq, int32_t *inv) {
    if (n == 0) return 0;
    int32_t prev = keys[perm[0]];
    uniq[0] = prev;
    inv[perm[0]] = 0;
    int64_t ng = 1;
    for (size_t r = 1; r < n; r++) {
        int32_t row = perm[r];
        int32_t k = keys[row];
        if (k != prev) {
            prev = k;
            uniq[ng++] = k;
        }
        inv[row] = (int32_t)(ng - 1);
    }
    return ng;
}
static int64_t run_encode_i64(const int64_t *keys, const int32_t *perm, size_t n,
                              int64_t *uniq, int32_t *inv) {
    if (n == 0) return 0;
    int64_t prev = keys[perm[0]];
    uniq[0] = prev;
    inv[perm[0]] = 0;
    int64_t ng = 1;
    for (size_t r = 1; r < n; r++) {
        int32_t row = perm[r];
        int64_t k = keys[row];
        if (k != prev) {
            prev = k;
            uniq[ng++] = k;
        }
        inv[row] = (int32_t)(ng - 1);
    }
    return ng;
}

/* lib.rs:1420-1456 */
EXPORT int64_t nf_unique_inverse_i32(const int32_t *keys, size_t n, int32_t *uniq_out,
                                     int32_t *inv_out, int32_t *perm, uint32_t *tmp0,
                                     uint32_t *tmp1, int32_t *tmp_p) {
    if (n == 0) return 0;
    if (!keys || !uniq_out || !inv_out || !perm || !tmp0 || !tmp1 || !tmp_p) return -2;
    if (n > (size_t)INT32_MAX) return NF_MALFORMED;
    sort_perm_i32(keys, n, 0, perm, tmp0, tmp1, tmp_p);
    return run_encode_i32(keys, perm, n, uniq_out, inv_out);
}
/* lib.rs:1464-1500 */
EXPORT int64_t nf_unique_inverse_i64(const int64_t *keys, size_t n, int64_t *uniq_out,
                                     int32_t *inv_out, int32_t *perm, uint64_t *tmp0,
                                     uint64_t *tmp1, int32_t *tmp_p) {
    if (n == 0) return 0;
    if (!keys || !uniq_out || !inv_out || !perm || !tmp0 || !tmp1 || !tmp_p) return -2;
    if (n > (size_t)INT32_MAX) return NF_MALFORMED;
    sort_perm_i64(keys, n, 0, perm, tmp0, tmp1, tmp_p);
    return run_encode_i64(keys, perm, n, uniq_out, inv_out);
}

/* ------------------------------------------------------------------ */
/* dense f64 scatter in row order: groupby/dense.rs:18-37.  The reference
 * panics on an out-of-range key; the restatement reports -2.          */
EXPORT int32_t nf_group_sum_count(const int32_t *keys, const double *values, size_t n,
                                  double *sums, int64_t *counts, size_t g) {
    if (!keys || !values || !sums || !counts) return NF_NULL_OR_ABORT;
    for (size_t j = 0; j < g; j++) { sums[j] = 0.0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        sums[k] += values[i];
        counts[k] += 1;
    }
    return NF_OK;
}

/* SoA multi-column, range-checked: groupby/fused.rs:25-50 */
static int32_t fused_soa(const int32_t *keys, const double *values, size_t n, size_t ncols,
                         double *sums, int64_t *counts, size_t g) {
    for (size_t j = 0; j < ncols * g; j++) sums[j] = 0.0;
    for (size_t j = 0; j < g; j++) counts[j] = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        counts[k] += 1;
        for (size_t c = 0; c < ncols; c++) sums[c * g + k] += values[c * n + i];
    }
    return NF_OK;
}
/* lib.rs:117-135 */
EXPORT int32_t nf_group_multi_sum_count(const int32_t *keys, const double *values, size_t n,
                                        size_t ncols, double *sums, int64_t *counts,
                                        size_t g) {
    if (!keys || !values || !sums || !counts) return NF_NULL_OR_ABORT;
    return fused_soa(keys, values, n, ncols, sums, counts, g);
}

/* groupby/mixed.rs:42-74 + lib.rs:150-178 */
EXPORT int32_t nf_group_mixed_sum_only(const int32_t *keys, const int32_t *i32_vals,
                                       const double *f64_vals, size_t n, size_t n_i32,
                                       size_t n_f64, int32_t *i32_sums, double *f64_sums,
                                       size_t g) {
    if (!keys) return NF_NULL_OR_ABORT;
    if ((n_i32 > 0 && (!i32_vals || !i32_sums)) || (n_f64 > 0 && (!f64_vals || !f64_sums)))
        return NF_NULL_OR_ABORT;
    for (size_t j = 0; j < n_i32 * g; j++) i32_sums[j] = 0;
    for (size_t j = 0; j < n_f64 * g; j++) f64_sums[j] = 0.0;
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        for (size_t c = 0; c < n_i32; c++)
            if (!add_i32_checked(&i32_sums[c * g + k], i32_vals[c * n + i])) return NF_OVERFLOW;
        for (size_t c = 0; c < n_f64; c++) f64_sums[c * g + k] += f64_vals[c * n + i];
    }
    return NF_OK;
}

/* groupby/mixed.rs:79-90 + lib.rs:184-197 */
EXPORT int32_t nf_group_count_only(const int32_t *keys, size_t n, int64_t *counts, size_t g) {
    if (!keys || !counts) return NF_NULL_OR_ABORT;
    for (size_t j = 0; j < g; j++) counts[j] = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        counts[k] += 1;
    }
    return NF_OK;
}

/* groupby/mixed.rs:106-149 + lib.rs:213-250 */
EXPORT int32_t nf_group_owner_2i32_1f64(const int32_t *keys, const int32_t *a,
                                        const int32_t *b, const double *c, size_t n,
                                        size_t lo, size_t hi, int32_t *s1, int32_t *s2,
                                        double *s3, int64_t *counts, size_t g) {
    if (!keys || !a || !b || !c || !s1 || !s2 || !s3 || !counts || lo > hi || hi > g)
        return NF_NULL_OR_ABORT;
    for (size_t j = lo; j < hi; j++) { s1[j] = 0; s2[j] = 0; s3[j] = 0.0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        if (k >= lo && k < hi) {
            counts[k] += 1;
            if (!add_i32_checked(&s1[k], a[i])) return NF_OVERFLOW;
            if (!add_i32_checked(&s2[k], b[i])) return NF_OVERFLOW;
            s3[k] += c[i];
        }
    }
    return NF_OK;
}

/* ------------------------------------------------------------------ */
/* int families: groupby/scaled.rs:27-254, dispatch variants.rs:45-101  */
static int32_t checked_scatter_i64(const int32_t *keys, const int64_t *t, size_t n,
                                   int64_t *sums, int64_t *counts, size_t g) {
    for (size_t j = 0; j < g; j++) { sums[j] = 0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        if (!add_i64_checked(&sums[k], t[i])) return NF_OVERFLOW;
        counts[k] += 1;
    }
    return NF_OK;
}
static int32_t checked_scatter_i32(const int32_t *keys, const int32_t *t, size_t n,
                                   int32_t *sums, int64_t *counts, size_t g) {
    for (size_t j = 0; j < g; j++) { sums[j] = 0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)(int64_t)keys[i];
        if (k >= g) return NF_BAD_RANGE;
        if (!add_i32_checked(&sums[k], t[i])) return NF_OVERFLOW;
        counts[k] += 1;
    }
    return NF_OK;
}
/* scaled.rs:113-156 (i64), :162-205 (i32), :213-254 (f64): run traversal
 * over nondecreasing keys; the restatement adds a -2 where the reference
 * would index out of bounds. */
#define SORTED_DENSE_BODY(ACC_T, ADD_OK)                                        \
    for (size_t j = 0; j < g; j++) { sums[j] = 0; counts[j] = 0; }               \
    if (n == 0) return NF_OK;                                                   \
    int32_t cur = keys[0];                                                      \
    ACC_T acc = vals[0];                                                        \
    int64_t cnt = 1;                                                            \
    for (size_t i = 1; i < n; i++) {                                            \
        int32_t k = keys[i];                                                    \
        if (k < cur) return NF_UNSORTED;                                        \
        if (k != cur) {                                                         \
            if ((uint64_t)(int64_t)cur >= g) return NF_BAD_RANGE;               \
            sums[cur] = acc;                                                    \
            counts[cur] = cnt;                                                  \
            cur = k;                                                            \
            acc = vals[i];                                                      \
            cnt = 1;                                                            \
        } else {                                                                \
            if (!(ADD_OK)) return NF_OVERFLOW;                                  \
            cnt += 1;                                                           \
        }                                                                       \
    }                                                                           \
    if ((uint64_t)(int64_t)cur >= g) return NF_BAD_RANGE;                       \
    sums[cur] = acc;                                                            \
    counts[cur] = cnt;                                                          \
    return NF_OK;

static int32_t sorted_dense_i64(const int32_t *keys, const int64_t *vals, size_t n,
                                int64_t *sums, int64_t *counts, size_t g) {
    SORTED_DENSE_BODY(int64_t, add_i64_checked(&acc, vals[i]))
}
static int32_t sorted_dense_i32(const int32_t *keys, const int32_t *vals, size_t n,
                                int32_t *sums, int64_t *counts, size_t g) {
    SORTED_DENSE_BODY(int32_t, add_i32_checked(&acc, vals[i]))
}
static int f64_add(double *acc, double v) { *acc += v; return 1; }
static int32_t sorted_dense_f64(const int32_t *keys, const double *vals, size_t n,
                                double *sums, int64_t *counts, size_t g) {
    SORTED_DENSE_BODY(double, f64_add(&acc, vals[i]))
}

/* lib.rs:539-557 + variants.rs:45-62 */
EXPORT int32_t nf_group_variant_f64(const int32_t *keys, const double *values, size_t n,
                                    uint32_t variant, double *sums, int64_t *counts,
                                    size_t g) {
    if (!keys || !values || !sums || !counts) return NF_NULL_OR_ABORT;
    switch (variant) {
    case 0: return nf_group_sum_count(keys, values, n, sums, counts, g);
    case 1: return fused_soa(keys, values, n, 1, sums, counts, g);
    case 2: return sorted_dense_f64(keys, values, n, sums, counts, g);
    default: return NF_BAD_VARIANT;
    }
}
/* lib.rs:569-587 + variants.rs:68-82; V_DENSE = dense_scatter_i64
 * (scaled.rs:27-47) which differs from V_CHECKED only in the missing
 * range check. */
EXPORT int32_t nf_group_variant_i64(const int32_t *keys, const int64_t *ticks, size_t n,
                                    uint32_t variant, int64_t *sums, int64_t *counts,
                                    size_t g) {
    if (!keys || !ticks || !sums || !counts) return NF_NULL_OR_ABORT;
    switch (variant) {
    case 0:
    case 1: return checked_scatter_i64(keys, ticks, n, sums, counts, g);
    case 2: return sorted_dense_i64(keys, ticks, n, sums, counts, g);
    default: return NF_BAD_VARIANT;
    }
}
/* lib.rs:596-614 + variants.rs:88-101: no V_CHECKED lane (variant 1 -> -3) */
EXPORT int32_t nf_group_variant_i32(const int32_t *keys, const int32_t *ticks, size_t n,
                                    uint32_t variant, int32_t *sums, int64_t *counts,
                                    size_t g) {
    if (!keys || !ticks || !sums || !counts) return NF_NULL_OR_ABORT;
    switch (variant) {
    case 0: return checked_scatter_i32(keys, ticks, n, sums, counts, g);
    case 2: return sorted_dense_i32(keys, ticks, n, sums, counts, g);
    default: return NF_BAD_VARIANT;
    }
}

/* ------------------------------------------------------------------ */
/* sorted runs: groupby/sorted.rs:24-58, lib.rs:399-447                 */
#define SORTED_RUNS_BODY(ACC_T)                                                 \
    int64_t cur = keys[0];                                                      \
    ACC_T acc = vals[0];                                                        \
    int64_t cnt = 1, ng = 0;                                                    \
    for (size_t i = 1; i < n; i++) {                                            \
        int64_t k = keys[i];                                                    \
        if (k < cur) return -1;                                                 \
        if (k != cur) {                                                         \
            ukeys[ng] = cur; sums[ng] = acc; counts[ng] = cnt; ng++;            \
            cur = k; acc = vals[i]; cnt = 1;                                    \
        } else {                                                                \
            acc += vals[i]; cnt += 1;                                           \
        }                                                                       \
    }                                                                           \
    ukeys[ng] = cur; sums[ng] = acc; counts[ng] = cnt;                          \
    return ng + 1;

EXPORT int64_t nf_sorted_run_i64(const int32_t *keys, const int64_t *vals, size_t n,
                                 int64_t *ukeys, int64_t *sums, int64_t *counts) {
    if (n == 0) return 0;
    if (!keys || !vals || !ukeys || !sums || !counts) return -2;
    /* Rust `acc += v` on i64 in release mode wraps (no overflow checks in
     * the release profile); restated with an explicit wrapping add. */
    int64_t cur = keys[0];
    uint64_t acc = (uint64_t)vals[0];
    int64_t cnt = 1, ng = 0;
    for (size_t i = 1; i < n; i++) {
        int64_t k = keys[i];
        if (k < cur) return -1;
        if (k != cur) {
            ukeys[ng] = cur; sums[ng] = (int64_t)acc; counts[ng] = cnt; ng++;
            cur = k; acc = (uint64_t)vals[i]; cnt = 1;
        } else {
            acc += (uint64_t)vals[i]; cnt += 1;
        }
    }
    ukeys[ng] = cur; sums[ng] = (int64_t)acc; counts[ng] = cnt;
    return ng + 1;
}
EXPORT int64_t nf_sorted_run_f64(const int32_t *keys, const double *vals, size_t n,
                                 int64_t *ukeys, double *sums, int64_t *counts) {
    if (n == 0) return 0;
    if (!keys || !vals || !ukeys || !sums || !counts) return -2;
    SORTED_RUNS_BODY(double)
}

/* ------------------------------------------------------------------ */
/* carry compaction: result/carry.rs:23-67 -> core/compact.rs:16-37     */
EXPORT int64_t nf_carry_build_i64(const int64_t *counts_m, const int64_t *sums_m, size_t m,
                                  int64_t *ukeys, int64_t *counts, int64_t *sums) {
    if (m == 0) return 0;
    if (!counts_m || !sums_m || !ukeys || !counts || !sums) return -2;
    int64_t ng = 0;
    for (size_t j = 0; j < m; j++) {
        if (counts_m[j] != 0) {
            ukeys[ng] = (int64_t)j; counts[ng] = counts_m[j]; sums[ng] = sums_m[j]; ng++;
        }
    }
    return ng;
}
EXPORT int64_t nf_carry_build_f64(const int64_t *counts_m, const double *sums_m, size_t m,
                                  int64_t *ukeys, int64_t *counts, double *sums) {
    if (m == 0) return 0;
    if (!counts_m || !sums_m || !ukeys || !counts || !sums) return -2;
    int64_t ng = 0;
    for (size_t j = 0; j < m; j++) {
        if (counts_m[j] != 0) {
            ukeys[ng] = (int64_t)j; counts[ng] = counts_m[j]; sums[ng] = sums_m[j]; ng++;
        }
    }
    return ng;
}

/* ------------------------------------------------------------------ */
/* composite pack: transform/pack.rs:15-32, core/numeric.rs:29-38       */
static int pack_elem(int32_t a, int32_t b, int64_t m, int32_t *out) {
    if (a < 0 || b < 0) return 0;
    int64_t v = (int64_t)a * m + (int64_t)b;
    if (v > INT32_MAX) return 0;
    *out = (int32_t)v;
    return 1;
}
EXPORT int32_t nf_pack_i32_direct(const int32_t *k1, const int32_t *k2, int32_t m2, size_t n,
                                  int32_t *out) {
    if (!k1 || !k2 || !out) return NF_NULL_OR_ABORT;
    for (size_t i = 0; i < n; i++)
        if (!pack_elem(k1[i], k2[i], (int64_t)m2, &out[i])) return NF_BAD_RANGE;
    return NF_OK;
}
/* groupby/pack_fused.rs:26-57 */
EXPORT int32_t nf_pack_sum_count_i32(const int32_t *k1, const int32_t *k2, int32_t m2,
                                     const int32_t *vals, size_t n, int32_t *sums,
                                     int64_t *counts, size_t g) {
    if (!k1 || !k2 || !vals || !sums || !counts) return NF_NULL_OR_ABORT;
    for (size_t j = 0; j < g; j++) { sums[j] = 0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        int32_t pk;
        if (!pack_elem(k1[i], k2[i], (int64_t)m2, &pk)) return NF_BAD_RANGE;
        if ((uint64_t)(int64_t)pk >= g) return NF_BAD_RANGE;
        if (!add_i32_checked(&sums[pk], vals[i])) return NF_OVERFLOW;
        counts[pk] += 1;
    }
    return NF_OK;
}
/* groupby/pack_fused.rs:61-90 */
EXPORT int32_t nf_pack_sum_count_f64(const int32_t *k1, const int32_t *k2, int32_t m2,
                                     const double *vals, size_t n, double *sums,
                                     int64_t *counts, size_t g) {
    if (!k1 || !k2 || !vals || !sums || !counts) return NF_NULL_OR_ABORT;
    for (size_t j = 0; j < g; j++) { sums[j] = 0.0; counts[j] = 0; }
    for (size_t i = 0; i < n; i++) {
        int32_t pk;
        if (!pack_elem(k1[i], k2[i], (int64_t)m2, &pk)) return NF_BAD_RANGE;
        if ((uint64_t)(int64_t)pk >= g) return NF_BAD_RANGE;
        sums[pk] += vals[i];
        counts[pk] += 1;
    }
    return NF_OK;
}

/* ------------------------------------------------------------------ */
/* prefix+int string encode: transform/pattern.rs:50-131, lib.rs:340-373 */
EXPORT int32_t nf_pattern_encode(const uint8_t *data, size_t total, const int32_t *offs,
                                 size_t n, const uint8_t *prefix, size_t prefix_len,
                                 int32_t *codes, uint8_t *valid, int32_t *width_out,
                                 int32_t *err_row_out) {
    if (!data || !offs || !prefix || !codes || !valid || !width_out || !err_row_out)
        return NF_NULL_OR_ABORT;
    *width_out = -1;
    *err_row_out = -1;
    int32_t prev = offs[0];
    if (prev < 0 || (size_t)prev > total) return NF_MALFORMED;
    for (size_t i = 0; i < n; i++) {
        codes[i] = 0;
        valid[i] = 0;
        int32_t end = offs[i + 1];
        if (end < prev || end < 0 || (size_t)end > total) return NF_MALFORMED;
        const uint8_t *row = data + prev;
        size_t len = (size_t)(end - prev);
        prev = end;
        if (len < prefix_len || memcmp(row, prefix, prefix_len) != 0) continue;
        const uint8_t *body = row + prefix_len;
        size_t blen = len - prefix_len;
        int neg = blen > 0 && body[0] == '-';
        if (neg) { body++; blen--; }
        if (blen == 0) continue;
        int64_t mag = 0;
        int32_t w = 0;
        int ok = 1;
        for (size_t j = 0; j < blen; j++) {
            uint8_t ch = body[j];
            if (ch < '0' || ch > '9') { ok = 0; break; }
            /* saturating Horner (pattern.rs:28-40 for 19+ digit bodies;
             * exact for <=18 digits, pattern.rs:94-109) */
            int64_t d = (int64_t)(ch - '0');
            if (mag > (INT64_MAX - d) / 10) mag = INT64_MAX;
            else mag = mag * 10 + d;
            w++;
        }
        if (!ok) continue;
        int64_t lim = neg ? ((int64_t)1 << 31) : (int64_t)INT32_MAX;
        if (mag > lim) { *err_row_out = (int32_t)i; return NF_BAD_RANGE; }
        codes[i] = neg ? (int32_t)(-mag) : (int32_t)mag;
        valid[i] = 1;
        if (w > *width_out) *width_out = w;
    }
    if (offs[n] < 0 || (size_t)offs[n] > total) return NF_MALFORMED;
    return NF_OK;
}

/* ---- select / mask: numfast-native/src/series/select.rs:13-67 behind
 * lib.rs:616-1062.  Count and scatter return m, -2 on a null pointer; the
 * mask ops return 0 / -1. */
EXPORT int64_t nf_select_count(const uint8_t *mask, size_t n) {
    if (n == 0) return 0;
    if (!mask) return NF_BAD_RANGE;
    int64_t m = 0;
    for (size_t i = 0; i < n; i++) m += mask[i] != 0;
    return m;
}

#define SELECT_SCATTER(name, T)                                                   \
    EXPORT int64_t name(const T *src, const uint8_t *mask, size_t n, T *out) {    \
        if (n == 0) return 0;                                                     \
        if (!src || !mask || !out) return NF_BAD_RANGE;                           \
        size_t m = 0;                                                             \
        for (size_t i = 0; i < n; i++)                                            \
            if (mask[i] != 0) out[m++] = src[i];                                  \
        return (int64_t)m;                                                        \
    }
SELECT_SCATTER(nf_select_scatter_i32, int32_t)
SELECT_SCATTER(nf_select_scatter_i64, int64_t)
SELECT_SCATTER(nf_select_scatter_f32, float)
SELECT_SCATTER(nf_select_scatter_f64, double)
SELECT_SCATTER(nf_select_scatter_u8, uint8_t)

EXPORT int32_t nf_mask_and(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    if (n == 0) return NF_OK;
    if (!a || !b || !out) return NF_NULL_OR_ABORT;
    for (size_t i = 0; i < n; i++) out[i] = (uint8_t)((a[i] != 0) & (b[i] != 0));
    return NF_OK;
}
EXPORT int32_t nf_mask_or(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    if (n == 0) return NF_OK;
    if (!a || !b || !out) return NF_NULL_OR_ABORT;
    for (size_t i = 0; i < n; i++) out[i] = (uint8_t)((a[i] != 0) | (b[i] != 0));
    return NF_OK;
}
EXPORT int32_t nf_mask_not(const uint8_t *a, size_t n, uint8_t *out) {
    if (n == 0) return NF_OK;
    if (!a || !out) return NF_NULL_OR_ABORT;
    for (size_t i = 0; i < n; i++) out[i] = (uint8_t)(a[i] == 0);
    return NF_OK;
}

/* ------------------------------------------------------------------ ghash
 * Partitioned pair-hash for grouped COUNT DISTINCT
 * (numfast-native/src/groupby/hash_grouped.rs; FFI lib.rs:2266-2470).
 * int64 lanes.  The reference shards rows (fp, count, scatter) or
 * partitions (pins, occ_count, occ_fill) over `nthreads` threads; the
 * restatement runs the shards one after another, with the same chunk
 * boundaries, which is what makes count / scatter / occ_fill outputs
 * depend on nthreads exactly as the reference's do.
 * Refusals where the reference would index out of bounds (panic, abort)
 * or probe forever: -2 for a partition id or reservation outside its
 * buffer, -3 for a table too small for its distinct pairs. */
#define GH_PART_MAX 256
#define GH_THREADS_MAX 64

/* hash_grouped.rs:55-68 */
static uint64_t gh_pair_fp(int64_t k, int64_t v) {
  const uint64_t uk = (uint64_t)k, uv = (uint64_t)v;
  return uk * 0x9E3779B97F4A7C15ull + (uv ^ (uv >> 33)) * 0x00BF58476D1FF973ull;
}

/* hash_grouped.rs:70-74: clamp to [1, 64], capped by the work items */
static size_t gh_fanout(size_t nthreads, size_t work) {
  size_t t = nthreads < 1 ? 1 : (nthreads > GH_THREADS_MAX ? GH_THREADS_MAX : nthreads);
  const size_t w = work < 1 ? 1 : work;
  return t < w ? t : w;
}

static int gh_pow2(size_t p) { return p != 0 && p <= GH_PART_MAX && (p & (p - 1)) == 0; }

/* lib.rs:2266-2287, hash_grouped.rs:78-130 */
EXPORT int32_t nf_ghash_fp_i64(const int64_t *k, const int64_t *v, size_t n, int64_t *fp,
                               int64_t pmask, size_t nthreads) {
  (void)nthreads;
  if (n == 0) return NF_OK;
  if (!k || !v || !fp) return NF_NULL_OR_ABORT;
  for (size_t i = 0; i < n; ++i) fp[i] = (int64_t)(gh_pair_fp(k[i], v[i]) & (uint64_t)pmask);
  return NF_OK;
}

/* lib.rs:2299-2316, hash_grouped.rs:136-190: chunk tt of
 * nt = min(fanout(nthreads, n), t) equal row chunks (ceil) adds its
 * partition histogram into tc[tt * p + q] */
EXPORT int32_t nf_ghash_count_i64(const int64_t *fp, size_t n, int64_t *tc, size_t t, size_t p,
                                  size_t nthreads) {
  if (n == 0) return NF_OK;
  if (!fp || !tc) return NF_NULL_OR_ABORT;
  if (!gh_pow2(p) || t == 0 || t > GH_THREADS_MAX) return NF_MALFORMED;
  size_t nt = gh_fanout(nthreads, n);
  if (nt > t) nt = t;
  const size_t cs = (n + nt - 1) / nt;
  for (size_t i = 0; i < n; ++i)
    if (fp[i] < 0 || (uint64_t)fp[i] >= p) return NF_BAD_RANGE;
  for (size_t tt = 0; tt < nt; ++tt) {
    const size_t s0 = tt * cs < n ? tt * cs : n, s1 = (tt + 1) * cs < n ? (tt + 1) * cs : n;
    for (size_t i = s0; i < s1; ++i) tc[tt * p + (size_t)fp[i]] += 1;
  }
  return NF_OK;
}

/* lib.rs:2334-2360, hash_grouped.rs:196-290: chunk tt writes its rows in
 * row order from the cursor row tpos[tt * p ..] */
EXPORT int32_t nf_ghash_scatter_i64(const int64_t *k, const int64_t *v, const int64_t *fp,
                                    size_t n, const int64_t *tpos, size_t t, size_t p,
                                    int64_t *pk, int64_t *pv, size_t nthreads) {
  if (n == 0) return NF_OK;
  if (!k || !v || !fp || !tpos || !pk || !pv) return NF_NULL_OR_ABORT;
  if (!gh_pow2(p) || t == 0 || t > GH_THREADS_MAX) return NF_MALFORMED;
  size_t nt = gh_fanout(nthreads, n);
  if (nt > t) nt = t;
  const size_t cs = (n + nt - 1) / nt;
  int64_t cur[GH_PART_MAX];
  for (size_t tt = 0; tt < nt; ++tt) {
    const size_t s0 = tt * cs < n ? tt * cs : n, s1 = (tt + 1) * cs < n ? (tt + 1) * cs : n;
    for (size_t q = 0; q < p; ++q) cur[q] = tpos[tt * p + q];
    for (size_t i = s0; i < s1; ++i) {
      if (fp[i] < 0 || (uint64_t)fp[i] >= p) return NF_BAD_RANGE;
      const int64_t lane = cur[fp[i]]++;
      if (lane < 0 || (uint64_t)lane >= n) return NF_BAD_RANGE;
      pk[lane] = k[i];
      pv[lane] = v[i];
    }
  }
  return NF_OK;
}

/* lib.rs:2374-2409, hash_grouped.rs:296-440: each partition inserts its
 * rows, in order, into its linear-probing table (full pair compare) */
EXPORT int32_t nf_ghash_pins_i64(const int64_t *pk, const int64_t *pv, const int64_t *starts,
                                 const int64_t *toff, const int64_t *tm, int64_t *tk,
                                 int64_t *tv, uint8_t *used, size_t p, size_t nthreads) {
  (void)nthreads;
  if (!pk || !pv || !starts || !toff || !tm || !tk || !tv || !used) return NF_NULL_OR_ABORT;
  if (!gh_pow2(p)) return NF_MALFORMED;
  for (size_t q = 0; q < p; ++q) {
    const uint64_t slots = (uint64_t)tm[q] + 1;
    if (tm[q] < 0 || (slots & (slots - 1)) != 0 || starts[q] > starts[q + 1] ||
        toff[q] + (int64_t)slots > toff[q + 1] || starts[q] < 0 || toff[q] < 0)
      return NF_MALFORMED;
  }
  for (size_t q = 0; q < p; ++q) {
    const uint64_t mask = (uint64_t)tm[q];
    const int64_t tbase = toff[q];
    for (int64_t j = starts[q]; j < starts[q + 1]; ++j) {
      const int64_t rk = pk[j], vv = pv[j];
      uint64_t h = gh_pair_fp(rk, vv) & mask;
      for (uint64_t probes = 0;; ++probes) {
        /* every slot seen, none free or equal: the reference probes
         * forever here */
        if (probes > mask) return NF_MALFORMED;
        const int64_t s = tbase + (int64_t)h;
        if (used[s] == 0) {
          used[s] = 1;
          tk[s] = rk;
          tv[s] = vv;
          break;
        } else if (tk[s] == rk && tv[s] == vv) {
          break;
        }
        h = (h + 1) & mask;
      }
    }
  }
  return NF_OK;
}

/* lib.rs:2419-2438, hash_grouped.rs:460-520 */
EXPORT int32_t nf_ghash_occ_count(const uint8_t *used, const int64_t *toff, size_t p,
                                  int64_t *cnt, size_t nthreads) {
  (void)nthreads;
  if (!used || !toff || !cnt) return NF_NULL_OR_ABORT;
  if (!gh_pow2(p)) return NF_MALFORMED;
  for (size_t q = 0; q < p; ++q) {
    int64_t c = 0;
    for (int64_t s = toff[q]; s < toff[q + 1]; ++s) c += used[s] != 0;
    cnt[q] = c;
  }
  return NF_OK;
}

/* lib.rs:2448-2470, hash_grouped.rs:522-634: ok[0, nocc) is carved into
 * consecutive pieces, one per shard tt (partitions [tt*cs, ..),
 * cs = ceil(p / fanout(nthreads, p))) of ostart[hi] - ostart[lo] entries;
 * each shard writes its partitions' occupied keys, in slot order, into
 * its piece.  A carve or write past a piece panics in the reference. */
EXPORT int32_t nf_ghash_occ_fill(const int64_t *tk, const uint8_t *used, const int64_t *toff,
                                 const int64_t *ostart, int64_t *ok, size_t p, size_t nthreads) {
  if (!tk || !used || !toff || !ostart || !ok) return NF_NULL_OR_ABORT;
  if (!gh_pow2(p)) return NF_MALFORMED;
  const size_t t = gh_fanout(nthreads, p), cs = (p + t - 1) / t;
  if (ostart[0] < 0) return NF_BAD_RANGE;
  for (size_t plo = 0; plo < p; plo += cs) {
    const size_t phi = plo + cs < p ? plo + cs : p;
    if (ostart[phi] < ostart[plo]) return NF_BAD_RANGE;
  }
  for (size_t plo = 0; plo < p; plo += cs) {
    const size_t phi = plo + cs < p ? plo + cs : p;
    const int64_t off = ostart[plo] - ostart[0], len = ostart[phi] - ostart[plo];
    int64_t w = 0;
    for (size_t q = plo; q < phi; ++q)
      for (int64_t s = toff[q]; s < toff[q + 1]; ++s)
        if (used[s] != 0) {
          if (w >= len) return NF_BAD_RANGE;
          ok[off + w++] = tk[s];
        }
  }
  return NF_OK;
}
