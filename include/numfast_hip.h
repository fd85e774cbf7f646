/*
 * numfast_hip.h -- C ABI of libnumfast_hip.so (MI355X / gfx950).
 *
 * The nf_* symbols are drop-in replacements for the reference crate
 * numfast-native: same names, argument order and meaning, same return
 * codes (numfast-native/src/core/errors.rs:13-38), loaded by NumFast's
 * ctypes binding through NUMFAST_NATIVE_DLL
 * (src/Drivers/CPU/_lib/native_cpu.py:93-143).  Buffers may be host memory
 * (staged through device memory) or device memory (used in place); every
 * call returns when its outputs are written.  Additional return codes:
 *   -6  device error (no GPU, out of memory, launch failure)
 *   -7  caller capacity too small (nfh_group_agg_i32 only)
 * Where the reference aborts the process (an out-of-range dense key
 * reaching an unchecked lane), these symbols return -2 instead.
 *
 * The nfh_* symbols are additive: they have no reference counterpart.
 */
#ifndef NUMFAST_HIP_H
#define NUMFAST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- eager group (native_cpu.py:125-140: all eight or none) ---------- */

/* numfast-native/src/lib.rs:85-103 -- dense f64 sum + count by code. */
int32_t nf_group_sum_count(const int32_t *keys, const double *values, size_t n,
                           double *sums, int64_t *counts, size_t g);
/* lib.rs:117-135 -- SoA multi-column f64 sums + shared counts. */
int32_t nf_group_multi_sum_count(const int32_t *keys, const double *values, size_t n,
                                 size_t ncols, double *sums, int64_t *counts, size_t g);
/* lib.rs:252-272 -- out[i] = k1[i]*m2 + k2[i], -2 on negative / overflow. */
int32_t nf_pack_i32_direct(const int32_t *k1, const int32_t *k2, int32_t m2, size_t n,
                           int32_t *out);
/* lib.rs:340-373 -- prefix+int string encode. */
int32_t nf_pattern_encode(const uint8_t *data, size_t total, const int32_t *offs, size_t n,
                          const uint8_t *prefix, size_t prefix_len, int32_t *codes,
                          uint8_t *valid, int32_t *width_out, int32_t *err_row_out);
/* lib.rs:399-447 -- run aggregation over nondecreasing keys; ng or -1. */
int64_t nf_sorted_run_i64(const int32_t *keys, const int64_t *vals, size_t n, int64_t *ukeys,
                          int64_t *sums, int64_t *counts);
int64_t nf_sorted_run_f64(const int32_t *keys, const double *vals, size_t n, int64_t *ukeys,
                          double *sums, int64_t *counts);
/* lib.rs:462-518 -- dense state -> compact (ukeys, counts, sums); ng. */
int64_t nf_carry_build_i64(const int64_t *counts_m, const int64_t *sums_m, size_t m,
                           int64_t *ukeys, int64_t *counts, int64_t *sums);
int64_t nf_carry_build_f64(const int64_t *counts_m, const double *sums_m, size_t m,
                           int64_t *ukeys, int64_t *counts, double *sums);

/* ---- variant group (native_cpu.py:324-352) --------------------------- */
/* lib.rs:539-614; variant 0 dense, 1 checked (not i32), 2 sorted-dense. */
int32_t nf_group_variant_f64(const int32_t *keys, const double *values, size_t n,
                             uint32_t variant, double *sums, int64_t *counts, size_t g);
int32_t nf_group_variant_i64(const int32_t *keys, const int64_t *ticks, size_t n,
                             uint32_t variant, int64_t *sums, int64_t *counts, size_t g);
int32_t nf_group_variant_i32(const int32_t *keys, const int32_t *ticks, size_t n,
                             uint32_t variant, int32_t *sums, int64_t *counts, size_t g);

/* ---- mixed group (native_cpu.py:503-532) ----------------------------- */
/* lib.rs:150-178 -- sum-only over n_i32 int32 + n_f64 float64 SoA lanes. */
int32_t nf_group_mixed_sum_only(const int32_t *keys, const int32_t *i32_vals,
                                const double *f64_vals, size_t n, size_t n_i32, size_t n_f64,
                                int32_t *i32_sums, double *f64_sums, size_t g);
/* lib.rs:184-197 -- keys-only counts. */
int32_t nf_group_count_only(const int32_t *keys, size_t n, int64_t *counts, size_t g);

/* ---- owner group (native_cpu.py:548-574) ----------------------------- */
/* lib.rs:213-250 -- writes only lanes [lo, hi) of the shared state. */
int32_t nf_group_owner_2i32_1f64(const int32_t *keys, const int32_t *a, const int32_t *b,
                                 const double *c, size_t n, size_t lo, size_t hi, int32_t *s1,
                                 int32_t *s2, double *s3, int64_t *counts, size_t g);

/* ---- fused group (native_cpu.py:621-651) ----------------------------- */
/* lib.rs:276-331 -- pack k1*m2+k2 fused into the dense sum + count. */
int32_t nf_pack_sum_count_i32(const int32_t *k1, const int32_t *k2, int32_t m2,
                              const int32_t *vals, size_t n, int32_t *sums, int64_t *counts,
                              size_t g);
int32_t nf_pack_sum_count_f64(const int32_t *k1, const int32_t *k2, int32_t m2,
                              const double *vals, size_t n, double *sums, int64_t *counts,
                              size_t g);

/* ---- unique group (native_cpu.py:1399-1431) -------------------------- */
/* lib.rs:1420-1500 -- np.unique(keys, return_inverse=True); returns ng. */
int64_t nf_unique_inverse_i32(const int32_t *keys, size_t n, int32_t *uniq_out,
                              int32_t *inv_out, int32_t *perm, uint32_t *tmp0, uint32_t *tmp1,
                              int32_t *tmp_p);
int64_t nf_unique_inverse_i64(const int64_t *keys, size_t n, int64_t *uniq_out,
                              int32_t *inv_out, int32_t *perm, uint64_t *tmp0, uint64_t *tmp1,
                              int32_t *tmp_p);

/* ---- sort group (native_cpu.py:1546-1576) ---------------------------- */
/* lib.rs:1326-1394 -- stable argsort; descending != 0 for descending. */
int32_t nf_sort_perm_i32(const int32_t *keys, size_t n, uint8_t descending, int32_t *perm_out,
                         uint32_t *tmp0, uint32_t *tmp1, int32_t *tmp_p);
int32_t nf_sort_perm_i64(const int64_t *keys, size_t n, uint8_t descending, int32_t *perm_out,
                         uint64_t *tmp0, uint64_t *tmp1, int32_t *tmp_p);

/* ---- select / mask (IR filter): lib.rs:616-1062, series/select.rs ---- */

/* lib.rs:630-641 -- rows with mask != 0; -2 on a null pointer. */
int64_t nf_select_count(const uint8_t *mask, size_t n);
/* lib.rs:643-700 -- order-preserving compaction of src where mask != 0
 * into out[0..m); returns m, -2 on a null pointer. */
int64_t nf_select_scatter_i32(const int32_t *src, const uint8_t *mask, size_t n, int32_t *out);
int64_t nf_select_scatter_i64(const int64_t *src, const uint8_t *mask, size_t n, int64_t *out);
int64_t nf_select_scatter_f32(const float *src, const uint8_t *mask, size_t n, float *out);
int64_t nf_select_scatter_f64(const double *src, const uint8_t *mask, size_t n, double *out);
int64_t nf_select_scatter_u8(const uint8_t *src, const uint8_t *mask, size_t n, uint8_t *out);
/* lib.rs:1014-1062 -- elementwise mask algebra, output 0/1; 0 or -1. */
int32_t nf_mask_and(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out);
int32_t nf_mask_or(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out);
int32_t nf_mask_not(const uint8_t *a, size_t n, uint8_t *out);

/* ---- partitioned pair-hash (grouped COUNT DISTINCT):
 *      lib.rs:2266-2470, groupby/hash_grouped.rs; bound by
 *      src/Drivers/GroupedHashMT/_lib/grouped_native_hash.py:95-128 ----
 * int64 lanes; p a power of two in [1, 256], t in [1, 64]; nthreads sets
 * the reference's sharding, which the outputs of count / scatter /
 * occ_fill follow exactly.  0 ok, -1 null pointer, -3 bad geometry; -2
 * where the reference would index out of bounds and -3 where its probe
 * would not end (a table with no room for a new pair). */
/* lib.rs:2266-2287 -- fp[i] = (k*C1 + (v ^ v>>33)*C2) & pmask (wrapping). */
int32_t nf_ghash_fp_i64(const int64_t *k, const int64_t *v, size_t n, int64_t *fp,
                        int64_t pmask, size_t nthreads);
/* lib.rs:2299-2316 -- tc[tt*p + q] += rows of row chunk tt with fp == q. */
int32_t nf_ghash_count_i64(const int64_t *fp, size_t n, int64_t *tc, size_t t, size_t p,
                           size_t nthreads);
/* lib.rs:2334-2360 -- chunk tt writes its rows in row order from the
 * cursors tpos[tt*p + q] into pk / pv. */
int32_t nf_ghash_scatter_i64(const int64_t *k, const int64_t *v, const int64_t *fp, size_t n,
                             const int64_t *tpos, size_t t, size_t p, int64_t *pk, int64_t *pv,
                             size_t nthreads);
/* lib.rs:2374-2409 -- partition q's rows [starts[q], starts[q+1]) inserted
 * in row order into its linear-probing table of tm[q]+1 slots at toff[q]
 * (used zeroed by the caller); the same layout as sequential insertion. */
int32_t nf_ghash_pins_i64(const int64_t *pk, const int64_t *pv, const int64_t *starts,
                          const int64_t *toff, const int64_t *tm, int64_t *tk, int64_t *tv,
                          uint8_t *used, size_t p, size_t nthreads);
/* lib.rs:2419-2438 -- cnt[q] = used slots of partition q. */
int32_t nf_ghash_occ_count(const uint8_t *used, const int64_t *toff, size_t p, int64_t *cnt,
                           size_t nthreads);
/* lib.rs:2448-2470 -- occupied keys in slot order, per shard from ok's
 * consecutive pieces of ostart[hi] - ostart[lo]. */
int32_t nf_ghash_occ_fill(const int64_t *tk, const uint8_t *used, const int64_t *toff,
                          const int64_t *ostart, int64_t *ok, size_t p, size_t nthreads);

/* ---- additive surface (no reference counterpart) --------------------- */

/* Order this thread's calls on a caller stream (hipStream_t); NULL
 * restores the library's own per-thread stream. */
int32_t nfh_set_stream(void *stream);
/* 1 when a GPU is usable, else 0. */
int32_t nfh_device_ok(void);
int32_t nfh_abi_version(void);

/* Per-kernel HIP-event timing on the launch stream (for benchmarks):
 * enable, reset, and read the total ms / launch count of a kernel name
 * (returns -1 for an unknown name). */
int32_t nfh_profile_enable(int32_t on);
int32_t nfh_profile_reset(void);
int32_t nfh_profile_get(const char *name, double *ms, uint64_t *launches);
/* Algorithmic bytes recorded for a kernel name (sum over its launches). */
int32_t nfh_profile_bytes(const char *name, double *bytes);
/* Recorded kernel names, newline-separated, into buf; returns the length. */
int64_t nfh_profile_names(char *buf, size_t cap);

/* Group-by over raw int32 keys with NULL masks (valid / key_valid may be
 * NULL): groups of rows with both masks set, ukeys ascending; sums int64,
 * min/max int32, mean = double(sum)/double(count).  Replaces the
 * executor's compaction + np.minimum/maximum.at + bincount path
 * (src/Drivers/CPU/_lib/cpu.py:3507-3563, carry.py:55-62).  Outputs hold
 * cap lanes (NULL outputs are skipped); returns ng, or -7 if ng > cap. */
int64_t nfh_group_agg_i32(const int32_t *keys, const int32_t *vals, const uint8_t *valid,
                          const uint8_t *key_valid, size_t n, int64_t *ukeys, int64_t *counts,
                          int64_t *sums, int32_t *mins, int32_t *maxs, double *means,
                          size_t cap);

/* The same group-by with a float64 value column: counts int64; sums,
 * mins, maxs, means float64.  Sums are the np.bincount row-order sums of
 * the reference within 1e-12 relative (exactly rounded, or bit-equal where
 * redone in row order); min / max are np.minimum.at / np.maximum.at over a
 * +-inf fill (cpu.py:2570-2601): the group's first NaN when it has one, a
 * zero extreme with the sign of the group's last zero.  Returns ng or -7. */
int64_t nfh_group_agg_f64(const int32_t *keys, const double *vals, const uint8_t *valid,
                          const uint8_t *key_valid, size_t n, int64_t *ukeys, int64_t *counts,
                          double *sums, double *mins, double *maxs, double *means, size_t cap);
/* int64 value column: sums int64 (wrapping; exact where the executor's 2^53
 * gate admits the column, cpu.py:424-446), mins / maxs int64 (iinfo fill),
 * means float64.  Returns ng or -7. */
int64_t nfh_group_agg_i64(const int32_t *keys, const int64_t *vals, const uint8_t *valid,
                          const uint8_t *key_valid, size_t n, int64_t *ukeys, int64_t *counts,
                          int64_t *sums, int64_t *mins, int64_t *maxs, double *means,
                          size_t cap);

/* Positional take (the IR gather, src/Drivers/CPU/_lib/cpu.py:1110-1138,
 * :3170-3176): out[i] = values[idx[i]] for k indices of idx_bytes (4 or 8),
 * elements of elem_size bytes (1, 4 or 8); valid_in / valid_out (both or
 * neither) take the validity bytes alike.  0, -1 null pointer, -2 an index
 * outside [0, n) (the reference raises; outputs unspecified), -3 bad sizes,
 * -6 device error. */
int32_t nfh_gather(const void *values, size_t n, size_t elem_size, const void *idx,
                   size_t idx_bytes, size_t k, void *out, const uint8_t *valid_in,
                   uint8_t *valid_out);

/* np.unique over valid rows; inv = -1 on NULL rows
 * (src/Drivers/CPU/_lib/cpu.py:4032-4036).  Returns ng. */
int64_t nfh_unique_valid_i32(const int32_t *keys, const uint8_t *valid, size_t n,
                             int32_t *uniq_out, int32_t *inv_out);

/* No counterpart in the reference: nf_sort_perm_i32 (lib.rs:1326-1357) for
 * callers that cannot allocate device memory (HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Device buffers only; the sort's
 * own scratch is a fixed 2.1 MiB device array per GPU, so n <= 2^24.  Runs
 * on the stream set by nfh_set_stream (else the null stream) and returns
 * when it is done.  Calls that overlap on one GPU share that array and
 * race.  0, -1 null pointer, -3 host buffer or n > 2^24, -6 device error. */
int32_t nfh_sort_perm_i32_noalloc(const int32_t *keys, size_t n, uint8_t descending,
                                  int32_t *perm_out, uint32_t *tmp0, uint32_t *tmp1,
                                  int32_t *tmp_p);

/* Learned-CDF model: a 1 -> 32 -> 32 -> 1 fp32 ReLU network over the
 * normalised key word (w - wmin) / span, w = (uint32)key ^ 0x80000000.
 * p = w1[32] b1[32] w2[32][32] b2[32] w3[32] b3. */
typedef struct nfh_cdf_model {
  float p[1153];
  uint32_t wmin;
  uint32_t span;
  uint32_t hints; /* bit 0: the training sample was nearly sorted (set by
                     nfh_cdf_train_i32; changes the sort's speed, never its
                     result) */
} nfh_cdf_model;

/* Train a model on a column (2^16-key sample sorted by the radix argsort,
 * plus the exact CDF at top-byte boundaries from the radix histogram
 * kernel): the network starts from a greedy 32-piece fit of the sample CDF
 * (flags bit 0: from the linear CDF) and takes `steps` full-batch gradient
 * steps on the device (normalised, with backtracking).  *loss_out (may be
 * NULL): final mean squared CDF error.  0, -1, -3 or -6. */
int32_t nfh_cdf_train_i32(const int32_t *keys, size_t n, uint32_t steps, uint32_t flags,
                          nfh_cdf_model *model_out, float *loss_out);
/* Stable ascending argsort through the model's bucket code; the result
 * equals nf_sort_perm_i32(keys, n, 0, ...) for any model.  stats (may be
 * NULL, 4 words): non-empty buckets, largest bucket, buckets over the LDS
 * capacity (each sorted by one workgroup through global memory, or copied
 * when its keys are all equal), and of those the ones over 2^17 keys with
 * unequal keys, re-sorted by the plain-sort fix-up. */
int32_t nfh_sort_perm_learned_i32(const int32_t *keys, size_t n, const nfh_cdf_model *model,
                                  int32_t *perm_out, uint64_t *stats);
/* Mean-squared-error gradient of the network at params[1153] over (x, y),
 * computed by the training kernels (grad_out[1153]; loss_out, pred_out[n]
 * may be NULL). */
int32_t nfh_mlp_grad(const float *params, const float *x, const float *y, size_t n,
                     float *grad_out, float *loss_out, float *pred_out);

/* filter(v > 0) -> derive(w = v * mul, int32 wrap) -> stable sort by w ->
 * limit: the first min(limit, m) rows as (w, source row index); returns
 * that count.  model NULL: LSD radix sort; else the learned-CDF sort
 * (identical output).  cpu.py:3034-3035, :3074, :1174-1225;
 * src/Semantic/TableExpr/_lib/chain.py:978-1008. */
int64_t nfh_filter_derive_sort_limit_i32(const int32_t *v, size_t n, int32_t mul, size_t limit,
                                         const nfh_cdf_model *model, int32_t *w_out,
                                         int32_t *rows_out);

/* Group-by strategy for this thread's following nfh_group_agg_i32 calls:
 * 0 the built-in rule, 1 dense (LDS hot keys + global spill), 2 partitioned
 * (bucket passes, 2048-key buckets), 3 sort-based, 4 hash-scatter (LDS hash
 * tables + one global hash table; sort-based when that table overflows),
 * 5 partitioned with 4096-key buckets (half the partition count); an
 * impossible choice falls back to the rule.  Returns the previous setting.
 * Results are identical for all.  With 4 set, nfh_group_agg_i32 refuses
 * inputs of more than 2^22 rows with -3. */
int32_t nfh_set_groupby_strategy(int32_t strategy);
/* The strategy the last group-by on this thread ran. */
int32_t nfh_groupby_last_strategy(void);

/* Learned strategy selector: 64 features of a key column (56-bin sampled
 * histogram + 8 summaries), and a 64 -> 64 -> 5 fp32 classifier
 * (params[4485] = W1[64][64] b1[64] W2[5][64] b2[5]) trained on the device
 * (full-batch Adam, softmax cross-entropy); class c = strategy c + 1. */
int32_t nfh_groupby_features_i32(const int32_t *keys, size_t n, float *out64);
int32_t nfh_selector_grad(const float *params, const float *X, const int32_t *y, size_t n,
                          float *grad_out, float *loss_out);
int32_t nfh_selector_train(const float *X, const int32_t *y, size_t n, uint32_t steps, float lr,
                           float *params_inout, float *loss_out);
int32_t nfh_selector_predict(const float *params, const float *X, size_t n, int32_t *cls_out);
/* The same network as a cost model: z_c predicts strategy c's centred log
 * time (T[n][5]), trained on the mean of 1/2 sum_c (z_c - T_c)^2; the pick
 * is the smallest predicted cost. */
int32_t nfh_selector_grad_cost(const float *params, const float *X, const float *T, size_t n,
                               float *grad_out, float *loss_out);
int32_t nfh_selector_train_cost(const float *X, const float *T, size_t n, uint32_t steps,
                                float lr, float *params_inout, float *loss_out);
int32_t nfh_selector_predict_cost(const float *params, const float *X, size_t n,
                                  int32_t *cls_out);

/* Batched group-by (no reference counterpart): nq independent dense
 * sum / count queries over concatenated device rows; query q owns rows
 * [row_off[q], row_off[q+1]) (< 2^24 rows) and keys [0, card[q])
 * (card <= 2^24).  Output (device, cap lanes): present groups query by
 * query, keys ascending, as ukeys / counts / sums (int64), and
 * group_off[nq + 1].  The plan autotunes rows per work item and the
 * persistent grid of the small-query pass (card <= 4096) on these buffers
 * and captures the batched run in a hipGraph; it keeps the pointers (the
 * buffers must outlive it).  NULL on failure with *rc_out = -1, -3 or -6. */
void *nfh_group_batch_plan(const int32_t *keys, const int32_t *vals, const int64_t *row_off,
                           const int32_t *card, size_t nq, int64_t *ukeys, int64_t *counts,
                           int64_t *sums, int64_t *group_off, size_t cap, int32_t *rc_out);
/* mode 0: replay the captured graph (one persistent launch of the batch);
 * mode 1: the plain form, kernels launched query by query.  Returns the
 * group count, -2 (a key outside its query's range) or -7 (cap too small). */
int64_t nfh_group_batch_run(void *plan, int32_t mode);
/* small-query pass: rows per work item, workgroups, work items, key slots. */
int32_t nfh_group_batch_info(void *plan, uint64_t *out4);
/* wide queries (card > 4096), cut into 16384-key units: partition tiles,
 * units, (tile, unit) record entries. */
int32_t nfh_group_batch_info2(void *plan, uint64_t *out3);
void nfh_group_batch_free(void *plan);

#ifdef __cplusplus
}
#endif
#endif /* NUMFAST_HIP_H */
