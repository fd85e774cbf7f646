// boundary_asan.cpp -- the C ABI's host side under AddressSanitizer.
//
// Linked against a build of the library whose host code is instrumented
// (hipcc -Xarch_host -fsanitize=address; device code is not), this driver
// calls every entry-point family with malloc'd host buffers -- so each call
// classifies its pointers, stages them through device scratch and copies
// the outputs back -- at odd sizes, with unaligned pointers (the scalar
// paths), with the argument-check refusals, and from several threads at
// once on disjoint slices (the reference's thread-pool pattern,
// src/Drivers/CPU/_lib/cpu.py:558-581).  Each result is checked against a
// plain loop here.  Exit 0 when everything matched; ASan aborts the process
// on a heap error in the library's host code.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#include "../../include/numfast_hip.h"

static int g_fail = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++g_fail;                                                            \
    }                                                                      \
  } while (0)

template <class T>
static T* hmalloc(size_t n) {  // plain heap memory, as NumPy gives
  return static_cast<T*>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
}

static void sort_and_unique(size_t n, uint32_t seed, int offset) {
  std::mt19937 rng(seed);
  int32_t* kb = hmalloc<int32_t>(n + 1);
  int32_t* k = kb + offset;  // offset 1: unaligned, the scalar paths
  for (size_t i = 0; i < n; ++i) k[i] = int32_t(rng() % 1000) - 500;
  int32_t* perm = hmalloc<int32_t>(n);
  uint32_t* t0 = hmalloc<uint32_t>(n);
  uint32_t* t1 = hmalloc<uint32_t>(n);
  int32_t* tp = hmalloc<int32_t>(n);
  for (int desc = 0; desc < 2; ++desc) {
    CHECK(nf_sort_perm_i32(k, n, uint8_t(desc), perm, t0, t1, tp) == 0);
    std::vector<int32_t> ref(n);
    std::iota(ref.begin(), ref.end(), 0);
    std::stable_sort(ref.begin(), ref.end(), [&](int32_t a, int32_t b) {
      return desc ? k[a] > k[b] : k[a] < k[b];
    });
    CHECK(std::equal(ref.begin(), ref.end(), perm));
  }
  int32_t* uniq = hmalloc<int32_t>(n);
  int32_t* inv = hmalloc<int32_t>(n);
  const int64_t ng = nf_unique_inverse_i32(k, n, uniq, inv, perm, t0, t1, tp);
  CHECK(ng > 0 && ng <= 1000);
  for (size_t i = 0; i < n && ng > 0; ++i) CHECK(inv[i] >= 0 && uniq[inv[i]] == k[i]);
  // NULL-aware unique and group-by over the same keys
  uint8_t* valid = hmalloc<uint8_t>(n);
  int32_t* vals = hmalloc<int32_t>(n);
  for (size_t i = 0; i < n; ++i) {
    valid[i] = uint8_t(rng() % 10 != 0);
    vals[i] = int32_t(rng() % 65536) - 32768;
  }
  CHECK(nfh_unique_valid_i32(k, valid, n, uniq, inv) >= 0);
  for (size_t i = 0; i < n; ++i) CHECK(valid[i] ? (inv[i] >= 0 && uniq[inv[i]] == k[i]) : inv[i] == -1);
  int64_t* uk = hmalloc<int64_t>(n);
  int64_t* cnt = hmalloc<int64_t>(n);
  int64_t* sm = hmalloc<int64_t>(n);
  int32_t* mn = hmalloc<int32_t>(n);
  int32_t* mx = hmalloc<int32_t>(n);
  double* mean = hmalloc<double>(n);
  const int64_t g = nfh_group_agg_i32(k, vals, valid, nullptr, n, uk, cnt, sm, mn, mx, mean, n);
  CHECK(g >= 0);
  int64_t tot = 0, tsum = 0, rsum = 0, rcnt = 0;
  for (int64_t j = 0; j < g; ++j) {
    tot += cnt[j];
    tsum += sm[j];
    CHECK(mn[j] <= mx[j]);
    if (j) CHECK(uk[j] > uk[j - 1]);
  }
  for (size_t i = 0; i < n; ++i)
    if (valid[i]) {
      ++rcnt;
      rsum += vals[i];
    }
  CHECK(tot == rcnt && tsum == rsum);
  // capacity too small: -7 with nothing written past cap
  if (g > 1) CHECK(nfh_group_agg_i32(k, vals, valid, nullptr, n, uk, cnt, sm, mn, mx, mean,
                                     size_t(g - 1)) == -7);
  for (void* p : {(void*)kb, (void*)perm, (void*)t0, (void*)t1, (void*)tp, (void*)uniq,
                  (void*)inv, (void*)valid, (void*)vals, (void*)uk, (void*)cnt, (void*)sm,
                  (void*)mn, (void*)mx, (void*)mean})
    std::free(p);
}

static void dense_lanes(size_t n, uint32_t seed) {
  std::mt19937 rng(seed);
  const size_t g = 37;
  int32_t* k = hmalloc<int32_t>(n);
  double* v = hmalloc<double>(3 * n);
  int32_t* iv = hmalloc<int32_t>(2 * n);
  for (size_t i = 0; i < n; ++i) {
    k[i] = int32_t(rng() % g);
    for (int c = 0; c < 3; ++c) v[c * n + i] = double(int(rng() % 2001) - 1000) * 0.5;
    for (int c = 0; c < 2; ++c) iv[c * n + i] = int32_t(rng() % 2001) - 1000;
  }
  double* s = hmalloc<double>(3 * g);
  int64_t* c = hmalloc<int64_t>(g);
  CHECK(nf_group_sum_count(k, v, n, s, c, g) == 0);
  std::vector<double> rs(g, 0.0);
  std::vector<int64_t> rc(g, 0);
  for (size_t i = 0; i < n; ++i) {
    rs[k[i]] += v[i];  // small halves: exact in any order
    ++rc[k[i]];
  }
  for (size_t j = 0; j < g; ++j) CHECK(s[j] == rs[j] && c[j] == rc[j]);
  CHECK(nf_group_multi_sum_count(k, v, n, 3, s, c, g) == 0);
  int32_t* is = hmalloc<int32_t>(2 * g);
  CHECK(nf_group_mixed_sum_only(k, iv, v, n, 2, 3, is, s, g) == 0);
  CHECK(nf_group_count_only(k, n, c, g) == 0);
  for (size_t j = 0; j < g; ++j) CHECK(c[j] == rc[j]);
  int32_t* s1 = hmalloc<int32_t>(g);
  int32_t* s2 = hmalloc<int32_t>(g);
  CHECK(nf_group_owner_2i32_1f64(k, iv, iv + n, v, n, 5, 20, s1, s2, s, c, g) == 0);
  CHECK(nf_group_owner_2i32_1f64(k, iv, iv + n, v, n, 20, 5, s1, s2, s, c, g) == -1);
  int32_t* s32 = hmalloc<int32_t>(g);
  CHECK(nf_group_variant_i32(k, iv, n, 0, s32, c, g) == 0);
  CHECK(nf_group_variant_i32(k, iv, n, 1, s32, c, g) == -3);
  // a key outside [0, g): -2 (the reference aborts in nf_group_sum_count)
  k[n / 2] = int32_t(g);
  CHECK(nf_group_sum_count(k, v, n, s, c, g) == -2);
  // float64 / int64 group-by with masks, host buffers
  k[n / 2] = 0;
  uint8_t* m = hmalloc<uint8_t>(n);
  for (size_t i = 0; i < n; ++i) m[i] = uint8_t(rng() % 7 != 0);
  int64_t* uk = hmalloc<int64_t>(g);
  int64_t* ct = hmalloc<int64_t>(g);
  double* fs = hmalloc<double>(g);
  double* fmn = hmalloc<double>(g);
  double* fmx = hmalloc<double>(g);
  double* fme = hmalloc<double>(g);
  const int64_t gf = nfh_group_agg_f64(k, v, m, nullptr, n, uk, ct, fs, fmn, fmx, fme, g);
  CHECK(gf >= 0);
  if (gf > 1)
    CHECK(nfh_group_agg_f64(k, v, m, nullptr, n, uk, ct, fs, fmn, fmx, fme, size_t(gf - 1)) ==
          -7);
  // gather with validity, and its refusal
  int32_t* idx = hmalloc<int32_t>(n);
  for (size_t i = 0; i < n; ++i) idx[i] = int32_t(rng() % n);
  double* out = hmalloc<double>(n);
  uint8_t* vo = hmalloc<uint8_t>(n);
  CHECK(nfh_gather(v, n, 8, idx, 4, n, out, m, vo) == 0);
  for (size_t i = 0; i < n; ++i) CHECK(out[i] == v[idx[i]] && vo[i] == m[idx[i]]);
  idx[n - 1] = int32_t(n);
  CHECK(nfh_gather(v, n, 8, idx, 4, n, out, m, vo) == -2);
  // selection
  uint8_t* mo = hmalloc<uint8_t>(n);
  CHECK(nf_mask_not(m, n, mo) == 0);
  const int64_t kept = nf_select_count(m, n);
  CHECK(nf_select_scatter_f64(v, m, n, out) == kept);
  for (void* p : {(void*)k, (void*)v, (void*)iv, (void*)s, (void*)c, (void*)is, (void*)s1,
                  (void*)s2, (void*)s32, (void*)m, (void*)uk, (void*)ct, (void*)fs, (void*)fmn,
                  (void*)fmx, (void*)fme, (void*)idx, (void*)out, (void*)vo, (void*)mo})
    std::free(p);
}

static void refusals() {
  int32_t x = 0;
  CHECK(nf_sort_perm_i32(nullptr, 5, 0, &x, nullptr, nullptr, nullptr) == -1);
  CHECK(nf_sort_perm_i32(&x, 0, 0, nullptr, nullptr, nullptr, nullptr) == 0);
  CHECK(nfh_group_agg_i32(nullptr, &x, nullptr, nullptr, 1, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, 1) == -1);
  CHECK(nfh_sort_perm_i32_noalloc(&x, 1, 0, &x, reinterpret_cast<uint32_t*>(&x),
                                  reinterpret_cast<uint32_t*>(&x), &x) == -3);  // host buffers
  CHECK(nfh_gather(&x, 1, 3, &x, 4, 1, &x, nullptr, nullptr) == -3);
}

int main() {
  if (!nfh_device_ok()) {
    std::fprintf(stderr, "no usable GPU\n");
    return 2;
  }
  for (size_t n : {1ul, 7ul, 4099ul, 100003ul}) {
    sort_and_unique(n, uint32_t(n), 0);
    sort_and_unique(n, uint32_t(n) + 1, 1);
    dense_lanes(n, uint32_t(n) + 2);
  }
  refusals();
  // the thread-pool pattern: disjoint slices from 8 threads at once
  std::vector<std::thread> th;
  for (int t = 0; t < 8; ++t)
    th.emplace_back([t] {
      sort_and_unique(50000 + 997 * size_t(t), 100u + uint32_t(t), t & 1);
      dense_lanes(30000 + 131 * size_t(t), 200u + uint32_t(t));
    });
  for (auto& x : th) x.join();
  std::printf("boundary under ASan: %s (%d failed checks)\n", g_fail ? "FAIL" : "ok", g_fail);
  return g_fail ? 1 : 0;
}
