// sort.hip -- stable LSD radix argsort (nf_sort_perm_i32/_i64) for gfx950.
//
// Semantics: numfast-native/src/sort.rs:3-11, :55-118.  Ascending equals
// np.argsort(kind="stable"); descending is a stable ascending sort of the
// inverted order-preserving word !(k ^ signbit), ties in input order.
//
// Design (one pass per 8-bit digit, LSD):
//   1. k_hist: one read of the keys builds all digit histograms at once
//      (LDS-privatised counters, one global atomic per non-empty bin).
//   2. k_bins: exclusive scan per digit position -> global bin starts, and
//      a flag for digit positions where every key shares one bin (such a
//      pass is the identity and is skipped).
//   3. k_onesweep per active pass: a workgroup claims a 16384-key tile
//      (atomic tile counter, so every earlier tile is already resident),
//      ranks keys stably inside the tile with per-wave LDS digit counters
//      (one lane-ordered LDS atomic per key, or a wave64 8-ballot digit
//      match where that LDS property is absent), publishes its per-digit counts and walks
//      back over earlier tiles (decoupled look-back) for its global offsets,
//      reorders the tile in LDS and writes it out in digit runs.
// The first active pass reads the raw keys and generates row ids; the last
// one writes only the permutation (and, for unique, the sorted keys).  So
// the traffic per key is 4 (histogram) + 12 + 16 + 16 + 12 bytes for int32.
#include <mutex>

#include "device.hpp"
#include "onesweep.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

// ------------------------------------------------------------ histograms
template <class K>
__global__ __launch_bounds__(256) void k_hist(const K* __restrict__ keys, uint32_t n,
                                              bool desc, uint32_t* __restrict__ hist) {
  using T = Key<K>;
  constexpr int P = T::PASSES;
  __shared__ uint32_t s_h[P * 256];
  for (int i = threadIdx.x; i < P * 256; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const typename T::W w = T::word(keys[i], desc);
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&s_h[p * 256 + (uint32_t(w >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * 256; i += blockDim.x)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

template <class K>
__global__ void k_iota(int32_t* __restrict__ perm, K* __restrict__ kout,
                       const K* __restrict__ keys, uint32_t n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    perm[i] = int32_t(i);
    if (kout != nullptr) kout[i] = keys[i];
  }
}

// ------------------------------------------------------------ host driver
uint64_t* lookback_words(size_t words, hipStream_t s) {
  Ctx& c = ctx();
  const size_t bytes = words * sizeof(uint64_t);
  const bool grow = c.cap[S_LOOKBACK] < bytes;
  uint64_t* p = static_cast<uint64_t*>(scratch(S_LOOKBACK, bytes));
  if (grow) fill_u64(p, 0, c.cap[S_LOOKBACK] / sizeof(uint64_t), s);
  return p;
}

uint32_t lookback_epoch(hipStream_t s) {
  Ctx& c = ctx();
  const uint32_t e = next_epoch();
  if (e == 1 && c.slot[S_LOOKBACK] != nullptr)  // wrapped: forget old words
    fill_u64(c.slot[S_LOOKBACK], 0, c.cap[S_LOOKBACK] / sizeof(uint64_t), s);
  return e;
}

static_assert(SORT_NCHAIN == NCHAIN, "chains of the sort workspace");
static_assert(SORT_TILE_I32 == uint32_t(SortGeom<int32_t>::TILE), "int32 sort tile");

template <class K>
size_t sort_lookback_words(size_t n) {
  return (n + SortGeom<K>::TILE - 1) / SortGeom<K>::TILE * 256;
}
template size_t sort_lookback_words<int32_t>(size_t);
template size_t sort_lookback_words<int64_t>(size_t);

template <class K>
void radix_argsort_ws(const K* keys, size_t n_, bool desc, int32_t* perm, word_t<K>* t0,
                      word_t<K>* t1, int32_t* tp, K* sorted_out, const SortWs& ws,
                      hipStream_t s) {
  using T = Key<K>;
  using W = typename T::W;
  constexpr int P = T::PASSES;
  constexpr size_t CH = size_t(NCHAIN) * 256;
  const uint32_t n = uint32_t(n_);
  if (n == 0) return;
  // meta: or / and words | per pass: per-chain hist | per pass: bins | tile counters
  constexpr size_t OW = 2 * sizeof(W) / 4;
  uint32_t* meta = ws.meta;
  W* orand = reinterpret_cast<W*>(meta);
  uint32_t* hist = meta + OW;
  uint32_t* bins = hist + P * CH;
  uint32_t* ctrs = bins + P * CH;
  fill_u32(hist, 0, 2 * P * CH + P * NCHAIN, s);
  int active[P];
  int na = 0;
  if (ws.all_passes) {
    // every digit pass runs (a digit all keys share is an identity pass):
    // no OR / AND readback, so no host synchronisation
    chain_hist_first<K>(keys, n, RadixDigit{0, 0}, desc, hist, static_cast<W*>(nullptr), s,
                        "radix_hist");
    for (int p = 0; p < P; ++p) active[na++] = p;
  } else {
    fill_u32(meta, 0, OW / 2, s);
    fill_u32(meta + OW / 2, 0xFFFFFFFFu, OW / 2, s);
    // per-chain histogram of digit 0 and the OR / AND of all key words
    chain_hist_first<K>(keys, n, RadixDigit{0, 0}, desc, hist, orand, s, "radix_hist");
    W* h_oa = static_cast<W*>(pinned(2 * sizeof(W)));
    NFH_CHECK(hipMemcpyAsync(h_oa, orand, 2 * sizeof(W), hipMemcpyDeviceToHost, s));
    NFH_CHECK(hipStreamSynchronize(s));
    const W differ = h_oa[0] ^ h_oa[1];
    for (int p = 0; p < P; ++p)  // a digit every key shares is an identity pass
      if ((differ >> (8 * p)) & W(255)) active[na++] = p;
  }
  if (na == 0) {
    hipLaunchKernelGGL((k_iota<K>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, perm,
                       sorted_out, keys, n);
    NFH_LAUNCH_CHECK();
    return;
  }
  if (active[0] != 0)  // rare: digit 0 is shared, count the first active one
    chain_hist_first<K>(keys, n, RadixDigit{8 * active[0], 0}, desc, hist + CH,
                        static_cast<W*>(nullptr), s, "radix_hist");
  uint64_t* lb = ws.lb;
  // Intermediate pairs: A = (t1, tp), B = (t0, perm).  The last pass must
  // read A so that it can write perm; k = na-1 intermediate outputs.
  // Pass a's per-chain counts are hist + CH * slot(a): slot 0 holds digit
  // 0's from the first histogram, so pass a uses slot a + (active[0] != 0).
  const int k = na - 1;
  const int off = active[0] != 0 ? 1 : 0;
  const W* cw = nullptr;
  const int32_t* ci = nullptr;
  for (int a = 0; a < na; ++a) {
    const int p = active[a];
    const bool first = (a == 0), last = (a == na - 1);
    uint32_t* h = hist + CH * size_t(a + off);
    uint32_t* nh = last ? nullptr : hist + CH * size_t(a + 1 + off);
    uint32_t* bs = bins + CH * size_t(a);
    uint32_t* ctr = ctrs + NCHAIN * a;
    const uint32_t ep = ws.epoch0 != 0 ? ws.epoch0 + uint32_t(a) : 0u;
    const RadixDigit dg{8 * p, last ? 0 : 8 * active[a + 1]};
    hipLaunchKernelGGL(k_chain_bins<256>, dim3(1), dim3(256), 0, s, h, bs);
    NFH_LAUNCH_CHECK();
    if (last) {
      if (first)
        onesweep_pass<K, true, true>(keys, nullptr, nullptr, nullptr, perm, sorted_out, bs, lb,
                                     ctr, n, dg, nh, desc, s, "radix_onesweep", ep);
      else
        onesweep_pass<K, false, true>(nullptr, cw, ci, nullptr, perm, sorted_out, bs, lb, ctr,
                                      n, dg, nh, desc, s, "radix_onesweep", ep);
    } else {
      const bool use_a = ((k - 1 - a) % 2) == 0;
      W* ow = use_a ? t1 : t0;
      int32_t* oi = use_a ? tp : perm;
      if (first)
        onesweep_pass<K, true, false>(keys, nullptr, nullptr, ow, oi, nullptr, bs, lb, ctr, n,
                                      dg, nh, desc, s, "radix_onesweep", ep);
      else
        onesweep_pass<K, false, false>(nullptr, cw, ci, ow, oi, nullptr, bs, lb, ctr, n, dg, nh,
                                       desc, s, "radix_onesweep", ep);
      cw = ow;
      ci = oi;
    }
  }
}

template <class K>
void radix_argsort(const K* keys, size_t n, bool desc, int32_t* perm, word_t<K>* t0,
                   word_t<K>* t1, int32_t* tp, K* sorted_out, hipStream_t s) {
  if (n == 0) return;
  SortWs ws;
  ws.meta = scratch_t<uint32_t>(S_META, sort_meta_words<K>());
  ws.lb = lookback_words(sort_lookback_words<K>(n), s);
  radix_argsort_ws<K>(keys, n, desc, perm, t0, t1, tp, sorted_out, ws, s);
}

void radix_histograms_i32(const int32_t* keys, uint32_t n, uint32_t* hist, hipStream_t s) {
  fill_u32(hist, 0, 4 * 256, s);
  if (n == 0) return;
  ProfScope prof("radix_hist", s, double(n) * 4.0);
  hipLaunchKernelGGL((k_hist<int32_t>), dim3(grid_for(n, 256, 4u * unsigned(num_cus()))),
                     dim3(256), 0, s, keys, n, false, hist);
  NFH_LAUNCH_CHECK();
}

template void radix_argsort<int32_t>(const int32_t*, size_t, bool, int32_t*, uint32_t*,
                                     uint32_t*, int32_t*, int32_t*, hipStream_t);
template void radix_argsort_ws<int32_t>(const int32_t*, size_t, bool, int32_t*, uint32_t*,
                                        uint32_t*, int32_t*, int32_t*, const SortWs&,
                                        hipStream_t);
template void radix_argsort<int64_t>(const int64_t*, size_t, bool, int32_t*, uint64_t*,
                                     uint64_t*, int32_t*, int64_t*, hipStream_t);

}  // namespace nfh

using namespace nfh;

// lib.rs:1326-1357 -- stable radix permutation of int32 keys.
NF_EXPORT int32_t nf_sort_perm_i32(const int32_t* keys, size_t n, uint8_t descending,
                                   int32_t* perm_out, uint32_t* tmp0, uint32_t* tmp1,
                                   int32_t* tmp_p) {
  if (n == 0) return OK;
  if (!keys || !perm_out || !tmp0 || !tmp1 || !tmp_p) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  int32_t* p = st.out(perm_out, n, S_OUT0);
  uint32_t* a = st.tmp(tmp0, n, S_T0);
  uint32_t* b = st.tmp(tmp1, n, S_T1);
  int32_t* c = st.tmp(tmp_p, n, S_T2);
  radix_argsort<int32_t>(k, n, descending != 0, p, a, b, c, nullptr, st.s);
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

// lib.rs:1363-1394 -- int64 keys, 8 digit passes, int32 positions.
NF_EXPORT int32_t nf_sort_perm_i64(const int64_t* keys, size_t n, uint8_t descending,
                                   int32_t* perm_out, uint64_t* tmp0, uint64_t* tmp1,
                                   int32_t* tmp_p) {
  if (n == 0) return OK;
  if (!keys || !perm_out || !tmp0 || !tmp1 || !tmp_p) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int64_t* k = st.in(keys, n, S_IN0);
  int32_t* p = st.out(perm_out, n, S_OUT0);
  uint64_t* a = st.tmp(tmp0, n, S_T0);
  uint64_t* b = st.tmp(tmp1, n, S_T1);
  int32_t* c = st.tmp(tmp_p, n, S_T2);
  radix_argsort<int64_t>(k, n, descending != 0, p, a, b, c, nullptr, st.s);
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}
