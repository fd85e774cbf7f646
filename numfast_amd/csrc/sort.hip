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
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

template <class K>
struct Key;
template <>
struct Key<int32_t> {
  using W = uint32_t;
  static constexpr int PASSES = 4;
  __device__ static W word(int32_t k, bool desc) {
    W w = W(k) ^ 0x80000000u;
    return desc ? ~w : w;
  }
  __device__ static int32_t key(W w, bool desc) {
    if (desc) w = ~w;
    return int32_t(w ^ 0x80000000u);
  }
};
template <>
struct Key<int64_t> {
  using W = uint64_t;
  static constexpr int PASSES = 8;
  __device__ static W word(int64_t k, bool desc) {
    W w = W(k) ^ 0x8000000000000000ull;
    return desc ? ~w : w;
  }
  __device__ static int64_t key(W w, bool desc) {
    if (desc) w = ~w;
    return int64_t(w ^ 0x8000000000000000ull);
  }
};

// Tiles are large (16384 int32 keys, 8192 int64 keys) so that each digit's
// run inside a tile spans whole 128-byte lines when written out: smaller
// tiles leave partial lines, which cost a fill read plus a partial
// write-back per line (measured with FETCH_SIZE / WRITE_SIZE).
constexpr int SORT_IPT = 16;
template <class K>
struct SortGeom {
  static constexpr int BLOCK = sizeof(K) == 4 ? 1024 : 512;
  static constexpr int WAVES = BLOCK / WAVE;
  static constexpr int TILE = BLOCK * SORT_IPT;
  static constexpr size_t LDS =
      size_t(TILE) * (sizeof(K) + 4) + size_t(WAVES) * 256 * 4 + 256 * 4 * 2 + WAVES * 4 + 16;
};

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

// One block per digit position: exclusive scan of its 256 bins.
__global__ __launch_bounds__(256) void k_bins(const uint32_t* __restrict__ hist, uint32_t n,
                                              uint32_t* __restrict__ bin_start,
                                              uint32_t* __restrict__ trivial) {
  __shared__ uint32_t s_tmp[4];
  const int p = blockIdx.x;
  const uint32_t c = hist[p * 256 + threadIdx.x];
  uint32_t total;
  const uint32_t ex = block_excl_scan<4>(c, s_tmp, &total);
  bin_start[p * 256 + threadIdx.x] = ex;
  if (threadIdx.x == 0) trivial[p] = 0;
  __syncthreads();
  if (c == n) trivial[p] = 1;
}

// ------------------------------------------------------------ one pass
// FIRST: read raw keys (convert to words) and generate row ids.
// LAST: write only the permutation (+ sorted keys when kout != nullptr).
// LR: rank with lane-ordered LDS atomics (rank_lds).
template <class K, bool FIRST, bool LAST, bool LR>
__global__ __launch_bounds__(SortGeom<K>::BLOCK) void k_onesweep(
    const K* __restrict__ keys, const typename Key<K>::W* __restrict__ win,
    const int32_t* __restrict__ iin, typename Key<K>::W* __restrict__ wout,
    int32_t* __restrict__ iout, K* __restrict__ kout, const uint32_t* __restrict__ bin_start,
    uint64_t* __restrict__ lookback, uint32_t* __restrict__ tile_ctr, uint32_t n, int shift,
    bool desc, uint32_t epoch) {
  using T = Key<K>;
  using W = typename T::W;
  using G = SortGeom<K>;
  constexpr int SORT_BLOCK = G::BLOCK, SORT_WAVES = G::WAVES, SORT_TILE = G::TILE;
  extern __shared__ __align__(16) unsigned char s_raw[];
  W* s_w = reinterpret_cast<W*>(s_raw);
  uint32_t* s_i = reinterpret_cast<uint32_t*>(s_w + SORT_TILE);
  uint32_t(*s_wh)[256] = reinterpret_cast<uint32_t(*)[256]>(s_i + SORT_TILE);
  uint32_t* s_tstart = &s_wh[SORT_WAVES][0];
  uint32_t* s_gbase = s_tstart + 256;
  uint32_t* s_tmp = s_gbase + 256;
  uint32_t* s_tile = s_tmp + SORT_WAVES;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = wave_id();
  if (tid == 0) *s_tile = atomicAdd(tile_ctr, 1u);
  for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) s_wh[i >> 8][i & 255] = 0;
  __syncthreads();
  const uint32_t tile = *s_tile;
  const uint32_t base = tile * uint32_t(SORT_TILE);
  const uint32_t wbase = base + uint32_t(wave) * (WAVE * SORT_IPT);

  W kw[SORT_IPT];
  uint32_t ix[SORT_IPT];
#pragma unroll
  for (int j = 0; j < SORT_IPT; ++j) {
    const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
    if (pos < n) {
      if constexpr (FIRST) {
        kw[j] = T::word(keys[pos], desc);
        ix[j] = pos;
      } else {
        kw[j] = win[pos];
        ix[j] = uint32_t(iin[pos]);
      }
    } else {
      kw[j] = 0;
      ix[j] = 0;
    }
  }

  // Stable in-wave ranks: item j of lane l precedes item j of lane l+1,
  // which precedes item j+1 of lane 0 (that is the position order).  LR:
  // one lane-ordered LDS atomic per item; else the 8-ballot digit match.
  uint32_t rk[SORT_IPT];
#pragma unroll
  for (int j = 0; j < SORT_IPT; ++j) {
    const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
    const bool valid = pos < n;
    const uint32_t d = uint32_t(kw[j] >> shift) & 255u;
    if constexpr (LR) {
      rk[j] = rank_lds(s_wh[wave], d, valid);
    } else {
      const uint64_t peers = match_digit8(d, valid);
      const uint32_t r = __popcll(peers & lanemask_lt());
      const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : lane;
      uint32_t old = 0;
      if (valid && r == 0) {
        old = s_wh[wave][d];
        s_wh[wave][d] = old + uint32_t(__popcll(peers));
      }
      old = __shfl(old, leader);
      rk[j] = old + r;
    }
  }
  __syncthreads();

  // Per digit (threads 0..255, thread == digit): wave-exclusive offsets and
  // the tile's count, published at once for later tiles' look-back.
  const bool digit_thread = tid < 256;
  uint32_t cnt = 0;
  uint64_t* my_lb = lookback + size_t(tile) * 256 + (digit_thread ? tid : 0);
  if (digit_thread) {
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) {
      const uint32_t c = s_wh[w][tid];
      s_wh[w][tid] = cnt;
      cnt += c;
    }
    lb_store(my_lb, lb_pack(epoch, tile == 0 ? LB_PREFIX : LB_AGG, cnt));
  }
  uint32_t tile_total;
  const uint32_t tstart = block_excl_scan<SORT_WAVES>(cnt, s_tmp, &tile_total);
  if (digit_thread) s_tstart[tid] = tstart;
  __syncthreads();

  // Reorder the tile in LDS by digit (stable).
#pragma unroll
  for (int j = 0; j < SORT_IPT; ++j) {
    const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
    if (pos < n) {
      const uint32_t d = uint32_t(kw[j] >> shift) & 255u;
      const uint32_t p = s_tstart[d] + s_wh[wave][d] + rk[j];
      s_w[p] = kw[j];
      s_i[p] = ix[j];
    }
  }

  // Decoupled look-back for this digit's offset among earlier tiles.
  if (digit_thread) {
    uint32_t excl = 0;
    if (tile > 0) {
      excl = lb_lookback(lookback + tid, tile, 256, epoch);
      lb_store(my_lb, lb_pack(epoch, LB_PREFIX, excl + cnt));
    }
    s_gbase[tid] = bin_start[tid] + excl - tstart;
  }
  __syncthreads();

  const uint32_t nvalid = (n - base) < uint32_t(SORT_TILE) ? (n - base) : uint32_t(SORT_TILE);
  for (uint32_t i = tid; i < nvalid; i += SORT_BLOCK) {
    const W w = s_w[i];
    const uint32_t d = uint32_t(w >> shift) & 255u;
    const uint32_t g = s_gbase[d] + i;
    if constexpr (LAST) {
      iout[g] = int32_t(s_i[i]);
      if (kout != nullptr) kout[g] = T::key(w, desc);
    } else {
      wout[g] = w;
      iout[g] = int32_t(s_i[i]);
    }
  }
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

template <class K, bool FIRST, bool LAST>
static void launch_pass(const K* keys, const typename Key<K>::W* win, const int32_t* iin,
                        typename Key<K>::W* wout, int32_t* iout, K* kout,
                        const uint32_t* bin_start, uint64_t* lb, uint32_t* ctr, uint32_t n,
                        int shift, bool desc, hipStream_t s) {
  using G = SortGeom<K>;
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(G::LDS)));
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, false>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(G::LDS)));
  });
  const bool lr = lds_rank_supported();
  const uint32_t tiles = (n + G::TILE - 1) / G::TILE;
  const uint32_t epoch = lookback_epoch(s);
  // pass model: first pass reads keys, later ones (word, row id); the last
  // writes row ids (+ keys for unique), the others (word, row id)
  const double in_b = FIRST ? sizeof(K) : sizeof(typename Key<K>::W) + 4.0;
  const double out_b = LAST ? 4.0 + (kout ? sizeof(K) : 0.0) : sizeof(typename Key<K>::W) + 4.0;
  ProfScope prof("radix_onesweep", s, double(n) * (in_b + out_b));
  if (lr)
    hipLaunchKernelGGL((k_onesweep<K, FIRST, LAST, true>), dim3(tiles), dim3(G::BLOCK), G::LDS, s,
                       keys, win, iin, wout, iout, kout, bin_start, lb, ctr, n, shift, desc, epoch);
  else
    hipLaunchKernelGGL((k_onesweep<K, FIRST, LAST, false>), dim3(tiles), dim3(G::BLOCK), G::LDS,
                       s, keys, win, iin, wout, iout, kout, bin_start, lb, ctr, n, shift, desc,
                       epoch);
  NFH_LAUNCH_CHECK();
}

template <class K>
void radix_argsort(const K* keys, size_t n_, bool desc, int32_t* perm, word_t<K>* t0,
                   word_t<K>* t1, int32_t* tp, K* sorted_out, hipStream_t s) {
  using T = Key<K>;
  using W = typename T::W;
  constexpr int P = T::PASSES;
  const uint32_t n = uint32_t(n_);
  if (n == 0) return;
  // meta: hist[P*256] | bin_start[P*256] | tile counters[P] | trivial[P]
  uint32_t* meta = scratch_t<uint32_t>(S_META, size_t(P) * 512 + 2 * P);
  uint32_t* hist = meta;
  uint32_t* bins = meta + P * 256;
  uint32_t* ctrs = meta + P * 512;
  uint32_t* triv = ctrs + P;
  fill_u32(meta, 0, size_t(P) * 512 + 2 * P, s);
  const unsigned hb = grid_for(n, 256, 4u * unsigned(num_cus()));
  {
    ProfScope prof("radix_hist", s, double(n) * sizeof(K));
    hipLaunchKernelGGL((k_hist<K>), dim3(hb), dim3(256), 0, s, keys, n, desc, hist);
    NFH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_bins, dim3(P), dim3(256), 0, s, hist, n, bins, triv);
  NFH_LAUNCH_CHECK();
  uint32_t* h_triv = static_cast<uint32_t*>(pinned(P * sizeof(uint32_t)));
  NFH_CHECK(hipMemcpyAsync(h_triv, triv, P * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  int active[P];
  int na = 0;
  for (int p = 0; p < P; ++p)
    if (!h_triv[p]) active[na++] = p;
  if (na == 0) {
    hipLaunchKernelGGL((k_iota<K>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, perm,
                       sorted_out, keys, n);
    NFH_LAUNCH_CHECK();
    return;
  }
  const size_t tiles = (n + SortGeom<K>::TILE - 1) / SortGeom<K>::TILE;
  uint64_t* lb = lookback_words(tiles * 256, s);
  // Intermediate pairs: A = (t1, tp), B = (t0, perm).  The last pass must
  // read A so that it can write perm; k = na-1 intermediate outputs.
  const int k = na - 1;
  const W* cw = nullptr;
  const int32_t* ci = nullptr;
  for (int a = 0; a < na; ++a) {
    const int p = active[a];
    const bool first = (a == 0), last = (a == na - 1);
    const uint32_t* bs = bins + p * 256;
    uint32_t* ctr = ctrs + p;
    if (last) {
      if (first)
        launch_pass<K, true, true>(keys, nullptr, nullptr, nullptr, perm, sorted_out, bs, lb,
                                   ctr, n, 8 * p, desc, s);
      else
        launch_pass<K, false, true>(nullptr, cw, ci, nullptr, perm, sorted_out, bs, lb, ctr,
                                    n, 8 * p, desc, s);
    } else {
      const bool use_a = ((k - 1 - a) % 2) == 0;
      W* ow = use_a ? t1 : t0;
      int32_t* oi = use_a ? tp : perm;
      if (first)
        launch_pass<K, true, false>(keys, nullptr, nullptr, ow, oi, nullptr, bs, lb, ctr, n,
                                    8 * p, desc, s);
      else
        launch_pass<K, false, false>(nullptr, cw, ci, ow, oi, nullptr, bs, lb, ctr, n, 8 * p,
                                     desc, s);
      cw = ow;
      ci = oi;
    }
  }
}

template void radix_argsort<int32_t>(const int32_t*, size_t, bool, int32_t*, uint32_t*,
                                     uint32_t*, int32_t*, int32_t*, hipStream_t);
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
