// onesweep.hpp -- one stable LSD digit pass (onesweep: per-tile ranking,
// decoupled look-back, LDS reorder, run-wise write-out), shared by the radix
// argsort and the learned-CDF bucket sort.
#pragma once
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

// The plain radix digit: byte `shift / 8` of the key word.
struct RadixDigit {
  int shift;
  static constexpr size_t EXTRA_LDS = 0;
  __device__ void init(unsigned char*) const {}
  template <class W>
  __device__ uint32_t digit(W w, const unsigned char*) const {
    return uint32_t(w >> shift) & 255u;
  }
};

// ------------------------------------------------------------ one pass
// FIRST: read raw keys (convert to words) and generate row ids.
// LAST: write only the permutation (+ sorted keys when kout != nullptr).
// LR: rank with lane-ordered LDS atomics (rank_lds).
// DG: digit generator -- `digit(w)` gives the pass's 8-bit digit of a key
// word (RadixDigit: one byte of the word; the learned sort passes a
// generator over its monotone bucket code).  DG::EXTRA_LDS bytes of dynamic
// LDS after the tile buffers are DG's own, filled by `init`.
template <class K, bool FIRST, bool LAST, bool LR, class DG>
__global__ __launch_bounds__(SortGeom<K>::BLOCK) void k_onesweep(
    const K* __restrict__ keys, const typename Key<K>::W* __restrict__ win,
    const int32_t* __restrict__ iin, typename Key<K>::W* __restrict__ wout,
    int32_t* __restrict__ iout, K* __restrict__ kout, const uint32_t* __restrict__ bin_start,
    uint64_t* __restrict__ lookback, uint32_t* __restrict__ tile_ctr, uint32_t n, DG dg,
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
  unsigned char* s_dg = s_raw + G::LDS;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = wave_id();
  if (tid == 0) *s_tile = atomicAdd(tile_ctr, 1u);
  for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) s_wh[i >> 8][i & 255] = 0;
  dg.init(s_dg);
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
  // rk[j] keeps (rank << 8) | digit, so the scatter needs no second digit
  // evaluation (a wave's rank is below 64 * SORT_IPT).
  uint32_t rk[SORT_IPT];
#pragma unroll
  for (int j = 0; j < SORT_IPT; ++j) {
    const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
    const bool valid = pos < n;
    const uint32_t d = dg.digit(kw[j], s_dg);
    if constexpr (LR) {
      rk[j] = (rank_lds(s_wh[wave], d, valid) << 8) | d;
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
      rk[j] = ((old + r) << 8) | d;
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
      const uint32_t d = rk[j] & 255u;
      const uint32_t p = s_tstart[d] + s_wh[wave][d] + (rk[j] >> 8);
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
    const uint32_t d = dg.digit(w, s_dg);
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

template <class K, bool FIRST, bool LAST, class DG>
void onesweep_pass(const K* keys, const word_t<K>* win, const int32_t* iin, word_t<K>* wout,
                   int32_t* iout, K* kout, const uint32_t* bin_start, uint64_t* lb,
                   uint32_t* ctr, uint32_t n, const DG& dg, bool desc, hipStream_t s,
                   const char* prof_name = "radix_onesweep") {
  using G = SortGeom<K>;
  constexpr size_t LDS = G::LDS + DG::EXTRA_LDS;
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, true, DG>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(LDS)));
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, false, DG>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(LDS)));
  });
  const bool lr = lds_rank_supported();
  const uint32_t tiles = (n + G::TILE - 1) / G::TILE;
  const uint32_t epoch = lookback_epoch(s);
  // pass model: first pass reads keys, later ones (word, row id); the last
  // writes row ids (+ keys for unique), the others (word, row id)
  const double in_b = FIRST ? sizeof(K) : sizeof(typename Key<K>::W) + 4.0;
  const double out_b = LAST ? 4.0 + (kout ? sizeof(K) : 0.0) : sizeof(typename Key<K>::W) + 4.0;
  ProfScope prof(prof_name, s, double(n) * (in_b + out_b));
  if (lr)
    hipLaunchKernelGGL((k_onesweep<K, FIRST, LAST, true, DG>), dim3(tiles), dim3(G::BLOCK), LDS,
                       s, keys, win, iin, wout, iout, kout, bin_start, lb, ctr, n, dg, desc,
                       epoch);
  else
    hipLaunchKernelGGL((k_onesweep<K, FIRST, LAST, false, DG>), dim3(tiles), dim3(G::BLOCK), LDS,
                       s, keys, win, iin, wout, iout, kout, bin_start, lb, ctr, n, dg, desc,
                       epoch);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh
