// onesweep.hpp -- one stable LSD digit pass (onesweep: per-tile ranking,
// decoupled look-back, LDS reorder, run-wise write-out), shared by the radix
// argsort and the learned-CDF bucket sort.
//
// Chains.  MI355X dispatches workgroup b of a launch to XCD b % 8, and each
// XCD has its own L2.  A single look-back chain over all tiles serialises
// the pass on the prefix hand-off between XCDs (a middle pass of 2^28 keys
// took 1.20 ms with one chain, 0.78 ms with no look-back at all).  So the
// tiles are split into NCHAIN = 8 contiguous ranges ("chains"), and chain c
// is processed only by workgroups b with b % 8 == c, in tile order (a
// per-chain atomic tile counter), with a look-back chain of its own that
// starts at the chain's first tile.  The global start of digit d in chain c
// is  sum_{d' < d} count(d') + sum_{c' < c} count_c'(d),  from per-chain
// digit histograms of the pass's input (0.90 ms for the same pass).
// Adjacent tiles of a chain run on one XCD, so the two partial 128-byte
// lines at the edges of a digit run merge in that XCD's L2.
//
// The per-chain histogram of the first pass comes from k_chain_hist over
// the keys.  Each later pass's comes from the pass before it: the write-out
// loop knows every item's output position, hence its chain in the next
// pass, and counts the item's next digit there in LDS.  The pass kernel is
// persistent (one workgroup per CU; a workgroup loops over its chain's
// tiles), so those counts reach global memory once per workgroup.
#pragma once
#include <algorithm>
#include <mutex>
#include <type_traits>

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

// ------------------------------------------------------------ chains
constexpr int NCHAIN = 8;

// Chain c of a pass over T tiles holds the tiles of the workgroups b < T
// with b % NCHAIN == c: T / NCHAIN of them, plus one for c < T % NCHAIN.
__host__ __device__ inline uint32_t chain_tiles(uint32_t T, uint32_t c) {
  return T / NCHAIN + (c < T % NCHAIN ? 1u : 0u);
}
__host__ __device__ inline uint32_t chain_first(uint32_t T, uint32_t c) {
  const uint32_t r = T % NCHAIN;
  return c * (T / NCHAIN) + (c < r ? c : r);
}
__device__ inline uint32_t chain_of_tile(uint32_t T, uint32_t t) {
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 1; k < uint32_t(NCHAIN); ++k) c += t >= chain_first(T, k) ? 1u : 0u;
  return c;
}

// Per-chain digit counts -> global starts: bins[c * R + d] = sum of all
// counts of digits below d plus the counts of digit d in chains below c.
// One workgroup of R threads (thread == digit).
template <int R>
__global__ __launch_bounds__(R) void k_chain_bins(const uint32_t* __restrict__ hist,
                                                  uint32_t* __restrict__ bins) {
  __shared__ uint32_t s_tmp[R / WAVE > 0 ? R / WAVE : 1];
  const int d = threadIdx.x;
  uint32_t tot = 0;
#pragma unroll
  for (int c = 0; c < NCHAIN; ++c) tot += hist[c * R + d];
  uint32_t all;
  uint32_t run = block_excl_scan<(R / WAVE > 0 ? R / WAVE : 1)>(tot, s_tmp, &all);
#pragma unroll
  for (int c = 0; c < NCHAIN; ++c) {
    bins[c * R + d] = run;
    run += hist[c * R + d];
  }
}

// Add one item's bin to an LDS histogram.  A wave whose valid lanes all
// hold one bin updates the counter once instead of serialising its atomics
// on one address; with PEEL (see rank_lds) also the group of lanes sharing
// the first valid lane's bin when it is at least half of the wave.  Every
// lane of the wave must call it.
template <bool PEEL = false>
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t bin, bool valid) {
  // the first and last lanes differ: no whole-wave run, skip the ballots
  // (see rank_lds)
  if constexpr (!PEEL) {
    if (__builtin_amdgcn_readfirstlane(bin) != __builtin_amdgcn_readlane(bin, WAVE - 1)) {
      if (valid) atomicAdd(&h[bin], 1u);
      return;
    }
  }
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const int first = __ffsll((unsigned long long)act) - 1;
  const uint32_t b0 = __builtin_amdgcn_readlane(bin, first);
  const uint64_t same = __ballot(valid && bin == b0);
  const uint32_t c = uint32_t(__popcll(same));
  if (same == act || (PEEL && c >= 32u)) {
    if (lane_id() == first) atomicAdd(&h[b0], c);
    if ((same >> lane_id()) & 1u) return;
  }
  if (valid) atomicAdd(&h[bin], 1u);
}

// Tiles are large (16384 int32 keys, 8192 int64 keys) so that each digit's
// run inside a tile spans whole 128-byte lines when written out.
constexpr int SORT_IPT = 16;
template <class K>
struct SortGeom {
  static constexpr int BLOCK = sizeof(K) == 4 ? 1024 : 512;
  static constexpr int WAVES = BLOCK / WAVE;
  static constexpr int TILE = BLOCK * SORT_IPT;
  // tile (word + row id), per-wave digit counters, digit starts and bases,
  // scan scratch and tile index, the next pass's per-chain counts
  static constexpr size_t LDS = size_t(TILE) * (sizeof(K) + 4) + size_t(WAVES) * 256 * 4 +
                                256 * 4 * 2 + WAVES * 4 + 16 + size_t(NCHAIN) * 256 * 4 + 256;
};

// The plain radix digit: byte `shift / 8` of the key word; `next_shift`
// is the next pass's byte (counted by the write-out when asked to).
struct RadixDigit {
  int shift, next_shift;
  static constexpr size_t EXTRA_LDS = 0;
  static constexpr bool PEEL = false;
  // PACK: the first pass keeps each item's code beside its tile-local row
  // in the LDS row slot, so the write-out does not evaluate it again (for
  // codes that cost more than a shift; RadixDigit's is the word itself)
  static constexpr bool PACK = false;
  __device__ void init(unsigned char*) const {}
  template <class W>
  __device__ W code(W w, const unsigned char*) const {
    return w;
  }
};

// ------------------------------------------------------------ first histogram
// Per-chain histogram of the digit dg selects (of the key words when FIRST,
// else of the words `win`), over tile-aligned ranges of T tiles of TILE
// items: one workgroup walks a contiguous run of tiles, so its chain
// changes at most a few times and its LDS counters are flushed once per
// chain.  With `orand` (FIRST only), also the OR and the AND of all key
// words: a digit where they agree is the same for every key.
template <class K, bool FIRST, class DG>
__global__ __launch_bounds__(256) void k_chain_hist(const K* __restrict__ keys,
                                                    const word_t<K>* __restrict__ win, uint32_t n,
                                                    uint32_t T, uint32_t tile_items, DG dg,
                                                    bool desc, uint32_t* __restrict__ hist,
                                                    word_t<K>* __restrict__ orand, bool vec) {
  using W = word_t<K>;
  extern __shared__ __align__(16) unsigned char s_raw[];
  uint32_t* s_h = reinterpret_cast<uint32_t*>(s_raw);
  unsigned char* s_dg = s_raw + 256 * 4;
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_h[i] = 0;
  dg.init(s_dg);
  __syncthreads();
  const uint32_t per = (T + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = blockIdx.x * per;
  const uint32_t t1 = t0 + per < T ? t0 + per : T;
  W o = 0, a = ~W(0);
  uint32_t cur = t0 < t1 ? chain_of_tile(T, t0) : 0u;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t c = chain_of_tile(T, t);
    if (c != cur) {  // uniform over the workgroup
      __syncthreads();
      for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        if (s_h[i]) atomicAdd(&hist[cur * 256 + i], s_h[i]);
        s_h[i] = 0;
      }
      __syncthreads();
      cur = c;
    }
    const uint32_t lo = t * tile_items;
    const uint32_t hi = lo + tile_items < n ? lo + tile_items : n;
    if constexpr (FIRST && sizeof(K) == 4) {
      if (vec) {
        // 16-byte loads of four keys: full quads of the tile (tile_items is
        // a multiple of 4, so lo / 4 is exact), then the last partial quad
        // key by key (the learned code histogram 0.33 -> 0.30 ms at 2^28)
        // (streaming loads: the radix histogram 0.276 -> 0.244 ms at 2^28;
        // eight quads per thread measured the same)
        constexpr int BQ = 4;  // quads in flight per thread
        const uint32_t q0 = lo / 4, qf = hi / 4;
        for (uint32_t b = q0; b < qf; b += BQ * blockDim.x) {
          int4 x[BQ];
#pragma unroll
          for (int u = 0; u < BQ; ++u) {
            const uint32_t qi = b + uint32_t(u) * blockDim.x + threadIdx.x;
            x[u] = qi < qf ? ld_stream4(reinterpret_cast<const int32_t*>(keys), qi)
                           : make_int4(0, 0, 0, 0);
          }
#pragma unroll
          for (int u = 0; u < BQ; ++u) {
            const bool valid = b + uint32_t(u) * blockDim.x + threadIdx.x < qf;
            const int32_t kk[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const W w = Key<K>::word(K(kk[c]), desc);
              if (valid) {
                o |= w;
                a &= w;
              }
              hist_add<DG::PEEL>(s_h, uint32_t(dg.code(w, s_dg) >> dg.shift) & 255u, valid);
            }
          }
        }
        for (uint32_t b = qf * 4; b < hi; b += blockDim.x) {
          const uint32_t i = b + threadIdx.x;
          const bool valid = i < hi;
          const W w = valid ? Key<K>::word(keys[i], desc) : W(0);
          if (valid) {
            o |= w;
            a &= w;
          }
          hist_add<DG::PEEL>(s_h, uint32_t(dg.code(w, s_dg) >> dg.shift) & 255u, valid);
        }
        continue;
      }
    }
    constexpr int B = 16;  // loads in flight per thread
    for (uint32_t b = lo; b < hi; b += B * blockDim.x) {
      W w[B];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const uint32_t i = b + uint32_t(j) * blockDim.x + threadIdx.x;
        w[j] = 0;
        if (i < hi) {
          if constexpr (FIRST)
            w[j] = W(keys[i]);
          else
            w[j] = win[i];
        }
      }
      if constexpr (FIRST) {
#pragma unroll
        for (int j = 0; j < B; ++j) w[j] = Key<K>::word(K(w[j]), desc);
      }
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const uint32_t i = b + uint32_t(j) * blockDim.x + threadIdx.x;
        const bool valid = i < hi;
        if (valid) {
          o |= w[j];
          a &= w[j];
        }
        hist_add<DG::PEEL>(s_h, uint32_t(dg.code(w[j], s_dg) >> dg.shift) & 255u, valid);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (s_h[i]) atomicAdd(&hist[cur * 256 + i], s_h[i]);
  if (orand != nullptr) {
    for (int off = 32; off > 0; off >>= 1) {
      o |= __shfl_xor(o, off);
      a &= __shfl_xor(a, off);
    }
    if (lane_id() == 0) {
      atomicOr(&orand[0], o);
      atomicAnd(&orand[1], a);
    }
  }
}

// ------------------------------------------------------------ one pass
// FIRST: read raw keys (convert to words) and generate row ids.
// LAST: write only the permutation (+ sorted keys when kout != nullptr).
// LR: rank with lane-ordered LDS atomics (rank_lds).
// NEXT: count each item's next digit (dg.next_shift) per chain of the next
// pass (same n and tiles) into next_hist[NCHAIN][256].
// DG: digit generator -- `code(w)` gives a value whose byte at dg.shift is
// the pass's digit (RadixDigit: the key word itself; the learned sort
// passes its monotone bucket code).  DG::EXTRA_LDS bytes of dynamic LDS
// after the tile buffers are DG's own, filled by `init`.
// bin_start: [NCHAIN][256] global starts (k_chain_bins); tile_ctr: NCHAIN
// counters, zero before the launch.  Grid: a multiple of NCHAIN.
template <class K, bool FIRST, bool LAST, bool LR, bool NEXT, class DG>
__global__ __launch_bounds__(SortGeom<K>::BLOCK) void k_onesweep(
    const K* __restrict__ keys, const typename Key<K>::W* __restrict__ win,
    const int32_t* __restrict__ iin, typename Key<K>::W* __restrict__ wout,
    int32_t* __restrict__ iout, K* __restrict__ kout, const uint32_t* __restrict__ bin_start,
    uint64_t* __restrict__ lookback, uint32_t* __restrict__ tile_ctr, uint32_t n, uint32_t T,
    DG dg, uint32_t* __restrict__ next_hist, bool desc, uint32_t epoch) {
  using Tk = Key<K>;
  using W = typename Tk::W;
  using G = SortGeom<K>;
  constexpr int SORT_BLOCK = G::BLOCK, SORT_WAVES = G::WAVES, SORT_TILE = G::TILE;
  // PACK (first pass only): the LDS row slot holds (code << TB) | local row
  constexpr bool PK = FIRST && DG::PACK;
  constexpr int RS = PK ? 16 : 8;  // rank shift in the rank word
  constexpr int TB = SORT_TILE == 16384 ? 14 : 13;
  static_assert((1 << TB) == SORT_TILE, "tile-local row bits");

  extern __shared__ __align__(16) unsigned char s_raw[];
  // 32-bit words: word and row id side by side (one 64-bit LDS access per
  // item in the reorder and the write-out); 64-bit words: two arrays
  constexpr bool IL = sizeof(W) == 4;
  W* s_w = reinterpret_cast<W*>(s_raw);
  uint32_t* s_i = reinterpret_cast<uint32_t*>(s_w + SORT_TILE);
  uint2* s_wi = reinterpret_cast<uint2*>(s_raw);
  uint32_t(*s_wh)[256] = reinterpret_cast<uint32_t(*)[256]>(s_i + SORT_TILE);
  uint32_t* s_tstart = &s_wh[SORT_WAVES][0];
  uint32_t* s_gbase = s_tstart + 256;
  uint32_t* s_tmp = s_gbase + 256;
  uint32_t* s_tile = s_tmp + SORT_WAVES;
  uint32_t(*s_nh)[256] = reinterpret_cast<uint32_t(*)[256]>(s_tile + 4);
  // per digit: the next-pass chain of this tile's whole run, or 0xFF when
  // the run crosses a chain boundary
  unsigned char* s_cc = reinterpret_cast<unsigned char*>(s_nh + NCHAIN);
  unsigned char* s_dg = s_raw + G::LDS;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = wave_id();
  const uint32_t chain = blockIdx.x % NCHAIN;
  const uint32_t cfirst = chain_first(T, chain);
  const uint32_t ctiles = chain_tiles(T, chain);
  const uint32_t* bins = bin_start + chain * 256;
  // next pass's chain boundaries as output positions
  uint32_t nb[NCHAIN];
#pragma unroll
  for (int k = 0; k < NCHAIN; ++k) nb[k] = chain_first(T, k) * uint32_t(SORT_TILE);
  if constexpr (NEXT)
    for (int i = tid; i < NCHAIN * 256; i += SORT_BLOCK) s_nh[i >> 8][i & 255] = 0;
  dg.init(s_dg);

  // The next tile's index and cleared counters are prepared during the
  // write-out of the current one (nothing reads s_wh or s_tile there), so
  // the barrier that ends a tile also publishes them.
  if (tid == 0) *s_tile = atomicAdd(&tile_ctr[chain], 1u);
  for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) s_wh[i >> 8][i & 255] = 0;
  __syncthreads();
  for (;;) {
    const uint32_t local = *s_tile;
    if (local >= ctiles) break;  // uniform over the workgroup
    const uint32_t tile = cfirst + local;
    const bool head = local == 0;
    const uint32_t base = tile * uint32_t(SORT_TILE);
    const uint32_t wbase = base + uint32_t(wave) * (WAVE * SORT_IPT);

    // The tile body, compiled twice: FULL tiles (all but a pass's last)
    // drop every bounds test, so their loads take one address and
    // immediate offsets and their ranks skip the validity masks.
    auto body = [&](auto full) {
      constexpr bool FULL = decltype(full)::value;

      W kw[SORT_IPT];
      uint32_t ix[SORT_IPT];
#pragma unroll
      for (int j = 0; j < SORT_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        if (FULL || pos < n) {
          if constexpr (FIRST) {
            kw[j] = W(keys[pos]);  // converted below, outside the branch
          } else {
            kw[j] = win[pos];
            ix[j] = uint32_t(iin[pos]);
          }
        } else {
          kw[j] = 0;
          ix[j] = 0;
        }
      }

      // (A conversion inside the bounds branch above made the compiler wait
      // for each key load before issuing the next.)
      if constexpr (FIRST) {
#pragma unroll
        for (int j = 0; j < SORT_IPT; ++j) kw[j] = Tk::word(K(kw[j]), desc);
      }

      // Stable in-wave ranks: item j of lane l precedes item j of lane l+1,
      // which precedes item j+1 of lane 0 (that is the position order).  LR:
      // one lane-ordered LDS atomic per item; else the 8-ballot digit match.
      // rk[j] keeps (rank << RS) | digit (a wave's rank is below 64 * SORT_IPT);
      // PK: (rank << 16) | code instead (codes below 2^16; the digit is a
      // byte of it, and the code goes to the row slot in the reorder).
      uint32_t rk[SORT_IPT];
#pragma unroll
      for (int j = 0; j < SORT_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        const bool valid = FULL || pos < n;
        const auto cd = dg.code(kw[j], s_dg);
        const uint32_t d = uint32_t(cd >> dg.shift) & 255u;
        if constexpr (LR) {
          rk[j] = (rank_lds<DG::PEEL>(s_wh[wave], d, valid) << RS) | (PK ? uint32_t(cd) : d);
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
          rk[j] = ((old + r) << RS) | (PK ? uint32_t(cd) : d);
        }
      }
      __syncthreads();

      // Per digit (threads 0..255, thread == digit): wave-exclusive offsets
      // and the tile's count, published at once for later tiles' look-back.
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
        lb_store(my_lb, lb_pack(epoch, head ? LB_PREFIX : LB_AGG, cnt));
      }
      uint32_t tile_total;
      // (no trailing barrier: s_tmp is next written in the next tile)
      const uint32_t tstart = block_excl_scan<SORT_WAVES, false>(cnt, s_tmp, &tile_total);
      if (digit_thread) s_tstart[tid] = tstart;
      __syncthreads();

      // Reorder the tile in LDS by digit (stable).
#pragma unroll
      for (int j = 0; j < SORT_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        if (FULL || pos < n) {
          const uint32_t d = PK ? ((rk[j] >> dg.shift) & 255u) : (rk[j] & 255u);
          const uint32_t p = s_tstart[d] + s_wh[wave][d] + (rk[j] >> RS);
          uint32_t xr;
          if constexpr (PK)
            xr = ((rk[j] & 0xFFFFu) << TB) | (pos - base);
          else if constexpr (FIRST)
            xr = pos;  // the first pass's rows are the positions
          else
            xr = ix[j];
          if constexpr (IL) {
            s_wi[p] = make_uint2(uint32_t(kw[j]), xr);
          } else {
            s_w[p] = kw[j];
            s_i[p] = xr;
          }
        }
      }

      // Decoupled look-back for this digit's offset among the chain's
      // earlier tiles (the chain's first tile publishes a prefix at once).
      if (digit_thread) {
        uint32_t excl = 0;
        if (!head) {
          excl = lb_lookback(lookback + tid, tile, 256, epoch);
          lb_store(my_lb, lb_pack(epoch, LB_PREFIX, excl + cnt));
        }
        s_gbase[tid] = bins[tid] + excl - tstart;
        if constexpr (NEXT) {
          // chain of the run's first and last output position (monotone in
          // the position, so equal ends mean one chain for the whole run)
          const uint32_t g0 = bins[tid] + excl;
          const uint32_t g1 = cnt ? g0 + cnt - 1u : g0;
          uint32_t c0 = 0, c1 = 0;
#pragma unroll
          for (int k = 1; k < NCHAIN; ++k) {
            c0 += g0 >= nb[k] ? 1u : 0u;
            c1 += g1 >= nb[k] ? 1u : 0u;
          }
          s_cc[tid] = (unsigned char)(c0 == c1 ? c0 : 0xFFu);
        }
      }
      __syncthreads();

      const uint32_t nvalid =
          FULL ? uint32_t(SORT_TILE)
               : ((n - base) < uint32_t(SORT_TILE) ? (n - base) : uint32_t(SORT_TILE));
      // WU items per thread per step, their LDS reads issued together so one
      // wait covers them (radix passes 1.05 -> 1.02 ms at 2^28 with two; the
      // learned passes, whose code map reads LDS knots, measured slower)
      constexpr int WU = DG::PEEL ? 1 : 2;
      if (tid == 0) *s_tile = atomicAdd(&tile_ctr[chain], 1u);
      for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) s_wh[i >> 8][i & 255] = 0;
      for (uint32_t i0 = 0; i0 < nvalid; i0 += WU * SORT_BLOCK) {  // uniform trip count
        W w[WU];
        uint32_t xi[WU], gb[WU], cc[WU], pc[WU];
        bool ok[WU];
#pragma unroll
        for (int u = 0; u < WU; ++u) {
          const uint32_t i = i0 + uint32_t(u * SORT_BLOCK + tid);
          ok[u] = FULL || i < nvalid;
          const uint32_t ic = ok[u] ? i : 0u;
          if constexpr (IL) {
            const uint2 e = s_wi[ic];
            w[u] = W(e.x);
            xi[u] = e.y;
          } else {
            w[u] = s_w[ic];
            xi[u] = s_i[ic];
          }
        }
#pragma unroll
        for (int u = 0; u < WU; ++u) {
          if (!ok[u]) w[u] = W(0);
          uint32_t d;
          if constexpr (PK) {
            pc[u] = xi[u] >> TB;
            xi[u] = base + (xi[u] & uint32_t(SORT_TILE - 1));
            d = (pc[u] >> dg.shift) & 255u;
          } else {
            d = uint32_t(dg.code(w[u], s_dg) >> dg.shift) & 255u;
          }
          gb[u] = s_gbase[d];
          if constexpr (NEXT) cc[u] = s_cc[d];
        }
#pragma unroll
        for (int u = 0; u < WU; ++u) {
          const uint32_t i = i0 + uint32_t(u * SORT_BLOCK + tid);
          const uint32_t g = gb[u] + i;
          if (ok[u]) {
            if constexpr (LAST) {
              iout[g] = int32_t(xi[u]);
              if (kout != nullptr) kout[g] = Tk::key(w[u], desc);
            } else {
              wout[g] = w[u];
              iout[g] = int32_t(xi[u]);
            }
          }
          if constexpr (NEXT) {
            // the run's chain from LDS; item by item only in the few runs
            // that cross a chain boundary
            uint32_t c = cc[u];
            if (c == 0xFFu) {
              c = 0;
#pragma unroll
              for (int k = 1; k < NCHAIN; ++k) c += g >= nb[k] ? 1u : 0u;
            }
            uint32_t nd;
            if constexpr (PK)
              nd = (pc[u] >> dg.next_shift) & 255u;
            else
              nd = uint32_t(dg.code(w[u], s_dg) >> dg.next_shift) & 255u;
            hist_add<DG::PEEL>(&s_nh[0][0], c * 256u + nd, ok[u]);
          }
        }
      }
    };
    // (the first and counting passes of digit generators with a costly
    // code keep one body: a second copy pushed the learned first pass over
    // its register budget; the learned second pass takes both: the learned
    // sort of 2^28 uniform keys 3.48 -> 3.40 ms)
    if (!(DG::PACK && (FIRST || NEXT)) && base + uint32_t(SORT_TILE) <= n)
      body(std::true_type{});
    else
      body(std::false_type{});
    __syncthreads();  // the tile buffers are reused by the next tile
  }
  if constexpr (NEXT) {
    for (int i = tid; i < NCHAIN * 256; i += SORT_BLOCK) {
      const uint32_t c = s_nh[i >> 8][i & 255];
      if (c) atomicAdd(&next_hist[i], c);
    }
  }
}

// One pass over n items in T tiles.  `bins` = [NCHAIN][256] starts of this
// pass; with next_hist != nullptr the pass also counts dg.next_shift's
// digit per chain of the next pass into next_hist (zeroed by the caller).
// ctr: NCHAIN zeroed tile counters of this pass.
template <class K, bool FIRST, bool LAST, class DG>
void onesweep_pass(const K* keys, const word_t<K>* win, const int32_t* iin, word_t<K>* wout,
                   int32_t* iout, K* kout, const uint32_t* bins, uint64_t* lb, uint32_t* ctr,
                   uint32_t n, const DG& dg, uint32_t* next_hist, bool desc, hipStream_t s,
                   const char* prof_name = "radix_onesweep", uint32_t epoch_in = 0) {
  using G = SortGeom<K>;
  constexpr size_t LDS = G::LDS + DG::EXTRA_LDS;
  static_assert(LDS <= 160 * 1024, "LDS budget of the pass kernel");
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    const void* fns[4] = {
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, true, true, DG>),
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, false, true, DG>),
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, true, false, DG>),
        reinterpret_cast<const void*>(&k_onesweep<K, FIRST, LAST, false, false, DG>)};
    for (const void* f : fns)
      NFH_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(LDS)));
  });
  const bool lr = lds_rank_supported();
  const bool next = next_hist != nullptr;
  const uint32_t T = (n + G::TILE - 1) / G::TILE;
  // epoch_in != 0: the caller's epoch, newer than every word in lb
  const uint32_t epoch = epoch_in != 0 ? epoch_in : lookback_epoch(s);
  // persistent: one workgroup per CU (the tile buffers take most of the
  // LDS), NCHAIN-aligned so every chain has the same number of workers
  const unsigned grid =
      unsigned(std::max<int>(NCHAIN, num_cus() / NCHAIN * NCHAIN));
  // pass model: first pass reads keys, later ones (word, row id); the last
  // writes row ids (+ keys for unique), the others (word, row id)
  const double in_b = FIRST ? sizeof(K) : sizeof(typename Key<K>::W) + 4.0;
  const double out_b = LAST ? 4.0 + (kout ? sizeof(K) : 0.0) : sizeof(typename Key<K>::W) + 4.0;
  ProfScope prof(prof_name, s, double(n) * (in_b + out_b));
#define NFH_OS_LAUNCH(LRV, NXV)                                                                \
  hipLaunchKernelGGL((k_onesweep<K, FIRST, LAST, LRV, NXV, DG>), dim3(grid), dim3(G::BLOCK), \
                     LDS, s, keys, win, iin, wout, iout, kout, bins, lb, ctr, n, T, dg,         \
                     next_hist, desc, epoch)
  if (lr && next)
    NFH_OS_LAUNCH(true, true);
  else if (lr)
    NFH_OS_LAUNCH(true, false);
  else if (next)
    NFH_OS_LAUNCH(false, true);
  else
    NFH_OS_LAUNCH(false, false);
#undef NFH_OS_LAUNCH
  NFH_LAUNCH_CHECK();
}

// Per-chain histogram of the first pass's digit (and the OR / AND of the
// key words when orand != nullptr) over the keys.
template <class K, class DG>
void chain_hist_first(const K* keys, uint32_t n, const DG& dg, bool desc, uint32_t* hist,
                      word_t<K>* orand, hipStream_t s, const char* prof_name) {
  using G = SortGeom<K>;
  const uint32_t T = (n + G::TILE - 1) / G::TILE;
  const unsigned grid = std::max(1u, std::min<unsigned>(T, 4u * unsigned(num_cus())));
  const size_t lds = 256 * 4 + DG::EXTRA_LDS;
  ProfScope prof(prof_name, s, double(n) * sizeof(K));
  const bool vec = sizeof(K) == 4 && (reinterpret_cast<uintptr_t>(keys) % 16) == 0;
  hipLaunchKernelGGL((k_chain_hist<K, true, DG>), dim3(grid), dim3(256), lds, s, keys,
                     static_cast<const word_t<K>*>(nullptr), n, T, uint32_t(G::TILE), dg, desc,
                     hist, orand, vec);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh
