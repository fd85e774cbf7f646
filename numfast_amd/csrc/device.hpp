// device.hpp -- wave64 / workgroup primitives shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nfh {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }
__device__ __forceinline__ uint64_t lanemask_lt() {
  return (uint64_t(1) << lane_id()) - 1;
}

// Tile of workgroup b in a grid of t workgroups, one per tile.  MI355X
// dispatches workgroup b to XCD b % 8, each XCD with its own L2; here XCD
// x takes the contiguous tiles [x*(t/8) + min(x, t%8), ...) in order, so
// neighbouring tiles, whose output runs meet inside 128-byte lines, merge
// those lines in one L2 instead of writing partial lines from two.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t t) {
  const uint32_t x = b & 7u, base = t >> 3, rem = t & 7u;
  return x * base + (x < rem ? x : rem) + (b >> 3);
}

// Lanes of this wave whose 8-bit digit equals mine (restricted to `valid`
// lanes).  Eight ballots, one per digit bit; every lane must execute it.
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

// Stable in-wave rank of digit d through a wave-private LDS counter array
// `ctr` (one counter per digit), returning the count of this wave's earlier
// items with the same digit.  Calls for successive items of a wave see each
// other's increments (LDS operations of a wave complete in issue order);
// within one call, lanes that hit the same counter receive ds_add_rtn's old
// values in ascending lane order -- a property of the LDS that
// lds_rank_supported() verifies once per device (the ballot-match ranking
// above is used when it does not hold).  Invalid lanes must be the highest
// lanes of the wave (a tail).  When every valid lane has the same digit the
// wave updates the counter once instead of serialising on one address.
// PEEL (for digits that long runs of rows share, such as the learned
// sort's bucket code of nearly sorted keys): also when at least half of the
// valid lanes share the first lane's digit, that group is ranked by one
// counter update and a popcount and only the other lanes use the atomic.
// It costs radix digits about 1.5% per pass, and peeling every shared
// digit off in a loop measured slower still.
template <bool PEEL = false>
__device__ __forceinline__ uint32_t rank_lds(uint32_t* ctr, uint32_t d, bool valid) {
  // Cheap guard (two lane reads, a scalar compare): when the first and the
  // last lane hold different digits the wave cannot be a single run, so
  // the ballots are skipped (radix passes 1.02 -> 0.97 ms on average at
  // 2^28 uniform int32 keys, with the same guard in hist_add).  With PEEL
  // the ballots always run: nearly sorted codes measured slower with it.
  if constexpr (!PEEL) {
    if (__builtin_amdgcn_readfirstlane(d) != __builtin_amdgcn_readlane(d, WAVE - 1))
      return valid ? atomicAdd(&ctr[d], 1u) : 0u;
  }
  const uint64_t act = __ballot(valid);
  const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
  const uint64_t same = __ballot(valid && d == d0);
  const uint32_t c = uint32_t(__popcll(same));
  if (same == act || (PEEL && c >= 32u)) {
    uint32_t base = 0;
    if (lane_id() == 0) {
      base = ctr[d0];
      ctr[d0] = base + c;
    }
    base = __builtin_amdgcn_readfirstlane(base);
    if ((same >> lane_id()) & 1u) return base + uint32_t(__popcll(same & lanemask_lt()));
  }
  return valid ? atomicAdd(&ctr[d], 1u) : 0u;
}

// Inclusive scan of a 32-bit value across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint32_t t = __shfl_up(x, off);
    if (lane >= off) x += t;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint64_t t = __shfl_up(x, off);
    if (lane >= off) x += t;
  }
  return x;
}

// Exclusive scan across a workgroup of NW waves; `s_tmp` holds NW words.
// Every thread must call it.  Returns the exclusive prefix, sets *total.
// TRAIL = false leaves out the closing barrier, for callers that pass
// another barrier before s_tmp is written again.
template <int NW, bool TRAIL = true>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp,
                                                    uint32_t* total) {
  const uint32_t inc = wave_incl_scan(v);
  if (lane_id() == WAVE - 1) s_tmp[wave_id()] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t c = s_tmp[w];
    if (w < wave_id()) off += c;
    tot += c;
  }
  if constexpr (TRAIL) __syncthreads();
  *total = tot;
  return off + inc - v;
}
template <int NW>
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* s_tmp,
                                                      uint64_t* total) {
  const uint64_t inc = wave_incl_scan64(v);
  if (lane_id() == WAVE - 1) s_tmp[wave_id()] = inc;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint64_t c = s_tmp[w];
    if (w < wave_id()) off += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// ---------------------------------------------------------------- look-back
// A look-back word: [63:34] epoch (30 bits), [33:32] status, [31:0] value.
// The epoch tags each pass/call so stale words from earlier passes read as
// "not ready" and the buffer never needs clearing.
constexpr uint64_t LB_AGG = 1;
constexpr uint64_t LB_PREFIX = 2;

__device__ __forceinline__ uint64_t lb_pack(uint32_t epoch, uint64_t status, uint32_t v) {
  return (uint64_t(epoch) << 34) | (status << 32) | uint64_t(v);
}
__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Walk back from tile `tile-1` summing aggregates until a prefix is found.
// `stride` is the distance between consecutive tiles' words.
// The first probe reads one word (usually the predecessor's prefix); once
// a run of aggregates shows up, eight words are requested back to back per
// round, so a long walk costs one memory round trip per eight tiles.
__device__ __forceinline__ uint32_t lb_lookback(const uint64_t* lb, uint32_t tile,
                                                uint32_t stride, uint32_t epoch) {
  constexpr int K = 8;
  uint32_t excl = 0;
  int64_t j = int64_t(tile) - 1;
  int width = 1;
  while (j >= 0) {
    const int cnt = j + 1 < width ? int(j + 1) : width;
    uint64_t w[K];
#pragma unroll
    for (int k = 0; k < K; ++k) w[k] = k < cnt ? lb_load(lb + size_t(j - k) * stride) : 0ull;
    int k = 0;
    for (; k < cnt; ++k) {
      const uint32_t ep = uint32_t(w[k] >> 34);
      const uint32_t st = uint32_t(w[k] >> 32) & 3u;
      if (ep != epoch || st == 0) break;  // not published yet: poll again
      excl += uint32_t(w[k]);
      if (st == LB_PREFIX) return excl;
    }
    j -= k;
    if (k < cnt)
      __builtin_amdgcn_s_sleep(1);
    else
      width = K;
  }
  return excl;
}

// Streaming (non-temporal) quad loads and stores: rows read or written
// once bypass the L2 so that randomly accessed tables stay resident.
typedef int32_t nf_v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int4 ld_stream4(const int32_t* p, uint32_t q) {
  const nf_v4i x = __builtin_nontemporal_load(reinterpret_cast<const nf_v4i*>(p) + q);
  return make_int4(x.x, x.y, x.z, x.w);
}

__device__ __forceinline__ void st_stream4(int32_t* p, uint32_t q, int4 v) {
  nf_v4i x;
  x.x = v.x;
  x.y = v.y;
  x.z = v.z;
  x.w = v.w;
  __builtin_nontemporal_store(x, reinterpret_cast<nf_v4i*>(p) + q);
}

}  // namespace nfh
