// groupagg.hip -- NULL-aware group-by by raw int32 keys (sum / count / min /
// max / mean) and unique with validity.  Additive entry points: the
// reference ABI has no grouped min/max, NULL or mean kernel; its executor
// does these in NumPy (src/Drivers/CPU/_lib/cpu.py:3507-3563 compaction +
// np.minimum/maximum.at + bincount, carry.py:55-62 mean = sum / count).
//
// Result contract (ColumnCarry, carry.py:21-62; groupindex.py:3-8): groups
// are the distinct keys of the rows whose key and value are both valid,
// ascending; counts int64; sums int64 (exact); min/max int32; mean float64
// = double(sum) / double(count).
//
// Strategies (identical results; the rule picks partitioned, then dense,
// then sort-based, and nfh_set_groupby_strategy can force one):
//  * partitioned (span <= 2^24, no key mask): gbpart.hip;
//  * dense (valid key span moderate): one accumulate pass into span-indexed
//    32-byte slots {sum, count, min, max}.  Hot keys -- found by sampling
//    2^16 rows, sorting the sample and keeping the keys seen most often --
//    are privatised per workgroup in LDS (hash table + LDS atomics) and
//    flushed once per workgroup; every other row goes straight to its
//    global slot (the high-cardinality spill: two atomics, min/max only
//    when the row would change them).  One scan over the slots compacts the
//    present groups in key order.
//  * sparse (huge span): the valid rows are compacted, radix-unique'd
//    (unique.hip) and aggregated by code; ukeys = the unique values;
//  * hash-scatter (any span, few distinct keys; only when forced or
//    selected): per-workgroup LDS hash tables flushed into one global
//    hash table, then the present keys sorted (gbhash.hip).
#include <algorithm>

#include "device.hpp"
#include "gbhash.hpp"
#include "gbpart.hpp"
#include "hot.hpp"
#include "groupagg.hpp"
#include "runtime.hpp"
#include "scan.hpp"
#include "sort.hpp"
#include "unique.hpp"

namespace nfh {

typedef unsigned long long u64;

using GroupSlot = GpSlot;

constexpr uint32_t HOT_MIN_COUNT = 4;  // sampled occurrences to qualify
constexpr int HOT_LEN_BUCKETS = 1024;
constexpr int ACC_BLOCK = 256;


// Rows are visited in quads (128-bit key/value loads, 32-bit mask loads)
// when every pointer is suitably aligned, else one by one.  Each thread
// issues the loads of R quads (RQ unless the caller asks for more) before
// using any of them, so enough bytes are in flight per CU to cover HBM
// latency.
constexpr int RQ = 4;

template <bool VEC, int R = RQ, class F>
__device__ __forceinline__ void for_each_row(const int32_t* __restrict__ keys,
                                             const int32_t* __restrict__ vals,
                                             const uint8_t* __restrict__ vm,
                                             const uint8_t* __restrict__ km, uint32_t n, F&& f) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t tail = 0;
  if constexpr (VEC) {
    const uint32_t nq = n / 4;
    uint32_t q = tid;
    for (; q + (R - 1) * stride < nq; q += R * stride) {
      int4 k4[R], v4[R];
      uint32_t a[R], b[R];
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const uint32_t qq = q + uint32_t(u) * stride;
        k4[u] = reinterpret_cast<const int4*>(keys)[qq];
        v4[u] = vals ? reinterpret_cast<const int4*>(vals)[qq] : make_int4(0, 0, 0, 0);
        a[u] = vm ? reinterpret_cast<const uint32_t*>(vm)[qq] : 0x01010101u;
        b[u] = km ? reinterpret_cast<const uint32_t*>(km)[qq] : 0x01010101u;
      }
#pragma unroll
      for (int u = 0; u < R; ++u) {
        f(k4[u].x, v4[u].x, (a[u] & 0xFFu) && (b[u] & 0xFFu));
        f(k4[u].y, v4[u].y, (a[u] & 0xFF00u) && (b[u] & 0xFF00u));
        f(k4[u].z, v4[u].z, (a[u] & 0xFF0000u) && (b[u] & 0xFF0000u));
        f(k4[u].w, v4[u].w, (a[u] & 0xFF000000u) && (b[u] & 0xFF000000u));
      }
    }
    for (; q < nq; q += stride) {
      const int4 k4 = reinterpret_cast<const int4*>(keys)[q];
      const int4 v4 = vals ? reinterpret_cast<const int4*>(vals)[q] : make_int4(0, 0, 0, 0);
      const uint32_t a = vm ? reinterpret_cast<const uint32_t*>(vm)[q] : 0x01010101u;
      const uint32_t b = km ? reinterpret_cast<const uint32_t*>(km)[q] : 0x01010101u;
      f(k4.x, v4.x, (a & 0xFFu) && (b & 0xFFu));
      f(k4.y, v4.y, (a & 0xFF00u) && (b & 0xFF00u));
      f(k4.z, v4.z, (a & 0xFF0000u) && (b & 0xFF0000u));
      f(k4.w, v4.w, (a & 0xFF000000u) && (b & 0xFF000000u));
    }
    tail = nq * 4;
  }
  for (uint32_t i = tail + tid; i < n; i += stride) {
    const bool ok = (vm == nullptr || vm[i] != 0) && (km == nullptr || km[i] != 0);
    f(keys[i], vals ? vals[i] : 0, ok);
  }
}

static bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }
static bool vec_ok(const int32_t* k, const int32_t* v, const uint8_t* m1, const uint8_t* m2) {
  return aligned(k, 16) && (v == nullptr || aligned(v, 16)) && (m1 == nullptr || aligned(m1, 4)) &&
         (m2 == nullptr || aligned(m2, 4));
}

// min / max of the valid keys.  Each workgroup reads one contiguous range
// of quads, sixteen in flight per thread, with streaming loads; then one
// min / max per workgroup reaches global memory, each atomic issued only
// when it would move the value.  (At 2^28 rows: a grid-stride form with an
// atomic pair per wave took 0.30 ms -- 8192 waves' atomics on two words --
// and 0.44 ms at twice the grid; this form 0.24 ms, 5.6 TB/s.)
template <bool VEC>
__global__ __launch_bounds__(256) void k_key_range(const int32_t* __restrict__ keys,
                                                   const uint8_t* __restrict__ v,
                                                   const uint8_t* __restrict__ kv, uint32_t n,
                                                   int* __restrict__ out /* min, max */) {
  __shared__ int s_mn[256 / WAVE], s_mx[256 / WAVE];
  int mn = INT32_MAX, mx = INT32_MIN;
  auto f = [&](int32_t k, bool ok) {
    if (ok) {
      mn = k < mn ? k : mn;
      mx = k > mx ? k : mx;
    }
  };
  uint32_t tail = 0;
  if constexpr (VEC) {
    constexpr int R = 16;
    const uint32_t nq = n / 4;
    const uint32_t per = (nq + gridDim.x - 1) / gridDim.x;
    const uint32_t q0 = blockIdx.x * per;
    const uint32_t q1 = q0 + per < nq ? q0 + per : nq;
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += R * blockDim.x) {
      int4 k4[R];
      uint32_t a[R], b[R];
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const uint32_t qq = q + uint32_t(u) * blockDim.x;
        const uint32_t qc = qq < q1 ? qq : q0;  // in range; masked out below
        k4[u] = ld_stream4(keys, qc);
        a[u] = v ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(v) + qc)
                 : 0x01010101u;
        b[u] = kv ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(kv) + qc)
                  : 0x01010101u;
        if (qq >= q1) a[u] = 0u;
      }
#pragma unroll
      for (int u = 0; u < R; ++u) {
        f(k4[u].x, (a[u] & 0xFFu) && (b[u] & 0xFFu));
        f(k4[u].y, (a[u] & 0xFF00u) && (b[u] & 0xFF00u));
        f(k4[u].z, (a[u] & 0xFF0000u) && (b[u] & 0xFF0000u));
        f(k4[u].w, (a[u] & 0xFF000000u) && (b[u] & 0xFF000000u));
      }
    }
    tail = nq * 4;
  }
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = tail + tid; i < n; i += stride)
    f(keys[i], (v == nullptr || v[i] != 0) && (kv == nullptr || kv[i] != 0));
  for (int off = 32; off > 0; off >>= 1) {
    const int a = __shfl_down(mn, off), b = __shfl_down(mx, off);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if (lane_id() == 0) {
    s_mn[wave_id()] = mn;
    s_mx[wave_id()] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < int(blockDim.x / WAVE); ++w) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    // both only move one way: a stale read costs a redundant atomic only
    if (mn != INT32_MAX && mn < __hip_atomic_load(&out[0], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(&out[0], mn);
    if (mx != INT32_MIN && mx > __hip_atomic_load(&out[1], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&out[1], mx);
  }
}

__global__ void k_slot_init(GroupSlot* __restrict__ s, u64 m) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < m; i += stride) {
    GroupSlot z;
    z.sum = 0;
    z.cnt = 0;
    z.mn = INT32_MAX;
    z.mx = INT32_MIN;
    z.pad[0] = z.pad[1] = z.pad[2] = 0;
    s[i] = z;
  }
}

// ---------------------------------------------------------------- hot keys
__global__ void k_sample(const int32_t* __restrict__ keys, uint32_t n, uint32_t S,
                         int32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S) out[i] = keys[uint32_t((u64(i) * n) / S)];
}

// Run length at each run head of the sorted sample (0 elsewhere).
// Hot-key counting: the sample is split over HOT_PARTS workgroups by a
// hash of the key; each counts its part in an LDS table of 64-bit words
// {key : 32 | count : 32} (count 0 = empty; claimed by compare-and-swap,
// counted by atomic add, no waiting on other lanes) and appends the keys
// seen at least HOT_MIN_COUNT times to the candidate list.
constexpr int HOT_PARTS = 64;
constexpr int HOT_PART_SLOTS = 2048;  // a part holds about S / 64 sampled keys

__global__ __launch_bounds__(1024) void k_hot_count(const int32_t* __restrict__ sample, uint32_t S,
                                                    int2* __restrict__ cand,
                                                    uint32_t* __restrict__ ncand) {
  __shared__ unsigned long long t[HOT_PART_SLOTS];
  for (int i = threadIdx.x; i < HOT_PART_SLOTS; i += blockDim.x) t[i] = 0ull;
  __syncthreads();
  // eight sample loads in flight per thread (one at a time, the pass was
  // a chain of 64 dependent loads per thread)
  constexpr uint32_t U = 8;
  for (uint32_t i0 = threadIdx.x; i0 < S; i0 += U * blockDim.x) {
    int32_t ks[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * blockDim.x;
      ks[u] = i < S ? sample[i] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (i0 + u * blockDim.x >= S) continue;
      const int32_t k = ks[u];
      const uint32_t h = hot_mix(k);
      if ((h >> 26) != blockIdx.x) continue;
      const unsigned long long mine = ((unsigned long long)uint32_t(k) << 32) | 1ull;
      uint32_t slot = h & (HOT_PART_SLOTS - 1);
      for (int probe = 0; probe < HOT_PART_SLOTS; ++probe) {
        const unsigned long long old = atomicCAS(&t[slot], 0ull, mine);
        if (old == 0ull) break;                          // claimed, count 1
        if (uint32_t(old >> 32) == uint32_t(k)) {       // counted
          atomicAdd(&t[slot], 1ull);
          break;
        }
        slot = (slot + 1) & (HOT_PART_SLOTS - 1);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < HOT_PART_SLOTS; i += blockDim.x) {
    const unsigned long long e = t[i];
    const uint32_t c = uint32_t(e);
    if (c >= HOT_MIN_COUNT) cand[atomicAdd(ncand, 1u)] = make_int2(int32_t(e >> 32), int32_t(c));
  }
}

// One block: keep the candidates counted most often (at most HOT_MAX) and
// build their cuckoo table.  hot_n[0] = hot keys, hot_n[1] = sampled rows
// that carry one of them (of HOT_SAMPLE).
__global__ __launch_bounds__(1024) void k_hot_pick(const int2* __restrict__ cand,
                                                   const uint32_t* __restrict__ ncand,
                                                   int32_t* __restrict__ hot_keys,
                                                   int2* __restrict__ table,
                                                   uint32_t* __restrict__ hot_n) {
  __shared__ uint32_t hist[HOT_LEN_BUCKETS];
  __shared__ int32_t s_hot[HOT_MAX];
  __shared__ unsigned long long cell[HOT_TBL];
  __shared__ int32_t t_idx[HOT_TBL];
  __shared__ int32_t t_key[HOT_TBL];
  __shared__ uint32_t s_thr, s_cnt, s_cov;
  __shared__ uint32_t s_tmp[1024 / WAVE];
  static_assert(HOT_LEN_BUCKETS == 1024, "one count bucket per thread below");
  const uint32_t nc = *ncand;
  for (int b = threadIdx.x; b < HOT_LEN_BUCKETS; b += blockDim.x) hist[b] = 0;
  if (threadIdx.x == 0) {
    s_thr = 0xFFFFFFFFu;
    s_cnt = 0;
    s_cov = 0;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
    const uint32_t L = uint32_t(cand[i].y);
    atomicAdd(&hist[L < HOT_LEN_BUCKETS ? L : HOT_LEN_BUCKETS - 1], 1u);
  }
  __syncthreads();
  // threshold: the smallest count bucket b >= HOT_MIN_COUNT whose suffix
  // (candidates counted b times or more) holds at most HOT_MAX keys --
  // thread t takes bucket 1023 - t, so an inclusive scan is the suffix sum
  {
    const uint32_t b = uint32_t(HOT_LEN_BUCKETS - 1) - threadIdx.x;
    uint32_t tot;
    const uint32_t suf = block_excl_scan<1024 / WAVE>(hist[b], s_tmp, &tot) + hist[b];
    if (b >= HOT_MIN_COUNT && suf <= uint32_t(HOT_MAX)) atomicMin(&s_thr, b);
  }
  __syncthreads();
  const uint32_t thr = s_thr;
  for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
    const int2 c = cand[i];
    const uint32_t L = uint32_t(c.y);
    const uint32_t Lb = L < HOT_LEN_BUCKETS ? L : HOT_LEN_BUCKETS - 1;
    if (Lb >= thr) {
      const uint32_t idx = atomicAdd(&s_cnt, 1u);
      hot_keys[idx] = c.x;
      s_hot[idx] = c.x;
      atomicAdd(&s_cov, L);
    }
  }
  __syncthreads();
  const uint32_t nh = s_cnt;
  hot_build(s_hot, nh, cell, t_key, t_idx);
  for (int b = threadIdx.x; b < HOT_TBL; b += blockDim.x) table[b] = make_int2(t_key[b], t_idx[b]);
  if (threadIdx.x == 0) {
    hot_n[0] = nh;
    hot_n[1] = s_cov;
  }
}

__device__ __forceinline__ void slot_add(GroupSlot* p, int32_t x) {
  atomicAdd(&p->sum, (u64)(long long)x);
  atomicAdd(&p->cnt, 1u);
  // min / max only move one way, so a stale read can only cause a
  // redundant atomic, never a missed one.
  if (x < p->mn) atomicMin(&p->mn, x);
  if (x > p->mx) atomicMax(&p->mx, x);
}

template <bool VEC>
__global__ __launch_bounds__(ACC_BLOCK) void k_group_acc(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const uint8_t* __restrict__ vm, const uint8_t* __restrict__ km, uint32_t n, int32_t kmin,
    GroupSlot* __restrict__ slots, const int32_t* __restrict__ code,
    const int2* __restrict__ table, const int32_t* __restrict__ hot_keys,
    const uint32_t* __restrict__ hot_n) {
  __shared__ int32_t t_key[HOT_TBL];
  __shared__ int32_t t_idx[HOT_TBL];
  __shared__ u64 a_sum[HOT_MAX];
  __shared__ uint32_t a_cnt[HOT_MAX];
  __shared__ int32_t a_mn[HOT_MAX];
  __shared__ int32_t a_mx[HOT_MAX];
  const uint32_t nh = table ? *hot_n : 0u;
  hot_load(table, t_key, t_idx);
  for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) {
    a_sum[b] = 0;
    a_cnt[b] = 0;
    a_mn[b] = INT32_MAX;
    a_mx[b] = INT32_MIN;
  }
  __syncthreads();
  auto row = [&](int32_t k, int32_t x, bool ok, uint32_t s) {
    if (!ok) return;
    if (nh != 0) {
      const int idx = hot_index(t_key, t_idx, k);
      if (idx >= 0) {
        atomicAdd(&a_sum[idx], (u64)(long long)x);
        atomicAdd(&a_cnt[idx], 1u);
        if (x < a_mn[idx]) atomicMin(&a_mn[idx], x);
        if (x > a_mx[idx]) atomicMax(&a_mx[idx], x);
        return;
      }
    }
    slot_add(slots + s, x);
  };
  if (code == nullptr) {
    for_each_row<VEC>(keys, vals, vm, km, n, [&](int32_t k, int32_t x, bool ok) {
      row(k, x, ok, uint32_t(k - kmin));
    });
  } else {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
      row(keys[i], vals[i], true, uint32_t(code[i]));
  }
  if (nh == 0) return;
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x) {
    const uint32_t c = a_cnt[b];
    if (c == 0) continue;
    GroupSlot* p = slots + uint32_t(hot_keys[b] - kmin);
    atomicAdd(&p->sum, a_sum[b]);
    atomicAdd(&p->cnt, c);
    atomicMin(&p->mn, a_mn[b]);
    atomicMax(&p->mx, a_mx[b]);
  }
}

struct SlotCompactOp {
  const GroupSlot* slots;
  const int32_t* ukey_of;  // sparse path: unique value per code (else null)
  int32_t kmin;
  GroupAggOut o;
  int64_t* ng;
  __device__ uint32_t value(uint32_t s) const { return slots[s].cnt != 0 ? 1u : 0u; }
  __device__ void emit(uint32_t s, uint32_t ex) const {
    const GroupSlot sl = slots[s];
    if (sl.cnt == 0) return;
    const int64_t sum = (int64_t)sl.sum;
    if (o.ukeys) o.ukeys[ex] = ukey_of ? int64_t(ukey_of[s]) : int64_t(kmin) + int64_t(s);
    if (o.counts) o.counts[ex] = int64_t(sl.cnt);
    if (o.sums) o.sums[ex] = sum;
    if (o.mins) o.mins[ex] = sl.mn;
    if (o.maxs) o.maxs[ex] = sl.mx;
    if (o.means) o.means[ex] = double(sum) / double(sl.cnt);
  }
  __device__ void total(uint32_t t) const { *ng = int64_t(t); }
};

// Compaction of the present slots in key order with coalesced output: each
// wave takes 64 consecutive slots per step (striped), ballots the present
// ones and writes them to consecutive positions; tiles of 2048 slots chain
// their counts with a decoupled look-back.  Each slot is read once, whole,
// and kept in registers for the write (reading the counts first and the
// present slots again fetched the slot array about twice: 0.33 -> 0.26 ms
// at 2^24 slots).
constexpr int SC2_BLOCK = 256, SC2_IPT = 8, SC2_TILE = SC2_BLOCK * SC2_IPT;

__global__ __launch_bounds__(SC2_BLOCK) void k_slot_compact(
    const GroupSlot* __restrict__ slots, uint64_t m, const int32_t* __restrict__ ukey_of,
    int32_t kmin, GroupAggOut o, uint64_t* __restrict__ lookback, uint32_t* __restrict__ tile_ctr,
    uint32_t epoch, int64_t* __restrict__ ng) {
  __shared__ uint32_t s_wtot[SC2_BLOCK / WAVE];
  __shared__ uint32_t s_tile, s_prefix;
  const int lane = lane_id(), wave = wave_id();
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t wbase = uint64_t(tile) * SC2_TILE + uint64_t(wave) * (WAVE * SC2_IPT);
  uint64_t bal[SC2_IPT];
  uint32_t wtot = 0;
  GroupSlot sv[SC2_IPT];
#pragma unroll
  for (int j = 0; j < SC2_IPT; ++j) {
    const uint64_t s = wbase + uint64_t(j * WAVE + lane);
    sv[j] = slots[s < m ? s : m - 1];
  }
#pragma unroll
  for (int j = 0; j < SC2_IPT; ++j) {
    const uint64_t s = wbase + uint64_t(j * WAVE + lane);
    const bool present = sv[j].cnt != 0 && s < m;
    bal[j] = __ballot(present);
    wtot += uint32_t(__popcll(bal[j]));
  }
  if (lane == 0) s_wtot[wave] = wtot;
  __syncthreads();
  uint32_t woff = 0, ttot = 0;
#pragma unroll
  for (int w = 0; w < SC2_BLOCK / WAVE; ++w) {
    if (w < wave) woff += s_wtot[w];
    ttot += s_wtot[w];
  }
  if (threadIdx.x == 0) {
    uint64_t* my = lookback + tile;
    if (tile == 0) {
      lb_store(my, lb_pack(epoch, LB_PREFIX, ttot));
      s_prefix = 0;
    } else {
      lb_store(my, lb_pack(epoch, LB_AGG, ttot));
      const uint32_t ex = lb_lookback(lookback, tile, 1, epoch);
      lb_store(my, lb_pack(epoch, LB_PREFIX, ex + ttot));
      s_prefix = ex;
    }
  }
  __syncthreads();
  uint32_t pos = s_prefix + woff;
#pragma unroll
  for (int j = 0; j < SC2_IPT; ++j) {
    const uint64_t s = wbase + uint64_t(j * WAVE + lane);
    if ((bal[j] >> lane) & 1ull) {
      const uint32_t ex = pos + uint32_t(__popcll(bal[j] & lanemask_lt()));
      const GroupSlot sl = sv[j];
      const int64_t sum = (int64_t)sl.sum;
      if (o.ukeys) o.ukeys[ex] = ukey_of ? int64_t(ukey_of[s]) : int64_t(kmin) + int64_t(s);
      if (o.counts) o.counts[ex] = int64_t(sl.cnt);
      if (o.sums) o.sums[ex] = sum;
      if (o.mins) o.mins[ex] = sl.mn;
      if (o.maxs) o.maxs[ex] = sl.mx;
      if (o.means) o.means[ex] = double(sum) / double(sl.cnt);
    }
    pos += uint32_t(__popcll(bal[j]));
  }
  const uint32_t ntiles = uint32_t((m + SC2_TILE - 1) / SC2_TILE);
  if (tile == ntiles - 1 && threadIdx.x == 0) *ng = int64_t(s_prefix + ttot);
}

static int64_t compact_slots(const GroupSlot* slots, uint64_t m, const int32_t* ukey_of,
                             int32_t kmin, const GroupAggOut& o, hipStream_t s) {
  if (m == 0) return 0;
  const uint32_t tiles = uint32_t((m + SC2_TILE - 1) / SC2_TILE);
  uint64_t* lb = lookback_words(tiles, s);
  const uint32_t epoch = lookback_epoch(s);
  uint32_t* ctr = scratch_t<uint32_t>(S_META2, 8) + 4;
  fill_u32(ctr, 0, 1, s);
  int64_t* ng_d = scratch_t<int64_t>(S_WORK2, 1);
  {
    // reads the slots, writes 40 B per present group
    ProfScope prof("slot_compact", s, double(m) * sizeof(GroupSlot));
    hipLaunchKernelGGL(k_slot_compact, dim3(tiles), dim3(SC2_BLOCK), 0, s, slots, m, ukey_of,
                       kmin, o, lb, ctr, epoch, ng_d);
    NFH_LAUNCH_CHECK();
  }
  return read_scalar(ng_d, s);
}

// Compact the valid rows (keys + values) in row order.
struct ValidRowsOp {
  const int32_t* keys;
  const int32_t* vals;
  const uint8_t* v;
  const uint8_t* kv;
  int32_t* ck;
  int32_t* cv;
  uint32_t* m;
  __device__ bool ok(uint32_t i) const {
    return (v == nullptr || v[i] != 0) && (kv == nullptr || kv[i] != 0);
  }
  __device__ uint32_t value(uint32_t i) const { return ok(i) ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t ex) const {
    if (ok(i)) {
      ck[ex] = keys[i];
      if (cv) cv[ex] = vals[i];
    }
  }
  __device__ void total(uint32_t t) const { *m = t; }
};

bool key_range(const int32_t* keys, const uint8_t* v, const uint8_t* kv, size_t n, int32_t* kmin,
               int32_t* kmax, hipStream_t s) {
  int* r = scratch_t<int>(S_MISC, 2);
  fill_u32(r, uint32_t(INT32_MAX), 1, s);
  fill_u32(r + 1, uint32_t(INT32_MIN), 1, s);
  {
    ProfScope prof("key_range", s, double(n) * (4.0 + (v ? 1.0 : 0.0) + (kv ? 1.0 : 0.0)));
    const unsigned grid = grid_for(n / 64 + 1, 256, 4 * num_cus());
    if (vec_ok(keys, nullptr, v, kv))
      hipLaunchKernelGGL(k_key_range<true>, dim3(grid), dim3(256), 0, s, keys, v, kv,
                         uint32_t(n), r);
    else
      hipLaunchKernelGGL(k_key_range<false>, dim3(grid), dim3(256), 0, s, keys, v, kv,
                         uint32_t(n), r);
    NFH_LAUNCH_CHECK();
  }
  int* h = static_cast<int*>(pinned(8));
  NFH_CHECK(hipMemcpyAsync(h, r, 8, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  *kmin = h[0];
  *kmax = h[1];
  return h[0] <= h[1];
}

size_t dense_span_limit(size_t n) {
  const size_t lim = std::max<size_t>(size_t(1) << 25, 2 * n);
  return std::min<size_t>(lim, size_t(1) << 28);
}

// Hot-key table from a strided sample of the keys; returns false when the
// input is too small for privatisation to pay.
static bool build_hot_table(const int32_t* keys, size_t n, int2** table, int32_t** hot_keys,
                            uint32_t** hot_n, hipStream_t s) {
  if (n < size_t(HOT_SAMPLE) * 8) return false;
  const uint32_t S = HOT_SAMPLE;
  unsigned char* base = static_cast<unsigned char*>(
      scratch(S_WORK3, size_t(S) * 4 + size_t(S) * 8 + HOT_TBL * 8 + HOT_MAX * 4 + 256));
  int32_t* sample = reinterpret_cast<int32_t*>(base);
  int2* cand = reinterpret_cast<int2*>(sample + S);
  *table = cand + S;
  *hot_keys = reinterpret_cast<int32_t*>(*table + HOT_TBL);
  *hot_n = reinterpret_cast<uint32_t*>(*hot_keys + HOT_MAX);
  uint32_t* ncand = *hot_n + 4;
  fill_u32(ncand, 0, 1, s);
  hipLaunchKernelGGL(k_sample, dim3(S / 256), dim3(256), 0, s, keys, uint32_t(n), S, sample);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_hot_count, dim3(HOT_PARTS), dim3(1024), 0, s, sample, S, cand, ncand);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_hot_pick, dim3(1), dim3(1024), 0, s, cand, ncand, *hot_keys, *table,
                     *hot_n);
  NFH_LAUNCH_CHECK();
  return true;
}

// Strategy of the next group-by on this thread (0: the rule below) and the
// one the last group-by used: 1 dense (LDS hot keys + global spill), 2
// partitioned (bucket passes, no per-row global atomics; 2048-key buckets),
// 3 sort-based (radix unique + aggregation by code), 4 hash-scatter (LDS
// hash tables + one global hash table), 5 partitioned with 4096-key buckets
// (half the partition count).  A forced strategy that cannot run on the
// input (span too large, key mask for the partitioned path) falls back to
// the rule; a hash table that overflows falls back to the sort-based path.
namespace {
thread_local int t_strategy = 0;
thread_local int t_last_strategy = 0;
}  // namespace
int set_groupby_strategy(int s) {
  const int old = t_strategy;
  t_strategy = (s >= 0 && s <= 5) ? s : 0;
  return old;
}
int last_groupby_strategy() { return t_last_strategy; }

// Forced hash-scatter is refused above this many rows: a labelling run that
// forced every strategy on columns of 2^20..2^26 rows ended in a GPU memory
// fault whose kernel and column are unknown, and hash-scatter (with its
// sort-based fallback) is the one strategy that only a forced call reaches.
// The built-in rule never picks it, so unforced calls are not affected.
constexpr size_t HASH_FORCE_MAX_ROWS = size_t(1) << 22;
bool forced_strategy_refused(size_t n) { return t_strategy == 4 && n > HASH_FORCE_MAX_ROWS; }

int64_t group_agg_device(const int32_t* keys, const int32_t* vals, const uint8_t* v,
                         const uint8_t* kv, size_t n, const GroupAggOut& o, hipStream_t s,
                         const int32_t* known_range) {
  if (n == 0) return 0;
  int32_t kmin, kmax;
  if (known_range) {
    kmin = known_range[0];
    kmax = known_range[1];
    if (kmin > kmax) return 0;
  } else if (!key_range(keys, v, kv, n, &kmin, &kmax, s)) {
    return 0;
  }
  const u64 span = u64(int64_t(kmax) - int64_t(kmin) + 1);
  const bool dense_ok = span <= dense_span_limit(n);
  const bool part_ok = dense_ok && gp_supported(span) && kv == nullptr;
  const bool wide_ok = dense_ok && gp_supported(span, GP_BUCKET_BITS_WIDE) && kv == nullptr;
  int strat = t_strategy;
  if ((strat == 1 && !dense_ok) || (strat == 2 && !part_ok) || (strat == 5 && !wide_ok))
    strat = 0;
  if (strat == 0) strat = part_ok ? 2 : dense_ok ? 1 : 3;
  t_last_strategy = strat;
  if (strat == 2 || strat == 5) {
    // Partitioned path: no per-row global atomics; the bucket pass writes
    // the groups themselves (no slot array)
    int2* table = nullptr;
    int32_t* hot_keys = nullptr;
    uint32_t* hot_n = nullptr;
    if (!build_hot_table(keys, n, &table, &hot_keys, &hot_n, s)) table = nullptr;
    return gp_aggregate(keys, vals, v, n, kmin, span, table, hot_keys, hot_n, o, s,
                        strat == 5 ? GP_BUCKET_BITS_WIDE : GP_BUCKET_BITS);
  }
  if (strat == 1) {
    GroupSlot* slots = scratch_t<GroupSlot>(S_WORK0, span);
    hipLaunchKernelGGL(k_slot_init, dim3(grid_for(span, 256, 16 * num_cus())), dim3(256), 0, s,
                       slots, span);
    NFH_LAUNCH_CHECK();
    int2* table = nullptr;
    int32_t* hot_keys = nullptr;
    uint32_t* hot_n = nullptr;
    if (!build_hot_table(keys, n, &table, &hot_keys, &hot_n, s)) table = nullptr;
    {
      ProfScope prof("group_acc", s, double(n) * (8.0 + (v ? 1.0 : 0.0) + (kv ? 1.0 : 0.0)));
      const unsigned grid = grid_for(n / 4 + 1, ACC_BLOCK, 4 * num_cus());
      if (vec_ok(keys, vals, v, kv))
        hipLaunchKernelGGL(k_group_acc<true>, dim3(grid), dim3(ACC_BLOCK), 0, s, keys, vals, v,
                           kv, uint32_t(n), kmin, slots, nullptr, table, hot_keys, hot_n);
      else
        hipLaunchKernelGGL(k_group_acc<false>, dim3(grid), dim3(ACC_BLOCK), 0, s, keys, vals, v,
                           kv, uint32_t(n), kmin, slots, nullptr, table, hot_keys, hot_n);
      NFH_LAUNCH_CHECK();
    }
    return compact_slots(slots, span, nullptr, kmin, o, s);
  }
  if (strat == 4) {
    int2* table = nullptr;
    int32_t* hot_keys = nullptr;
    uint32_t* hot_n = nullptr;
    if (!build_hot_table(keys, n, &table, &hot_keys, &hot_n, s)) table = nullptr;
    const int64_t ng = hash_group_agg(keys, vals, v, kv, n, span, table, hot_keys, hot_n, o, s);
    if (ng >= 0) return ng;
    t_last_strategy = 3;  // the table overflowed: sort-based
  }
  // sparse: compact valid rows, radix unique, aggregate by code
  int32_t* ck = scratch_t<int32_t>(S_WORK1, n);
  int32_t* cv = scratch_t<int32_t>(S_WORK3, n);
  uint32_t* m_d = scratch_t<uint32_t>(S_MISC, 4) + 2;
  scan_tiles(ValidRowsOp{keys, vals, v, kv, ck, cv, m_d}, n, s);
  const uint32_t m = read_scalar(m_d, s);
  int32_t* uniq = scratch_t<int32_t>(S_T0, m);
  int32_t* inv = scratch_t<int32_t>(S_T1, m);
  int32_t* perm = scratch_t<int32_t>(S_T2, m);
  uint32_t* t0 = scratch_t<uint32_t>(S_T3, m);
  uint32_t* t1 = scratch_t<uint32_t>(S_T4, m);
  int32_t* tp = scratch_t<int32_t>(S_T5, m);
  const int64_t ng = unique_inverse_device<int32_t>(ck, m, uniq, inv, perm, t0, t1, tp, s);
  GroupSlot* slots = scratch_t<GroupSlot>(S_WORK0, size_t(ng));
  hipLaunchKernelGGL(k_slot_init, dim3(grid_for(ng, 256, 16 * num_cus())), dim3(256), 0, s,
                     slots, u64(ng));
  NFH_LAUNCH_CHECK();
  if (m > 0) {
    hipLaunchKernelGGL(k_group_acc<false>, dim3(grid_for(m, ACC_BLOCK, 4 * num_cus())),
                       dim3(ACC_BLOCK), 0, s, ck, cv, nullptr, nullptr, m, 0, slots, inv,
                       nullptr, nullptr, nullptr);
    NFH_LAUNCH_CHECK();
  }
  return compact_slots(slots, u64(ng), uniq, 0, o, s);
}

// ---------------------------------------------------------------- unique
// Dense unique.  k_mark sets one flag byte per present key offset
// o = key - kmin with plain byte stores (set-once: most rows only read, and
// a stale read only repeats a store of 1).  The scan packs the presence
// bits into one 16-byte rank entry per 96 offsets: the number of present
// offsets below the entry, then the entry's 96 presence bits, so the dense
// code of a present key is base + popc(bits below o) -- one random 16-byte
// lookup per row over span / 6 bytes (2.7 MiB at 2^24 ids, which stays in
// an XCD's 4 MiB L2; an 8-byte word per 32 offsets made 4 MiB and took
// 0.90 ms for the inverse at 2^28 rows against 0.75).  Keys of the sampled
// hot-key table are resolved in LDS in both row passes; they stay off the
// global tables.

template <bool VEC>
__global__ __launch_bounds__(256) void k_mark(const int32_t* __restrict__ keys,
                                              const uint8_t* __restrict__ v, uint32_t n,
                                              int32_t kmin, uint8_t* __restrict__ flag,
                                              const int2* __restrict__ table,
                                              const int32_t* __restrict__ hot_keys,
                                              const uint32_t* __restrict__ hot_n) {
  __shared__ int32_t t_key[HOT_TBL];
  __shared__ int32_t t_idx[HOT_TBL];
  __shared__ uint32_t s_seen[HOT_MAX];
  hot_load(table, t_key, t_idx);
  const uint32_t nh = table ? *hot_n : 0u;
  for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) s_seen[b] = 0;
  __syncthreads();
  // true when the row still needs the global flag
  auto cold = [&](int32_t k) -> bool {
    if (nh == 0) return true;
    const int32_t idx = hot_index(t_key, t_idx, k);
    if (idx < 0) return true;
    s_seen[idx] = 1u;
    return false;
  };
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t tail = 0;
  if constexpr (VEC) {
    // all 4*RQ flag reads of a thread are issued before any store (rows
    // with equal keys may both store 1)
    const uint32_t nq = n / 4;
    uint32_t q = tid;
    for (; q + (RQ - 1) * stride < nq; q += RQ * stride) {
      uint32_t off[4 * RQ], f[4 * RQ];
      uint32_t need = 0;
#pragma unroll
      for (int u = 0; u < RQ; ++u) {
        const uint32_t qq = q + uint32_t(u) * stride;
        const int4 k4 = ld_stream4(keys, qq);
        const uint32_t m =
            v ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(v) + qq) : 0x01010101u;
        const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          off[4 * u + c] = uint32_t(kk[c] - kmin);
          if (((m >> (8 * c)) & 0xFFu) && cold(kk[c])) need |= 1u << (4 * u + c);
        }
      }
      // every lane loads (rows that need no flag read the in-range byte 0),
      // so the 4 * RQ random reads are all in flight at once
#pragma unroll
      for (int j = 0; j < 4 * RQ; ++j) f[j] = flag[((need >> j) & 1u) ? off[j] : 0u];
#pragma unroll
      for (int j = 0; j < 4 * RQ; ++j)
        if (((need >> j) & 1u) && f[j] == 0u) flag[off[j]] = 1u;
    }
    for (; q < nq; q += stride) {
      const int4 k4 = reinterpret_cast<const int4*>(keys)[q];
      const uint32_t m = v ? reinterpret_cast<const uint32_t*>(v)[q] : 0x01010101u;
      const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (!((m >> (8 * c)) & 0xFFu) || !cold(kk[c])) continue;
        uint8_t* p = flag + uint32_t(kk[c] - kmin);
        if (*p == 0u) *p = 1u;
      }
    }
    tail = nq * 4;
  }
  for (uint32_t i = tail + tid; i < n; i += stride) {
    if ((v != nullptr && v[i] == 0) || !cold(keys[i])) continue;
    uint8_t* p = flag + uint32_t(keys[i] - kmin);
    if (*p == 0u) *p = 1u;
  }
  if (nh == 0) return;
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x)
    if (s_seen[b]) flag[uint32_t(hot_keys[b] - kmin)] = 1u;
}

// The 32 flag bytes of word w (each 0 or 1) as a bit mask.
__device__ __forceinline__ uint32_t flag_word(const uint8_t* __restrict__ flag, uint32_t w) {
  const uint4* p = reinterpret_cast<const uint4*>(flag) + 2 * size_t(w);
  const uint4 a = p[0], b = p[1];
  const uint32_t c[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t x = c[i];
    const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
    bits |= nib << (4 * i);
  }
  return bits;
}

// Rank entries: {present offsets below, bits of offsets 96e .. 96e+31,
// 96e+32 .. 96e+63, 96e+64 .. 96e+95}.
typedef uint32_t rank_entry __attribute__((ext_vector_type(4)));
constexpr uint32_t RANK_SPAN = 96;

// Exclusive scan of the per-word presence counts (one item per 32 offsets,
// so a scan tile covers 2^17 offsets); word w stores its bits into its
// rank entry, the entry's base when it is the entry's first word, and its
// present keys (ascending) to uniq.  Bits(w) gives the 32 presence bits of
// word w.  Components of the last entry past the last word are zeroed.
template <class Bits>
struct RankOp {
  Bits bits;
  uint32_t* ent;  // rank entries as 4 words each
  uint32_t nw;
  int32_t kmin;
  int32_t* uniq;
  int64_t* ng;
  __device__ uint32_t value(uint32_t w) const { return uint32_t(__popc(bits(w))); }
  __device__ void emit(uint32_t w, uint32_t ex) const {
    uint32_t b = bits(w);
    const uint32_t e = w / 3, j = w - 3 * e;
    ent[4 * e + 1 + j] = b;
    if (j == 0) ent[4 * e] = ex;
    if (w == nw - 1)
      for (uint32_t k = j + 1; k < 3; ++k) ent[4 * e + 1 + k] = 0u;
    const uint32_t k0 = uint32_t(kmin) + (w << 5);  // wraps only for absent keys
    while (b) {
      uniq[ex++] = int32_t(k0 + uint32_t(__ffs(b) - 1));
      b &= b - 1u;
    }
  }
  __device__ void total(uint32_t t) const { *ng = int64_t(t); }
};

// The 32 flag bytes of word w (each 0 or 1) as a bit mask.
struct FlagBits {
  const uint8_t* flag;
  __device__ uint32_t operator()(uint32_t w) const { return flag_word(flag, w); }
};

// Partitioned mark (inputs of 2^22 rows and more).  k_mark's random flag
// reads miss L2 for most tail keys: at 2^28 Zipf(1.1) rows over 2^24 ids
// it moved 7.3 GiB for 1.25 GiB of keys and masks.  Here the presence bits
// are set in LDS instead:
//   pass A (k_umark_part) reads keys and mask once; hot keys only set a
//     seen bit, every other valid row's offset goes to its tile's segment
//     for range r = offset >> UM_RANGE_BITS (ranks from LDS counters), and
//     the tile records (start, count) per range;
//   pass B (k_umark_bits) gives workgroup (r, g) every G-th tile of range
//     r; it ORs those offsets into a 2^20-bit LDS bitmap (128 KB) and
//     stores the bitmap to its own staging slice (g == 0 also adds the hot
//     keys of range r that pass A saw);
//   k_umark_merge ORs the G slices of each range into the final bitmap.
// Measured at 2^28 rows: 0.50 + 0.20 + 0.014 ms against 1.30 ms for k_mark.
constexpr int UM_RANGE_BITS = 20;
constexpr uint32_t UM_RANGE_WORDS = 1u << (UM_RANGE_BITS - 5);
constexpr int UM_MAX_RANGES = 256;
constexpr int UM_RQ = 8;                          // quads per thread per tile
constexpr uint32_t UM_TILE_QUADS = 256u * UM_RQ;  // 8192 rows
constexpr int UM_B_THREADS = 1024;

__global__ __launch_bounds__(256) void k_umark_part(const int32_t* __restrict__ keys,
                                                    const uint8_t* __restrict__ v, uint32_t n,
                                                    int32_t kmin, uint32_t nranges,
                                                    const int2* __restrict__ table,
                                                    const uint32_t* __restrict__ hot_n,
                                                    uint32_t* __restrict__ rec,
                                                    uint2* __restrict__ entry, uint32_t ntiles,
                                                    uint8_t* __restrict__ hot_seen) {
  __shared__ int32_t t_key[HOT_TBL], t_idx[HOT_TBL];
  __shared__ uint32_t s_seen[HOT_MAX];
  __shared__ uint32_t s_cnt[UM_MAX_RANGES], s_base[UM_MAX_RANGES];
  hot_load(table, t_key, t_idx);
  const uint32_t nh = table ? *hot_n : 0u;
  for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) s_seen[b] = 0;
  const uint32_t nq = n / 4, nrem = n & 3u;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    for (uint32_t r = threadIdx.x; r < nranges; r += blockDim.x) s_cnt[r] = 0;
    __syncthreads();
    int4 k4[UM_RQ];
    uint32_t m[UM_RQ];
#pragma unroll
    for (int u = 0; u < UM_RQ; ++u) {
      const uint32_t qq = t * UM_TILE_QUADS + uint32_t(u) * 256u + threadIdx.x;
      k4[u] = make_int4(0, 0, 0, 0);
      m[u] = 0;
      if (qq < nq) {
        k4[u] = ld_stream4(keys, qq);
        m[u] = v ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(v) + qq)
                 : 0x01010101u;
      } else if (qq == nq && nrem != 0) {  // the last 1-3 rows
        const uint32_t i = 4 * nq;
        k4[u].x = keys[i];
        k4[u].y = nrem > 1 ? keys[i + 1] : 0;
        k4[u].z = nrem > 2 ? keys[i + 2] : 0;
        for (uint32_t c = 0; c < nrem; ++c) m[u] |= (v == nullptr || v[i + c]) ? 1u << (8 * c) : 0u;
      }
    }
    uint32_t off[4 * UM_RQ], loc[4 * UM_RQ];
#pragma unroll
    for (int u = 0; u < UM_RQ; ++u) {
      const int32_t kk[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = 4 * u + c;
        off[j] = 0xFFFFFFFFu;
        if ((m[u] >> (8 * c)) & 0xFFu) {
          const int idx = nh ? hot_index(t_key, t_idx, kk[c]) : -1;
          if (idx >= 0) {
            s_seen[idx] = 1u;
          } else {
            off[j] = uint32_t(kk[c] - kmin);
            loc[j] = atomicAdd(&s_cnt[off[j] >> UM_RANGE_BITS], 1u);
          }
        }
      }
    }
    __syncthreads();
    const uint32_t tb = t * UM_TILE_QUADS * 4u;
    if (threadIdx.x < WAVE) {  // exclusive scan of the range counts, 4 per lane
      const uint32_t l = threadIdx.x;
      uint32_t c[4], x = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        c[i] = 4 * l + i < nranges ? s_cnt[4 * l + i] : 0u;
        x += c[i];
      }
      const uint32_t own = x;
      for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (l >= uint32_t(o)) x += y;
      }
      uint32_t ex = x - own;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t r = 4 * l + i;
        if (r < nranges) {
          s_base[r] = ex;
          entry[size_t(r) * ntiles + t] = make_uint2(tb + ex, c[i]);
        }
        ex += c[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4 * UM_RQ; ++j)
      if (off[j] != 0xFFFFFFFFu) rec[tb + s_base[off[j] >> UM_RANGE_BITS] + loc[j]] = off[j];
  }
  if (nh == 0) return;
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x)
    if (s_seen[b]) hot_seen[b] = 1u;
}

// Workgroup blockIdx = r * G + g; one wave per tile entry.  `wr` words of
// the bitmap are live per range (UM_RANGE_WORDS unless the span is
// smaller), and slice blockIdx of the staging holds them.
__global__ __launch_bounds__(UM_B_THREADS) void k_umark_bits(
    const uint32_t* __restrict__ rec, const uint2* __restrict__ entry, uint32_t ntiles,
    uint32_t G, uint32_t wr, const int32_t* __restrict__ hot_keys,
    const uint32_t* __restrict__ hot_n, const uint8_t* __restrict__ hot_seen, int32_t kmin,
    uint32_t* __restrict__ staging) {
  constexpr uint32_t NW = UM_B_THREADS / WAVE;
  constexpr uint32_t LOW = (1u << UM_RANGE_BITS) - 1u;
  __shared__ uint32_t s_bits[UM_RANGE_WORDS];
  const uint32_t r = blockIdx.x / G, g = blockIdx.x % G;
  for (uint32_t w = threadIdx.x; w < wr; w += UM_B_THREADS) s_bits[w] = 0;
  __syncthreads();
  if (g == 0 && hot_n != nullptr) {
    const uint32_t nh = *hot_n;
    for (uint32_t b = threadIdx.x; b < nh; b += UM_B_THREADS)
      if (hot_seen[b]) {
        const uint32_t o = uint32_t(hot_keys[b] - kmin);
        if ((o >> UM_RANGE_BITS) == r) atomicOr(&s_bits[(o & LOW) >> 5], 1u << (o & 31u));
      }
  }
  const uint32_t lane = uint32_t(lane_id());
  // The wave's entries are t0, t0 + stride, ...: 64 of them are loaded at
  // once (one per lane) and handed out by shuffles, so only the records'
  // loads wait on memory in the loop below.
  const uint32_t t0 = g + G * uint32_t(wave_id()), stride = G * NW;
  for (uint32_t kb = 0; t0 + kb * stride < ntiles; kb += WAVE) {  // uniform over the wave
    const uint32_t tl = t0 + (kb + lane) * stride;
    const uint2 el = tl < ntiles ? entry[size_t(r) * ntiles + tl] : make_uint2(0u, 0u);
    const uint32_t kn = min(uint32_t(WAVE), (ntiles - t0 - kb * stride + stride - 1) / stride);
    for (uint32_t k = 0; k < kn; ++k) {
      const uint32_t ex = uint32_t(__shfl(int(el.x), int(k)));
      const uint32_t ey = uint32_t(__shfl(int(el.y), int(k)));
      for (uint32_t j = lane; j < ey; j += 4 * WAVE) {
        uint32_t o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t jj = j + uint32_t(WAVE) * u;
          o[u] = jj < ey ? __builtin_nontemporal_load(rec + ex + jj) : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (o[u] != 0xFFFFFFFFu) atomicOr(&s_bits[(o[u] & LOW) >> 5], 1u << (o[u] & 31u));
      }
    }
  }
  __syncthreads();
  uint32_t* dst = staging + size_t(blockIdx.x) * wr;
  for (uint32_t w = threadIdx.x; w < wr; w += UM_B_THREADS) dst[w] = s_bits[w];
}

__global__ __launch_bounds__(256) void k_umark_merge(const uint32_t* __restrict__ staging,
                                                     uint32_t G, uint32_t wr, uint32_t nw,
                                                     uint32_t* __restrict__ bits) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += stride) {
    const uint32_t r = w / UM_RANGE_WORDS, wl = w % UM_RANGE_WORDS;
    const uint32_t* p = staging + size_t(r) * G * wr + wl;
    uint32_t x = 0;
    for (uint32_t g = 0; g < G; ++g) x |= p[size_t(g) * wr];
    bits[w] = x;
  }
}

// Words of a presence bitmap (one 32-bit word per 32 offsets).
struct BitmapBits {
  const uint32_t* bits;
  __device__ uint32_t operator()(uint32_t w) const { return bits[w]; }
};

// Entry of offset o (o / 96 as (o / 32) / 3 by multiply-high) and o's bit
// position within it.
__device__ __forceinline__ uint32_t rank_entry_of(uint32_t o) {
  return __umulhi(o >> 5, 0xAAAAAAABu) >> 1;
}
__device__ __forceinline__ int32_t rank_in_entry(rank_entry en, uint32_t b) {
  const uint32_t m0 = b >= 32 ? 0xFFFFFFFFu : ((1u << b) - 1u);
  const uint32_t m1 = b >= 64 ? 0xFFFFFFFFu : (b <= 32 ? 0u : ((1u << (b - 32)) - 1u));
  const uint32_t m2 = b <= 64 ? 0u : ((1u << (b - 64)) - 1u);
  return int32_t(en.x + uint32_t(__popc(en.y & m0) + __popc(en.z & m1) + __popc(en.w & m2)));
}
__device__ __forceinline__ int32_t rank_code(const rank_entry* __restrict__ ent, uint32_t o) {
  const uint32_t e = rank_entry_of(o);
  return rank_in_entry(ent[e], o - e * RANK_SPAN);
}

// Dense codes of the hot keys (read by k_inv from LDS).  A hot key with no
// valid row is never looked up; it may even lie outside [kmin, kmin + span)
// (the sample ignores validity), so only in-range keys read the rank entries.
__global__ void k_hot_codes(const int32_t* __restrict__ hot_keys,
                            const uint32_t* __restrict__ hot_n, int32_t kmin, uint32_t span,
                            const rank_entry* __restrict__ ent, int32_t* __restrict__ codes) {
  const uint32_t nh = *hot_n;
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x) {
    const uint32_t o = uint32_t(hot_keys[b]) - uint32_t(kmin);
    codes[b] = o < span ? rank_code(ent, o) : -1;
  }
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_inv(const int32_t* __restrict__ keys,
                                             const uint8_t* __restrict__ v, uint32_t n,
                                             int32_t kmin, const rank_entry* __restrict__ ent,
                                             const int2* __restrict__ table,
                                             const uint32_t* __restrict__ hot_n,
                                             const int32_t* __restrict__ hot_code,
                                             int32_t* __restrict__ inv) {
  __shared__ int32_t t_key[HOT_TBL];
  __shared__ int32_t t_idx[HOT_TBL];
  __shared__ int32_t s_code[HOT_MAX];
  hot_load(table, t_key, t_idx);
  const uint32_t nh = table ? *hot_n : 0u;
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x) s_code[b] = hot_code[b];
  __syncthreads();
  auto code = [&](int32_t k, bool ok) -> int32_t {
    if (!ok) return -1;
    if (nh != 0) {
      const int32_t idx = hot_index(t_key, t_idx, k);
      if (idx >= 0) return s_code[idx];
    }
    return rank_code(ent, uint32_t(k - kmin));
  };
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t tail = 0;
  if constexpr (VEC) {
    const uint32_t nq = n / 4;
    uint32_t q = tid;
    // RQ quads per thread: key loads, then all code lookups, then stores
    for (; q + (RQ - 1) * stride < nq; q += RQ * stride) {
      int4 k4[RQ];
      uint32_t m[RQ];
#pragma unroll
      for (int u = 0; u < RQ; ++u) {
        const uint32_t qq = q + uint32_t(u) * stride;
        k4[u] = ld_stream4(keys, qq);
        m[u] = v ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(v) + qq)
                 : 0x01010101u;
      }
      // per row: -1 (NULL), the hot code (LDS), or a word-table lookup;
      // every lane issues its lookup (rows without one read word 0), so
      // the 4 * RQ random reads are all in flight at once
      int32_t c[4 * RQ];
      uint32_t off[4 * RQ];
#pragma unroll
      for (int u = 0; u < RQ; ++u) {
        const int32_t kk[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool ok = (m[u] >> (8 * e)) & 0xFFu;
          const int32_t idx = (ok && nh != 0) ? hot_index(t_key, t_idx, kk[e]) : -1;
          c[4 * u + e] = !ok ? -1 : (idx >= 0 ? s_code[idx] : INT32_MIN);
          off[4 * u + e] = (ok && idx < 0) ? uint32_t(kk[e] - kmin) : 0u;
        }
      }
      rank_entry en[4 * RQ];
      uint32_t bi[4 * RQ];
#pragma unroll
      for (int j = 0; j < 4 * RQ; ++j) {
        const uint32_t e = rank_entry_of(off[j]);
        bi[j] = off[j] - e * RANK_SPAN;
        en[j] = ent[e];
      }
#pragma unroll
      for (int j = 0; j < 4 * RQ; ++j)
        if (c[j] == INT32_MIN) c[j] = rank_in_entry(en[j], bi[j]);
      int4 o[RQ];
#pragma unroll
      for (int u = 0; u < RQ; ++u) o[u] = make_int4(c[4 * u], c[4 * u + 1], c[4 * u + 2], c[4 * u + 3]);
#pragma unroll
      for (int u = 0; u < RQ; ++u) st_stream4(inv, q + uint32_t(u) * stride, o[u]);
    }
    for (; q < nq; q += stride) {
      const int4 k4 = reinterpret_cast<const int4*>(keys)[q];
      const uint32_t m = v ? reinterpret_cast<const uint32_t*>(v)[q] : 0x01010101u;
      int4 o;
      o.x = code(k4.x, m & 0xFFu);
      o.y = code(k4.y, m & 0xFF00u);
      o.z = code(k4.z, m & 0xFF0000u);
      o.w = code(k4.w, m & 0xFF000000u);
      reinterpret_cast<int4*>(inv)[q] = o;
    }
    tail = nq * 4;
  }
  for (uint32_t i = tail + tid; i < n; i += stride)
    inv[i] = code(keys[i], v == nullptr || v[i]);
}

__global__ void k_inv_scatter(const uint8_t* __restrict__ v, uint32_t n,
                              const int32_t* __restrict__ cinv, const uint32_t* __restrict__ pos,
                              int32_t* __restrict__ inv) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    inv[i] = v[i] ? cinv[pos[i]] : -1;
}

struct ValidPosOp {
  const int32_t* keys;
  const uint8_t* v;
  int32_t* ck;
  uint32_t* pos;
  uint32_t* m;
  __device__ uint32_t value(uint32_t i) const { return v[i] ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t ex) const {
    pos[i] = ex;
    if (v[i]) ck[ex] = keys[i];
  }
  __device__ void total(uint32_t t) const { *m = t; }
};

int64_t unique_valid_device(const int32_t* keys, const uint8_t* v, size_t n, int32_t* uniq,
                            int32_t* inv, hipStream_t s) {
  if (n == 0) return 0;
  int32_t kmin, kmax;
  if (!key_range(keys, v, nullptr, n, &kmin, &kmax, s)) {
    fill_u32(inv, 0xFFFFFFFFu, n, s);
    return 0;
  }
  const u64 span = u64(int64_t(kmax) - int64_t(kmin) + 1);
  int64_t* ng_d = scratch_t<int64_t>(S_WORK2, 1);
  if (span <= dense_span_limit(n)) {
    const size_t nw = size_t((span + 31) >> 5);
    const size_t ne = (nw + 2) / 3;
    rank_entry* ent = scratch_t<rank_entry>(S_WORK1, ne);
    int32_t* hot_code = scratch_t<int32_t>(S_T0, HOT_MAX);
    int2* table = nullptr;
    int32_t* hot_keys = nullptr;
    uint32_t* hot_n = nullptr;
    if (n < (size_t(1) << 22) || !build_hot_table(keys, n, &table, &hot_keys, &hot_n, s))
      table = nullptr;
    const bool vec = vec_ok(keys, nullptr, v, nullptr) && aligned(inv, 16);
    const unsigned grid = grid_for(n / 4 + 1, 256, 8 * num_cus());
    const uint32_t nranges = uint32_t((span + (u64(1) << UM_RANGE_BITS) - 1) >> UM_RANGE_BITS);
    if (vec && n >= (size_t(1) << 22) && nranges <= uint32_t(UM_MAX_RANGES)) {
      // partitioned mark into a presence bitmap
      const uint32_t G = std::max<uint32_t>(1u, 256u / nranges);
      const uint32_t wr = nranges > 1 ? UM_RANGE_WORDS : uint32_t(nw);
      const size_t nq_eff = n / 4 + ((n & 3) ? 1 : 0);
      const uint32_t ntiles = uint32_t((nq_eff + UM_TILE_QUADS - 1) / UM_TILE_QUADS);
      uint32_t* bits = scratch_t<uint32_t>(S_WORK0, nw);
      uint32_t* rec = scratch_t<uint32_t>(S_T2, n);
      uint2* entry = scratch_t<uint2>(S_T3, size_t(nranges) * ntiles);
      uint32_t* staging = scratch_t<uint32_t>(S_T4, size_t(nranges) * G * wr);
      uint8_t* hot_seen = scratch_t<uint8_t>(S_T1, HOT_MAX);
      fill_u32(hot_seen, 0, HOT_MAX / 4, s);
      {
        ProfScope prof("unique_part", s, double(n) * (4.0 + (v ? 1.0 : 0.0)));
        hipLaunchKernelGGL(k_umark_part, dim3(std::min<uint32_t>(ntiles, 8 * num_cus())),
                           dim3(256), 0, s, keys, v, uint32_t(n), kmin, nranges, table, hot_n,
                           rec, entry, ntiles, hot_seen);
        NFH_LAUNCH_CHECK();
      }
      {
        ProfScope prof("unique_bits", s);
        hipLaunchKernelGGL(k_umark_bits, dim3(nranges * G), dim3(UM_B_THREADS), 0, s, rec, entry,
                           ntiles, G, wr, hot_keys, table ? hot_n : nullptr, hot_seen, kmin,
                           staging);
        NFH_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_umark_merge, dim3(grid_for(nw, 256, 4 * num_cus())), dim3(256), 0, s,
                         staging, G, wr, uint32_t(nw), bits);
      NFH_LAUNCH_CHECK();
      scan_tiles(RankOp<BitmapBits>{BitmapBits{bits}, reinterpret_cast<uint32_t*>(ent),
                                    uint32_t(nw), kmin, uniq, ng_d},
                 nw, s);
    } else {
      uint8_t* flag = scratch_t<uint8_t>(S_WORK0, 32 * nw);
      fill_u32(flag, 0, 8 * nw, s);
      {
        ProfScope prof("unique_mark", s, double(n) * (4.0 + (v ? 1.0 : 0.0)));
        if (vec)
          hipLaunchKernelGGL(k_mark<true>, dim3(grid), dim3(256), 0, s, keys, v, uint32_t(n),
                             kmin, flag, table, hot_keys, hot_n);
        else
          hipLaunchKernelGGL(k_mark<false>, dim3(grid), dim3(256), 0, s, keys, v, uint32_t(n),
                             kmin, flag, table, hot_keys, hot_n);
        NFH_LAUNCH_CHECK();
      }
      scan_tiles(RankOp<FlagBits>{FlagBits{flag}, reinterpret_cast<uint32_t*>(ent), uint32_t(nw),
                                  kmin, uniq, ng_d},
                 nw, s);
    }
    if (table != nullptr) {
      hipLaunchKernelGGL(k_hot_codes, dim3(1), dim3(HOT_MAX), 0, s, hot_keys, hot_n, kmin,
                         uint32_t(span), ent, hot_code);
      NFH_LAUNCH_CHECK();
    }
    {
      ProfScope prof("unique_inv_dense", s, double(n) * (8.0 + (v ? 1.0 : 0.0)));
      // 64 workgroups per CU (0.77 ms at 2^28 Zipf rows with 16, 0.75 with
      // 32, 0.73 with 64)
      const unsigned grid_inv = grid_for(n / 4 + 1, 256, 64 * num_cus());
      if (vec)
        hipLaunchKernelGGL(k_inv<true>, dim3(grid_inv), dim3(256), 0, s, keys, v, uint32_t(n), kmin,
                           ent, table, hot_n, hot_code, inv);
      else
        hipLaunchKernelGGL(k_inv<false>, dim3(grid_inv), dim3(256), 0, s, keys, v, uint32_t(n),
                           kmin, ent, table, hot_n, hot_code, inv);
      NFH_LAUNCH_CHECK();
    }
    return read_scalar(ng_d, s);
  }
  int32_t* perm = scratch_t<int32_t>(S_T2, n);
  uint32_t* t0 = scratch_t<uint32_t>(S_T3, n);
  uint32_t* t1 = scratch_t<uint32_t>(S_T4, n);
  int32_t* tp = scratch_t<int32_t>(S_T5, n);
  if (v == nullptr) return unique_inverse_device<int32_t>(keys, n, uniq, inv, perm, t0, t1, tp, s);
  int32_t* ck = scratch_t<int32_t>(S_WORK1, n);
  uint32_t* pos = scratch_t<uint32_t>(S_WORK3, n);
  uint32_t* m_d = scratch_t<uint32_t>(S_MISC, 4) + 2;
  scan_tiles(ValidPosOp{keys, v, ck, pos, m_d}, n, s);
  const uint32_t m = read_scalar(m_d, s);
  int32_t* cinv = scratch_t<int32_t>(S_WORK0, m);
  const int64_t ng = unique_inverse_device<int32_t>(ck, m, uniq, cinv, perm, t0, t1, tp, s);
  hipLaunchKernelGGL(k_inv_scatter, dim3(grid_for(n, 256, 16 * num_cus())), dim3(256), 0, s, v,
                     uint32_t(n), cinv, pos, inv);
  NFH_LAUNCH_CHECK();
  return ng;
}

}  // namespace nfh

using namespace nfh;

// Strategy override for this thread's following group-bys (0 = rule,
// 1 dense, 2 partitioned, 3 sort-based, 4 hash-scatter, 5 partitioned with
// 4096-key buckets); returns the previous setting.
NF_EXPORT int32_t nfh_set_groupby_strategy(int32_t strategy) {
  return set_groupby_strategy(strategy);
}
// The strategy this thread's last group-by ran (0 before any).
NF_EXPORT int32_t nfh_groupby_last_strategy(void) { return last_groupby_strategy(); }

// Group-by over raw int32 keys with NULL masks (valid / key_valid may be
// null = all valid).  Outputs hold `cap` lanes; any output pointer may be
// null.  Returns ng >= 0, -1 null input, -3 n > i32::MAX or hash-scatter
// forced on more than 2^22 rows, -6 device error, -7 cap < ng (outputs
// unwritten).
NF_EXPORT int64_t nfh_group_agg_i32(const int32_t* keys, const int32_t* vals,
                                    const uint8_t* valid, const uint8_t* key_valid, size_t n,
                                    int64_t* ukeys, int64_t* counts, int64_t* sums,
                                    int32_t* mins, int32_t* maxs, double* means, size_t cap) {
  if (n == 0) return 0;
  if (!keys || !vals) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  if (forced_strategy_refused(n)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  const int32_t* v = st.in(vals, n, S_IN1);
  const uint8_t* vm = valid ? st.in(valid, n, S_IN2) : nullptr;
  const uint8_t* km = key_valid ? st.in(key_valid, n, S_IN3) : nullptr;
  GroupAggOut d{};
  if (cap >= n) {
    // ng <= n fits: write straight into the caller's (or staged) buffers.
    d.ukeys = st.out(ukeys, n, S_OUT0);
    d.counts = st.out(counts, n, S_OUT1);
    d.sums = st.out(sums, n, S_OUT2);
    d.mins = st.out(mins, n, S_OUT3);
    d.maxs = st.out(maxs, n, S_OUT4);
    d.means = st.out(means, n, S_OUT5);
    const int64_t ng = group_agg_device(k, v, vm, km, n, d, st.s);
    st.set_out_bytes(ukeys, size_t(ng) * 8);
    st.set_out_bytes(counts, size_t(ng) * 8);
    st.set_out_bytes(sums, size_t(ng) * 8);
    st.set_out_bytes(mins, size_t(ng) * 4);
    st.set_out_bytes(maxs, size_t(ng) * 4);
    st.set_out_bytes(means, size_t(ng) * 8);
    st.finish();
    return ng;
  }
  // Device outputs of smaller capacity: the kernels write them directly
  // when cap covers the group bound known from the key range.
  const bool all_dev = (!ukeys || is_device_ptr(ukeys)) && (!counts || is_device_ptr(counts)) &&
                       (!sums || is_device_ptr(sums)) && (!mins || is_device_ptr(mins)) &&
                       (!maxs || is_device_ptr(maxs)) && (!means || is_device_ptr(means));
  int32_t kmin = 0, kmax = -1;
  const bool any = key_range(k, vm, km, n, &kmin, &kmax, st.s);
  const u64 bound = any ? u64(int64_t(kmax) - int64_t(kmin) + 1) : 0;
  if (all_dev && bound <= cap) {
    d.ukeys = ukeys;
    d.counts = counts;
    d.sums = sums;
    d.mins = mins;
    d.maxs = maxs;
    d.means = means;
    const int32_t range[2] = {kmin, kmax};
    const int64_t ng = group_agg_device(k, v, vm, km, n, d, st.s, range);
    st.finish();
    return ng;
  }
  // Otherwise results land in device scratch first, then move to the
  // caller's buffers once ng <= cap is known.
  d.ukeys = ukeys ? scratch_t<int64_t>(S_OUT0, n) : nullptr;
  d.counts = counts ? scratch_t<int64_t>(S_OUT1, n) : nullptr;
  d.sums = sums ? scratch_t<int64_t>(S_OUT2, n) : nullptr;
  d.mins = mins ? scratch_t<int32_t>(S_OUT3, n) : nullptr;
  d.maxs = maxs ? scratch_t<int32_t>(S_OUT4, n) : nullptr;
  d.means = means ? scratch_t<double>(S_OUT5, n) : nullptr;
  const int64_t ng = group_agg_device(k, v, vm, km, n, d, st.s);
  if (size_t(ng) > cap) {
    st.finish();
    return TOO_SMALL;
  }
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) NFH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st.s));
  };
  copy(ukeys, d.ukeys, size_t(ng) * 8);
  copy(counts, d.counts, size_t(ng) * 8);
  copy(sums, d.sums, size_t(ng) * 8);
  copy(mins, d.mins, size_t(ng) * 4);
  copy(maxs, d.maxs, size_t(ng) * 4);
  copy(means, d.means, size_t(ng) * 8);
  st.finish();
  return ng;
  NFH_GUARD_END(int64_t)
}

// np.unique over the valid rows of int32 keys: uniq ascending (first ng
// lanes), inv[row] = code, -1 on NULL rows (cpu.py:4032-4036).  valid may
// be null.  Returns ng, -1 null input, -3 n > i32::MAX, -6 device error.
NF_EXPORT int64_t nfh_unique_valid_i32(const int32_t* keys, const uint8_t* valid, size_t n,
                                       int32_t* uniq_out, int32_t* inv_out) {
  if (n == 0) return 0;
  if (!keys || !uniq_out || !inv_out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  const uint8_t* vm = valid ? st.in(valid, n, S_IN1) : nullptr;
  int32_t* u = st.out(uniq_out, n, S_OUT0);
  int32_t* iv = st.out(inv_out, n, S_OUT1);
  const int64_t ng = unique_valid_device(k, vm, n, u, iv, st.s);
  st.set_out_bytes(uniq_out, size_t(ng) * 4);
  st.finish();
  return ng;
  NFH_GUARD_END(int64_t)
}
