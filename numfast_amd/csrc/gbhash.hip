// gbhash.hip -- hash-scatter group-by (strategy 4 of nfh_group_agg_i32):
// sum / count / min / max / mean of the rows whose key and value are valid,
// keys ascending, the same result contract as the other strategies
// (src/Drivers/CPU/_lib/carry.py:21-62; groupindex.py:3-8).
//
// For key columns whose span is too wide to index directly (the dense and
// partitioned strategies need span <= max(2^25, 2n)) but whose distinct
// keys are few, this replaces compaction + radix unique + aggregation by
// code with one pass over the rows:
//   1. hot keys (the sampled table shared with the other strategies)
//      aggregate in LDS;
//   2. every other row goes to its workgroup's LDS hash table (1024
//      slots, linear probing, at most HL_PROBE probes); rows that find no
//      room there go straight to one global open-addressing table
//      (HSlot: 64-bit key word, then sum / count / min / max), placed by a
//      64-bit compare-and-swap of the key word and aggregated with device
//      atomics;
//   3. each workgroup flushes its LDS entries and hot aggregates into the
//      global table (one insert per entry, not per row);
//   4. a scan compacts the present slots into (key, slot), the radix
//      argsort orders them by key, and one kernel writes the outputs.
// The global table has room for twice the distinct keys the rows can hold
// (capped at 2^22 slots); a probe that finds no room after HG_PROBE slots
// raises an overflow flag and the caller recomputes with the sort-based
// strategy, so the result never depends on the table size.
#include <algorithm>

#include "device.hpp"
#include "gbhash.hpp"
#include "hot.hpp"
#include "runtime.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace nfh {

typedef unsigned long long u64;

struct HSlot {
  u64 key;  // 0: empty; (1 << 32) | uint32(key): occupied
  u64 sum;  // wrapping int64
  uint32_t cnt;
  int32_t mn, mx;
  uint32_t pad;
};
static_assert(sizeof(HSlot) == 32, "one 32-byte sector per slot");

constexpr int HL_SLOTS = 1024;  // LDS hash slots per workgroup
constexpr int HL_PROBE = 8;
constexpr int HG_PROBE = 256;  // far above the longest probe at load 1/2
constexpr uint32_t HG_MAX_SLOTS = 1u << 22;
constexpr int HA_BLOCK = 256;

__device__ __forceinline__ u64 hkey(int32_t k) { return (u64(1) << 32) | u64(uint32_t(k)); }

// Slot of key k in the global table (inserted when absent), or null after
// HG_PROBE occupied slots (overflow raised).  A plain read that is stale
// can only show an empty slot that is occupied: the compare-and-swap then
// returns the occupant, and keys never change once written.
__device__ HSlot* g_slot(HSlot* __restrict__ tab, uint32_t mask, int32_t k,
                         uint32_t* __restrict__ overflow) {
  // once the table has overflowed the result is recomputed anyway
  if (__hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
    return nullptr;
  const u64 kw = hkey(k);
  uint32_t h = hot_mix(k) & mask;
  for (int p = 0; p < HG_PROBE; ++p) {
    u64 cur = tab[h].key;
    if (cur == kw) return tab + h;
    if (cur == 0) {
      cur = atomicCAS(&tab[h].key, 0ull, kw);
      if (cur == 0 || cur == kw) return tab + h;
    }
    h = (h + 1) & mask;
  }
  atomicOr(overflow, 1u);
  return nullptr;
}

// min / max only when the value would change them: a stale read shows an
// older value, which is no smaller (min) / no larger (max) than the current
// one, so skipping on it is always right.
__device__ __forceinline__ void g_add(HSlot* p, u64 sum, uint32_t cnt, int32_t mn, int32_t mx) {
  atomicAdd(&p->sum, sum);
  atomicAdd(&p->cnt, cnt);
  if (mn < p->mn) atomicMin(&p->mn, mn);
  if (mx > p->mx) atomicMax(&p->mx, mx);
}

__global__ void k_hslot_init(HSlot* __restrict__ tab, uint32_t m) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    HSlot z;
    z.key = 0;
    z.sum = 0;
    z.cnt = 0;
    z.mn = INT32_MAX;
    z.mx = INT32_MIN;
    z.pad = 0;
    tab[i] = z;
  }
}

template <bool VEC>
__global__ __launch_bounds__(HA_BLOCK) void k_hash_acc(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const uint8_t* __restrict__ vm, const uint8_t* __restrict__ km, uint32_t n,
    HSlot* __restrict__ tab, uint32_t mask, const int2* __restrict__ table,
    const int32_t* __restrict__ hot_keys, const uint32_t* __restrict__ hot_n,
    uint32_t* __restrict__ overflow) {
  __shared__ int32_t t_key[HOT_TBL];
  __shared__ int32_t t_idx[HOT_TBL];
  __shared__ u64 a_sum[HOT_MAX];
  __shared__ uint32_t a_cnt[HOT_MAX];
  __shared__ int32_t a_mn[HOT_MAX], a_mx[HOT_MAX];
  __shared__ u64 l_key[HL_SLOTS];
  __shared__ u64 l_sum[HL_SLOTS];
  __shared__ uint32_t l_cnt[HL_SLOTS];
  __shared__ int32_t l_mn[HL_SLOTS], l_mx[HL_SLOTS];
  const uint32_t nh = table ? *hot_n : 0u;
  hot_load(table, t_key, t_idx);
  for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) {
    a_sum[b] = 0;
    a_cnt[b] = 0;
    a_mn[b] = INT32_MAX;
    a_mx[b] = INT32_MIN;
  }
  for (int b = threadIdx.x; b < HL_SLOTS; b += blockDim.x) {
    l_key[b] = 0;
    l_sum[b] = 0;
    l_cnt[b] = 0;
    l_mn[b] = INT32_MAX;
    l_mx[b] = INT32_MIN;
  }
  __syncthreads();
  auto row = [&](int32_t k, int32_t x, bool ok) {
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
    const u64 kw = hkey(k);
    uint32_t h = hot_mix(k) & uint32_t(HL_SLOTS - 1);
    for (int p = 0; p < HL_PROBE; ++p) {
      u64 cur = l_key[h];
      if (cur == 0) cur = atomicCAS(&l_key[h], 0ull, kw);
      if (cur == 0 || cur == kw) {
        atomicAdd(&l_sum[h], (u64)(long long)x);
        atomicAdd(&l_cnt[h], 1u);
        if (x < l_mn[h]) atomicMin(&l_mn[h], x);
        if (x > l_mx[h]) atomicMax(&l_mx[h], x);
        return;
      }
      h = (h + 1) & uint32_t(HL_SLOTS - 1);
    }
    HSlot* g = g_slot(tab, mask, k, overflow);  // no room in LDS: spill
    if (g) g_add(g, (u64)(long long)x, 1u, x, x);
  };
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t tail = 0;
  if constexpr (VEC) {
    constexpr int R = 4;  // quads in flight per thread
    const uint32_t nq = n / 4;
    uint32_t q = tid;
    for (; q + (R - 1) * stride < nq; q += R * stride) {
      int4 k4[R], v4[R];
      uint32_t a[R], b[R];
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const uint32_t qq = q + uint32_t(u) * stride;
        k4[u] = reinterpret_cast<const int4*>(keys)[qq];
        v4[u] = reinterpret_cast<const int4*>(vals)[qq];
        a[u] = vm ? reinterpret_cast<const uint32_t*>(vm)[qq] : 0x01010101u;
        b[u] = km ? reinterpret_cast<const uint32_t*>(km)[qq] : 0x01010101u;
      }
      asm volatile("" ::: "memory");  // keep every load above the adds
#pragma unroll
      for (int u = 0; u < R; ++u) {
        row(k4[u].x, v4[u].x, (a[u] & 0xFFu) && (b[u] & 0xFFu));
        row(k4[u].y, v4[u].y, (a[u] & 0xFF00u) && (b[u] & 0xFF00u));
        row(k4[u].z, v4[u].z, (a[u] & 0xFF0000u) && (b[u] & 0xFF0000u));
        row(k4[u].w, v4[u].w, (a[u] & 0xFF000000u) && (b[u] & 0xFF000000u));
      }
    }
    for (; q < nq; q += stride) {
      const int4 k4 = reinterpret_cast<const int4*>(keys)[q];
      const int4 v4 = reinterpret_cast<const int4*>(vals)[q];
      const uint32_t a = vm ? reinterpret_cast<const uint32_t*>(vm)[q] : 0x01010101u;
      const uint32_t b = km ? reinterpret_cast<const uint32_t*>(km)[q] : 0x01010101u;
      row(k4.x, v4.x, (a & 0xFFu) && (b & 0xFFu));
      row(k4.y, v4.y, (a & 0xFF00u) && (b & 0xFF00u));
      row(k4.z, v4.z, (a & 0xFF0000u) && (b & 0xFF0000u));
      row(k4.w, v4.w, (a & 0xFF000000u) && (b & 0xFF000000u));
    }
    tail = nq * 4;
  }
  for (uint32_t i = tail + tid; i < n; i += stride)
    row(keys[i], vals[i], (vm == nullptr || vm[i] != 0) && (km == nullptr || km[i] != 0));
  __syncthreads();
  // flush: one global insert per LDS entry and per hot key with rows
  for (int b = threadIdx.x; b < HL_SLOTS; b += blockDim.x) {
    const u64 kw = l_key[b];
    if (kw == 0) continue;
    HSlot* g = g_slot(tab, mask, int32_t(uint32_t(kw)), overflow);
    if (g) g_add(g, l_sum[b], l_cnt[b], l_mn[b], l_mx[b]);
  }
  for (uint32_t b = threadIdx.x; b < nh; b += blockDim.x) {
    const uint32_t c = a_cnt[b];
    if (c == 0) continue;
    HSlot* g = g_slot(tab, mask, hot_keys[b], overflow);
    if (g) g_add(g, a_sum[b], c, a_mn[b], a_mx[b]);
  }
}

// Present slots -> (key, slot) in slot order; total = number of groups.
struct HashExtractOp {
  const HSlot* tab;
  int32_t* ek;
  uint32_t* es;
  uint32_t* ng;
  __device__ uint32_t value(uint32_t s) const { return tab[s].key != 0 ? 1u : 0u; }
  __device__ void emit(uint32_t s, uint32_t ex) const {
    const u64 kw = tab[s].key;
    if (kw == 0) return;
    ek[ex] = int32_t(uint32_t(kw));
    es[ex] = s;
  }
  __device__ void total(uint32_t t) const { *ng = t; }
};

// Group r (in key order) = slot es[perm[r]].
__global__ void k_hash_emit(const HSlot* __restrict__ tab, const uint32_t* __restrict__ es,
                            const int32_t* __restrict__ perm, uint32_t ng, GroupAggOut o) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < ng; r += stride) {
    const HSlot sl = tab[es[perm[r]]];
    const int64_t sum = int64_t(sl.sum);
    if (o.ukeys) o.ukeys[r] = int64_t(int32_t(uint32_t(sl.key)));
    if (o.counts) o.counts[r] = int64_t(sl.cnt);
    if (o.sums) o.sums[r] = sum;
    if (o.mins) o.mins[r] = sl.mn;
    if (o.maxs) o.maxs[r] = sl.mx;
    if (o.means) o.means[r] = double(sum) / double(sl.cnt);
  }
}

int64_t hash_group_agg(const int32_t* keys, const int32_t* vals, const uint8_t* v,
                       const uint8_t* kv, size_t n_, uint64_t span, const int2* table,
                       const int32_t* hot_keys, const uint32_t* hot_n, const GroupAggOut& o,
                       hipStream_t s) {
  const uint32_t n = uint32_t(n_);
  // room for twice the distinct keys the rows can hold, a power of two
  const uint64_t bound = std::max<uint64_t>(1, std::min<uint64_t>(n, span));
  uint32_t slots = 1024;
  while (slots < HG_MAX_SLOTS && uint64_t(slots) < 2 * bound) slots <<= 1;
  HSlot* tab = scratch_t<HSlot>(S_WORK0, slots);
  uint32_t* flags = scratch_t<uint32_t>(S_MISC, 8);  // [0] overflow, [1] ng
  fill_u32(flags, 0, 2, s);
  hipLaunchKernelGGL(k_hslot_init, dim3(grid_for(slots, 256, 16 * num_cus())), dim3(256), 0, s,
                     tab, slots);
  NFH_LAUNCH_CHECK();
  {
    ProfScope prof("hash_acc", s, double(n) * (8.0 + (v ? 1.0 : 0.0) + (kv ? 1.0 : 0.0)));
    const unsigned grid = grid_for(n / 4 + 1, HA_BLOCK, 4 * num_cus());
    const bool vec = (reinterpret_cast<uintptr_t>(keys) % 16) == 0 &&
                     (reinterpret_cast<uintptr_t>(vals) % 16) == 0 &&
                     (v == nullptr || (reinterpret_cast<uintptr_t>(v) % 4) == 0) &&
                     (kv == nullptr || (reinterpret_cast<uintptr_t>(kv) % 4) == 0);
    if (vec)
      hipLaunchKernelGGL(k_hash_acc<true>, dim3(grid), dim3(HA_BLOCK), 0, s, keys, vals, v, kv,
                         n, tab, slots - 1, table, hot_keys, hot_n, flags);
    else
      hipLaunchKernelGGL(k_hash_acc<false>, dim3(grid), dim3(HA_BLOCK), 0, s, keys, vals, v, kv,
                         n, tab, slots - 1, table, hot_keys, hot_n, flags);
    NFH_LAUNCH_CHECK();
  }
  int32_t* ek = scratch_t<int32_t>(S_WORK1, slots);
  uint32_t* es = scratch_t<uint32_t>(S_WORK2, slots);
  scan_tiles(HashExtractOp{tab, ek, es, flags + 1}, slots, s);
  uint32_t* h = static_cast<uint32_t*>(pinned(8));
  NFH_CHECK(hipMemcpyAsync(h, flags, 8, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  if (h[0] != 0) return -1;  // overflow: the caller recomputes
  const uint32_t ng = h[1];
  if (ng == 0) return 0;
  int32_t* perm = scratch_t<int32_t>(S_T0, ng);
  uint32_t* t0 = scratch_t<uint32_t>(S_T1, ng);
  uint32_t* t1 = scratch_t<uint32_t>(S_T2, ng);
  int32_t* tp = scratch_t<int32_t>(S_T3, ng);
  radix_argsort<int32_t>(ek, ng, false, perm, t0, t1, tp, nullptr, s);
  {
    ProfScope prof("hash_emit", s);
    hipLaunchKernelGGL(k_hash_emit, dim3(grid_for(ng, 256, 16 * num_cus())), dim3(256), 0, s, tab,
                       es, perm, ng, o);
    NFH_LAUNCH_CHECK();
  }
  return int64_t(ng);
}

}  // namespace nfh
