// batch.hip -- batched group-by: many independent (keys, values) queries,
// each a dense sum + count over its own key range [0, card_q), in one
// persistent launch set that a plan captures into a hipGraph.
//
// Per query the semantics are those of the dense group-by lane
// (numfast-native/src/groupby/dense.rs:18-37, lib.rs:85-103: sums and
// counts by key code, -2 on a key outside [0, g)), with exact int64 sums
// of int32 values, and the result compacted like the carry build
// (core/compact.rs:16-37): the present groups, ascending, as (ukey, count,
// sum), with per-query group offsets.
//
// Aggregates are packed: one u64 per key, bits [40, 64) the row count,
// [0, 40) the sum of (v + 2^15).  A row adds (1 << 40) + (v + 2^15) -- one
// add for count and sum, exact while a query has fewer than 2^24 rows
// (checked when the plan is built).  A value outside [-2^15, 2^15) adds
// (1 << 40) + 2^15 (as if 0) and goes to an overflow list (query, key,
// value) whose values join the compacted sums at the end.
//
// Two classes of query:
//   small (card <= BT_LDS_SLOTS): work items (query, row range) pulled by
//     persistent workgroups aggregate in LDS and flush into the query's
//     slots (one packed u64 per key, zeroed per run) -- the rows-per-item
//     and grid of this pass are autotuned;
//   ranged (card > BT_LDS_SLOTS): the key range is cut into units of
//     BT_RANGE keys.  A histogram pass counts each 8192-row tile's rows per
//     unit (one entry per (tile, unit), laid out unit-major within the
//     query), a scan of the entries gives every (tile, unit) its place, and
//     a partition pass writes each row as a 32-bit record (key within its
//     unit : biased value) there, so each unit's records are contiguous.
//     No slot array and no device-scope atomics: the card-sized dense slots
//     of the first form (3.6e8 slots, zero-filled and scanned for 2.7e8
//     rows, with a global atomic per row of every wide query) took 13.7 ms;
//     records left in per-tile segments (read through one entry per tile)
//     made the unit passes latency-bound (2.5 ms).
// Output: every query has at least one unit (a small query one, a ranged
// query ceil(card / BT_RANGE)); a count pass gives each unit's present
// groups (small: present slots; ranged: an LDS presence bitmap over its
// records), one scan gives unit and query offsets, and an emit pass writes
// each unit's groups in key order (ranged units aggregate their records in
// LDS first).  The plain form (mode 1) launches the same kernels query by
// query, chaining the group offset on the device.
#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "groupagg.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

typedef unsigned long long u64;

constexpr int BT_BLOCK = 256;
constexpr int BT_LDS_SLOTS = 4096;  // small queries: up to this card
constexpr int BT_SHIFT = 40;
constexpr u64 BT_MASK = (u64(1) << BT_SHIFT) - 1;
constexpr uint32_t BT_MAX_ROWS = 1u << 24;  // rows per query for the packed aggregates
constexpr int BT_RANGE_BITS = 12;
constexpr uint32_t BT_RANGE = 1u << BT_RANGE_BITS;  // keys per ranged unit (32 KB of LDS)
constexpr uint32_t BT_MAX_RANGES = 1u << (24 - BT_RANGE_BITS);  // card <= 2^24
constexpr int BT_PBLOCK = 512;                      // partition pass
constexpr int BT_PIPT = 16;
constexpr uint32_t BT_PTILE = uint32_t(BT_PBLOCK) * BT_PIPT;  // 8192 rows
constexpr int BT_CBLOCK = 256;   // count pass
constexpr int BT_EBLOCK = 256;   // emit pass (five workgroups per CU)
static_assert(uint32_t(BT_LDS_SLOTS) <= BT_RANGE, "a small query's slots are staged in LDS");
constexpr int BT_SCAN = 1024;    // unit scan (one workgroup)
// A ranged unit with more than BT_SLICE records is cut into slices that
// aggregate on separate workgroups (partial presence bitmaps / packed
// aggregates in staging, merged per unit by a second launch): one unit of
// 6e5 records on one workgroup, hot keys and all, set the length of the
// whole pass.
constexpr uint32_t BT_SLICE = 32768;

struct BtItem {  // accumulate work item of a small query
  uint32_t q;
  uint32_t r0, r1;  // row range (global row indices)
  uint32_t whole;   // 1: the item holds all of the query's rows
};
struct BtTile {  // partition tile of a ranged query
  uint32_t q;
  uint32_t r0, r1;  // row range
  uint32_t p0;      // entry of (this tile, unit 0); unit j's entry is p0 + j * nt
  uint32_t nt;      // tiles of the query
  uint32_t pad;
};
struct BtQuery {
  u64 slot0;      // small: first slot of the query
  uint32_t card;
  uint32_t nr;    // ranged: units of the query; 0: small
};
struct BtUnit {  // keys [key0, key0 + nkeys) of query q, in output order
  uint32_t q, j;
  uint32_t key0, nkeys;
  uint32_t p0, nt;  // ranged: entries [p0, p0 + nt) (its records follow off[p0] .. off[p0 + nt])
  u64 slot0;        // small: the query's first slot; ranged: ~0
};
struct BtSlice {  // one slice of a unit, everything its passes read first
  u64 slot0;        // small: the query's first slot; ranged: ~0
  u64 pos;          // output position of the unit's first group (one-slice units)
  uint32_t u;       // the unit
  uint32_t key0, nkeys;
  uint32_t a, e;    // ranged: records [a, e)
  uint32_t stage;   // slice of a split unit: its staging slot
  uint32_t split;   // 1: the unit has several slices
  uint32_t pad;
};
struct BtOver {  // a row whose value is outside [-2^15, 2^15)
  uint32_t q, key;
  int32_t v;
  uint32_t pad;
};
// device counters
struct BtMeta {
  uint32_t work;     // next work item
  uint32_t nmulti;   // units of several slices among the units in flight
  uint32_t nover;    // overflow entries
  uint32_t bad_row;  // first row with a key out of range (0xFFFFFFFF: none)
  u64 base;          // group offset before the first unit (0 when batched)
  u64 total;         // groups written
  uint32_t nslices;  // slices of the units in flight
  uint32_t pad;
};

__device__ __forceinline__ uint32_t bt_biased(int32_t v, bool* over) {
  const uint32_t b = uint32_t(v) + 32768u;
  *over = b >= 65536u;
  return *over ? 32768u : b;
}

// Small queries: persistent accumulation over work items [item0, item1),
// sixteen rows in flight per thread, LDS aggregation, one flush per item.
__global__ __launch_bounds__(BT_BLOCK) void k_bt_accumulate(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const BtItem* __restrict__ items, uint32_t item0, uint32_t item1,
    const BtQuery* __restrict__ qs, u64* __restrict__ slots, BtOver* __restrict__ over,
    BtMeta* __restrict__ meta) {
  constexpr int U = 16;
  __shared__ u64 s_acc[BT_LDS_SLOTS];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = item0 + atomicAdd(&meta->work, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    __syncthreads();
    if (it >= item1) break;
    const BtItem w = items[it];
    const BtQuery qd = qs[w.q];
    const uint32_t card = qd.card;
    // R copies of each slot (lane l adds into copy l % R): a query of few
    // keys would otherwise send most lanes of a wave to one LDS word
    uint32_t R = 1;
    while (R < 32 && card * (2 * R) <= uint32_t(BT_LDS_SLOTS)) R *= 2;
    const uint32_t rl = uint32_t(lane_id()) & (R - 1);
    for (uint32_t s = threadIdx.x; s < card * R; s += BT_BLOCK) s_acc[s] = 0;
    __syncthreads();
    for (uint32_t r = w.r0 + threadIdx.x; r < w.r1; r += U * BT_BLOCK) {
      int32_t k[U], x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t rr = r + uint32_t(u) * BT_BLOCK;
        k[u] = rr < w.r1 ? keys[rr] : 0;
        x[u] = rr < w.r1 ? vals[rr] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t rr = r + uint32_t(u) * BT_BLOCK;
        if (rr >= w.r1) continue;
        if (uint32_t(k[u]) >= card) {
          atomicMin(&meta->bad_row, rr);
          continue;
        }
        bool ov;
        const uint32_t b = bt_biased(x[u], &ov);
        atomicAdd(&s_acc[uint32_t(k[u]) * R + rl], (u64(1) << BT_SHIFT) + u64(b));
        if (ov) over[atomicAdd(&meta->nover, 1u)] = BtOver{w.q, uint32_t(k[u]), x[u], 0u};
      }
    }
    __syncthreads();
    // an item that holds the whole query stores its slots; the items of a
    // larger query add into them (device atomics: the cost that sets the
    // autotuned item size)
    u64* qslots = slots + qd.slot0;
    for (uint32_t s = threadIdx.x; s < card; s += BT_BLOCK) {
      u64 a = 0;
      for (uint32_t r = 0; r < R; ++r) a += s_acc[s * R + r];
      if (w.whole)
        qslots[s] = a;
      else if (a)
        atomicAdd(&qslots[s], a);
    }
    __syncthreads();
  }
}

// Ranged queries, histogram: one workgroup per tile tiles[t_first +
// blockIdx.x] counts its rows per unit into ecnt[p0 + j * nt].  A key
// outside [0, card) is not counted and marks the run.
__global__ __launch_bounds__(BT_PBLOCK) void k_bt_hist(const int32_t* __restrict__ keys,
                                                       const BtTile* __restrict__ tiles,
                                                       uint32_t t_first,
                                                       const BtQuery* __restrict__ qs,
                                                       uint32_t* __restrict__ ecnt,
                                                       BtMeta* __restrict__ meta) {
  __shared__ uint32_t s_cnt[BT_MAX_RANGES];
  const BtTile t = tiles[t_first + blockIdx.x];
  const BtQuery qd = qs[t.q];
  const uint32_t nr = qd.nr, card = qd.card;
  for (uint32_t j = threadIdx.x; j < nr; j += BT_PBLOCK) s_cnt[j] = 0;
  __syncthreads();
  int32_t k[BT_PIPT];
#pragma unroll
  for (int u = 0; u < BT_PIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * BT_PBLOCK + threadIdx.x;
    k[u] = r < t.r1 ? keys[r] : -1;
  }
#pragma unroll
  for (int u = 0; u < BT_PIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * BT_PBLOCK + threadIdx.x;
    if (r >= t.r1) continue;
    if (uint32_t(k[u]) >= card)
      atomicMin(&meta->bad_row, r);
    else
      atomicAdd(&s_cnt[uint32_t(k[u]) >> BT_RANGE_BITS], 1u);
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nr; j += BT_PBLOCK) ecnt[t.p0 + j * t.nt] = s_cnt[j];
}

// Entry scan in three steps over chunks of BT_ECHUNK entries: chunk sums,
// one workgroup's scan of the sums, then each chunk's own scan from its
// base.  off[p] = base + records before entry p; off[np] = base + all
// records.
constexpr uint32_t BT_ECHUNK = uint32_t(BT_SCAN) * 4;
__global__ __launch_bounds__(BT_SCAN) void k_bt_esum(const uint32_t* __restrict__ ecnt,
                                                     uint32_t np, uint32_t* __restrict__ csum) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  const uint32_t b = blockIdx.x * BT_ECHUNK;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t p = b + threadIdx.x + uint32_t(i) * BT_SCAN;
    c += p < np ? ecnt[p] : 0u;
  }
  uint32_t tot;
  (void)block_excl_scan<BT_SCAN / WAVE>(c, s_tmp, &tot);
  if (threadIdx.x == 0) csum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(BT_SCAN) void k_bt_etop(uint32_t* __restrict__ csum, uint32_t nc,
                                                     uint32_t* __restrict__ off, uint32_t np,
                                                     uint32_t base) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = base;
  __syncthreads();
  for (uint32_t b = 0; b < nc; b += BT_SCAN) {
    const uint32_t c = b + threadIdx.x;
    const uint32_t v = c < nc ? csum[c] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<BT_SCAN / WAVE>(v, s_tmp, &tot);
    if (c < nc) csum[c] = s_carry + ex;  // in place: chunk base
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) off[np] = s_carry;
}
__global__ __launch_bounds__(BT_SCAN) void k_bt_escan(const uint32_t* __restrict__ ecnt,
                                                      uint32_t np,
                                                      const uint32_t* __restrict__ cbase,
                                                      uint32_t* __restrict__ off) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  const uint32_t b = blockIdx.x * BT_ECHUNK + threadIdx.x * 4;  // blocked: 4 per thread
  uint32_t c[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c[i] = b + i < np ? ecnt[b + i] : 0u;
    sum += c[i];
  }
  uint32_t tot;
  uint32_t run = cbase[blockIdx.x] + block_excl_scan<BT_SCAN / WAVE>(sum, s_tmp, &tot);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (b + i < np) off[b + i] = run;
    run += c[i];
  }
}

// Ranged queries, partition: one workgroup per tile.  Each row becomes the
// record (key & (BT_RANGE - 1)) << 16 | biased value at off[p0 + j * nt] +
// its rank among the tile's rows of unit j (LDS counters: in no particular
// order, which aggregation does not depend on).
__global__ __launch_bounds__(BT_PBLOCK) void k_bt_part(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const BtTile* __restrict__ tiles, uint32_t t_first, const BtQuery* __restrict__ qs,
    const uint32_t* __restrict__ off, uint32_t* __restrict__ rec, BtOver* __restrict__ over,
    BtMeta* __restrict__ meta) {
  __shared__ uint32_t s_cnt[BT_MAX_RANGES];
  const BtTile t = tiles[t_first + blockIdx.x];
  const BtQuery qd = qs[t.q];
  const uint32_t nr = qd.nr, card = qd.card;
  for (uint32_t j = threadIdx.x; j < nr; j += BT_PBLOCK) s_cnt[j] = 0;
  __syncthreads();
  int32_t k[BT_PIPT], x[BT_PIPT];
#pragma unroll
  for (int u = 0; u < BT_PIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * BT_PBLOCK + threadIdx.x;
    k[u] = r < t.r1 ? keys[r] : -1;
    x[u] = r < t.r1 ? vals[r] : 0;
  }
  uint32_t rc[BT_PIPT], loc[BT_PIPT];
#pragma unroll
  for (int u = 0; u < BT_PIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * BT_PBLOCK + threadIdx.x;
    rc[u] = 0xFFFFFFFFu;  // no record
    if (r >= t.r1 || uint32_t(k[u]) >= card) continue;
    bool ov;
    const uint32_t b = bt_biased(x[u], &ov);
    if (ov) over[atomicAdd(&meta->nover, 1u)] = BtOver{t.q, uint32_t(k[u]), x[u], 0u};
    rc[u] = ((uint32_t(k[u]) & (BT_RANGE - 1)) << 16) | b;
    loc[u] = atomicAdd(&s_cnt[uint32_t(k[u]) >> BT_RANGE_BITS], 1u);
  }
  // the unit bases of this tile
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nr; j += BT_PBLOCK) s_cnt[j] = off[t.p0 + j * t.nt];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BT_PIPT; ++u)
    if (rc[u] != 0xFFFFFFFFu) rec[s_cnt[uint32_t(k[u]) >> BT_RANGE_BITS] + loc[u]] = rc[u];
}

// The records [a, e) of a ranged unit (contiguous), U per thread in
// flight: f(record) for each.
template <int NT, int U, class F>
__device__ __forceinline__ void bt_for_records(uint32_t a, uint32_t e,
                                               const uint32_t* __restrict__ rec, F f) {
  for (uint32_t i = a + threadIdx.x; i < e; i += U * NT) {
    uint32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t ii = i + uint32_t(u) * NT;
      r[u] = ii < e ? __builtin_nontemporal_load(rec + ii) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] != 0xFFFFFFFFu) f(r[u]);
  }
}

// Slices of units [u0, u1): a small unit or a ranged unit of up to
// BT_SLICE records is one slice, a larger one ceil(records / BT_SLICE).
// Two launches over chunks of BT_SCAN units: k_bt_slice_sums leaves each
// chunk's totals (slices, staging slots, split units) in csl[3 * c ..];
// k_bt_slice_scan scans within its chunk from the sum of the chunks before
// it: sstart[u] = first slice (from 0), sstart[u1] = all slices, desc[sl]
// = slice sl's descriptor (its output position is filled in by the unit
// scan), pstart[u] = first staging slot of a unit of several slices,
// multi[] = those units in order.
struct BtUnitSpan {
  uint32_t S, a, e;
};
__device__ __forceinline__ BtUnitSpan bt_unit_span(const BtUnit& un,
                                                   const uint32_t* __restrict__ off) {
  if (un.slot0 != ~0ull) return {1u, 0u, 0u};
  const uint32_t a = off[un.p0], e = off[un.p0 + un.nt];
  const uint32_t R = e - a;
  return {R > BT_SLICE ? (R + BT_SLICE - 1) / BT_SLICE : 1u, a, e};
}
__global__ __launch_bounds__(BT_SCAN) void k_bt_slice_sums(const BtUnit* __restrict__ units,
                                                           uint32_t u0, uint32_t u1,
                                                           const uint32_t* __restrict__ off,
                                                           uint32_t* __restrict__ csl) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  const uint32_t u = u0 + blockIdx.x * BT_SCAN + threadIdx.x;
  const uint32_t S = u < u1 ? bt_unit_span(units[u], off).S : 0u;
  uint32_t ts, tp, tm;
  (void)block_excl_scan<BT_SCAN / WAVE>(S, s_tmp, &ts);
  (void)block_excl_scan<BT_SCAN / WAVE>(S > 1 ? S : 0u, s_tmp, &tp);
  (void)block_excl_scan<BT_SCAN / WAVE>(S > 1 ? 1u : 0u, s_tmp, &tm);
  if (threadIdx.x == 0) {
    csl[3 * blockIdx.x] = ts;
    csl[3 * blockIdx.x + 1] = tp;
    csl[3 * blockIdx.x + 2] = tm;
  }
}
__global__ __launch_bounds__(BT_SCAN) void k_bt_slice_scan(
    const BtUnit* __restrict__ units, uint32_t u0, uint32_t u1, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ csl, uint32_t* __restrict__ sstart,
    uint32_t* __restrict__ pstart, BtSlice* __restrict__ desc, uint32_t* __restrict__ multi,
    BtMeta* __restrict__ meta) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  __shared__ uint32_t s_base[3];
  if (threadIdx.x < 3) {
    uint32_t b = 0;
    for (uint32_t c = 0; c < blockIdx.x; ++c) b += csl[3 * c + threadIdx.x];
    s_base[threadIdx.x] = b;
  }
  __syncthreads();
  const uint32_t u = u0 + blockIdx.x * BT_SCAN + threadIdx.x;
  BtUnit un{};
  BtUnitSpan sp{0u, 0u, 0u};
  if (u < u1) {
    un = units[u];
    sp = bt_unit_span(un, off);
  }
  const uint32_t S = sp.S, P = S > 1 ? S : 0u, M = S > 1 ? 1u : 0u;
  uint32_t ts, tp, tm;
  const uint32_t xs = s_base[0] + block_excl_scan<BT_SCAN / WAVE>(S, s_tmp, &ts);
  const uint32_t xp = s_base[1] + block_excl_scan<BT_SCAN / WAVE>(P, s_tmp, &tp);
  const uint32_t xm = s_base[2] + block_excl_scan<BT_SCAN / WAVE>(M, s_tmp, &tm);
  if (u < u1) {
    sstart[u] = xs;
    pstart[u] = xp;
    for (uint32_t i = 0; i < S; ++i) {
      BtSlice d;
      d.slot0 = un.slot0;
      d.pos = 0;
      d.u = u;
      d.key0 = un.key0;
      d.nkeys = un.nkeys;
      d.a = S > 1 ? sp.a + i * BT_SLICE : sp.a;
      d.e = S > 1 ? min(sp.a + (i + 1) * BT_SLICE, sp.e) : sp.e;
      d.stage = S > 1 ? xp + i : 0u;
      d.split = S > 1 ? 1u : 0u;
      d.pad = 0;
      desc[xs + i] = d;
    }
    if (M) multi[xm] = u;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    sstart[u1] = s_base[0] + ts;
    meta->nslices = s_base[0] + ts;
    meta->nmulti = s_base[2] + tm;
  }
}

// Count pass, one workgroup per slice (workgroups past the last slice
// exit): present groups of a one-slice unit go to ucount (small units count
// their present slots, ranged units the distinct keys of their records); a
// slice of a larger unit leaves its presence bitmap in cstage.
__global__ __launch_bounds__(BT_CBLOCK) void k_bt_count(
    const BtSlice* __restrict__ desc, const u64* __restrict__ slots,
    const uint32_t* __restrict__ rec, uint32_t* __restrict__ cstage,
    uint32_t* __restrict__ ucount, const BtMeta* __restrict__ meta) {
  constexpr uint32_t W = BT_RANGE / 32;
  __shared__ uint32_t s_bits[W];
  __shared__ uint32_t s_tmp[BT_CBLOCK / WAVE];
  const uint32_t sl = blockIdx.x;
  if (sl >= meta->nslices) return;  // uniform
  const BtSlice d = desc[sl];
  uint32_t c = 0;
  if (d.slot0 != ~0ull) {
    for (uint32_t k = threadIdx.x; k < d.nkeys; k += BT_CBLOCK)
      c += (slots[d.slot0 + k] >> BT_SHIFT) != 0 ? 1u : 0u;
  } else {
    for (uint32_t w = threadIdx.x; w < W; w += BT_CBLOCK) s_bits[w] = 0;
    __syncthreads();
    bt_for_records<BT_CBLOCK, 16>(d.a, d.e, rec, [&](uint32_t r) {
      const uint32_t kl = r >> 16;
      atomicOr(&s_bits[kl >> 5], 1u << (kl & 31u));
    });
    __syncthreads();
    if (d.split) {
      uint32_t* dst = cstage + size_t(d.stage) * W;
      for (uint32_t w = threadIdx.x; w < W; w += BT_CBLOCK) dst[w] = s_bits[w];
      return;  // uniform
    }
    for (uint32_t w = threadIdx.x; w < W; w += BT_CBLOCK) c += __popc(s_bits[w]);
  }
  uint32_t tot;
  (void)block_excl_scan<BT_CBLOCK / WAVE>(c, s_tmp, &tot);
  if (threadIdx.x == 0) ucount[d.u] = tot;
}

// Units of several slices (one workgroup each, past the last: exit): the
// OR of their slices' bitmaps.
__global__ __launch_bounds__(BT_CBLOCK) void k_bt_cmerge(
    const uint32_t* __restrict__ multi, const uint32_t* __restrict__ sstart,
    const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ cstage,
    uint32_t* __restrict__ ucount, const BtMeta* __restrict__ meta) {
  constexpr uint32_t W = BT_RANGE / 32;
  __shared__ uint32_t s_tmp[BT_CBLOCK / WAVE];
  if (blockIdx.x < meta->nmulti) {  // uniform
    const uint32_t u = multi[blockIdx.x];
    const uint32_t S = sstart[u + 1] - sstart[u];
    const uint32_t* src = cstage + size_t(pstart[u]) * W;
    uint32_t c = 0;
    for (uint32_t w = threadIdx.x; w < W; w += BT_CBLOCK) {
      uint32_t x = 0;
      for (uint32_t i = 0; i < S; ++i) x |= src[size_t(i) * W + w];
      c += __popc(x);
    }
    uint32_t tot;
    (void)block_excl_scan<BT_CBLOCK / WAVE>(c, s_tmp, &tot);
    if (threadIdx.x == 0) ucount[u] = tot;
  }
}

// Exclusive scan of the unit counts of [u0, u1) from meta->base, over
// chunks of BT_SCAN units: k_bt_usum leaves each chunk's total in csl[c];
// k_bt_uscan scans its chunk from the sum of the chunks before it: uoff per
// unit, the output position in the descriptor of a one-slice unit,
// group_off[q] at each query's first unit, meta->total and
// group_off[nq_end] at the end.
__global__ __launch_bounds__(BT_SCAN) void k_bt_usum(uint32_t u0, uint32_t u1,
                                                     const uint32_t* __restrict__ ucount,
                                                     uint32_t* __restrict__ csl) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  const uint32_t u = u0 + blockIdx.x * BT_SCAN + threadIdx.x;
  uint32_t tot;
  (void)block_excl_scan<BT_SCAN / WAVE>(u < u1 ? ucount[u] : 0u, s_tmp, &tot);
  if (threadIdx.x == 0) csl[blockIdx.x] = tot;
}
__global__ __launch_bounds__(BT_SCAN) void k_bt_uscan(
    const BtUnit* __restrict__ units, uint32_t u0, uint32_t u1,
    const uint32_t* __restrict__ ucount, const uint32_t* __restrict__ csl,
    const uint32_t* __restrict__ sstart, BtSlice* __restrict__ desc, u64* __restrict__ uoff,
    int64_t* __restrict__ group_off, uint32_t nq_end, BtMeta* __restrict__ meta) {
  __shared__ uint32_t s_tmp[BT_SCAN / WAVE];
  __shared__ u64 s_base;
  if (threadIdx.x == 0) {
    u64 b = meta->base;
    for (uint32_t c = 0; c < blockIdx.x; ++c) b += csl[c];
    s_base = b;
  }
  __syncthreads();
  const uint32_t u = u0 + blockIdx.x * BT_SCAN + threadIdx.x;
  const uint32_t c = u < u1 ? ucount[u] : 0u;
  uint32_t tot;
  const u64 run = s_base + block_excl_scan<BT_SCAN / WAVE>(c, s_tmp, &tot);
  if (u < u1) {
    uoff[u] = run;
    if (units[u].j == 0) group_off[units[u].q] = int64_t(run);
    const uint32_t s0 = sstart[u];
    if (sstart[u + 1] - s0 == 1) desc[s0].pos = run;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    meta->total = s_base + tot;
    group_off[nq_end] = int64_t(s_base + tot);
  }
}

// Groups of unit `un` in key order from pos (positions at or past cap are
// not written); packed(k) is key k's aggregate.  Small units leave each
// present slot's output position in it (for the overflow pass).  Each wave
// compacts one contiguous segment of the unit's keys by ballot.
template <class F>
__device__ __forceinline__ void bt_compact(u64 slot0, uint32_t key0, u64 pos, F packed,
                                           uint32_t* s_wc, u64* __restrict__ slots,
                                           int64_t* __restrict__ ukeys,
                                           int64_t* __restrict__ counts,
                                           int64_t* __restrict__ sums, u64 cap) {
  // wave w holds keys [w * PER * 64, (w + 1) * PER * 64), all PER reads in
  // flight at once and kept in registers for both passes
  constexpr int PER = int(BT_RANGE) / BT_EBLOCK;
  const bool small = slot0 != ~0ull;
  const uint32_t wave = uint32_t(wave_id()), lane = uint32_t(lane_id());
  const uint32_t k0 = wave * uint32_t(PER * WAVE);
  u64 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) v[j] = packed(k0 + uint32_t(j * WAVE) + lane);
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) c += uint32_t(__popcll(__ballot((v[j] >> BT_SHIFT) != 0)));
  if (lane == 0) s_wc[wave] = c;
  __syncthreads();
  for (uint32_t w = 0; w < wave; ++w) pos += s_wc[w];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t key = k0 + uint32_t(j * WAVE) + lane;
    const u64 pk = v[j];
    const u64 cnt = pk >> BT_SHIFT;
    const uint64_t b = __ballot(cnt != 0);
    if (cnt != 0) {
      const u64 p = pos + u64(__popcll(b & lt));
      if (p < cap) {
        ukeys[p] = int64_t(key0 + key);
        counts[p] = int64_t(cnt);
        sums[p] = int64_t(pk & BT_MASK) - int64_t(cnt) * 32768;
      }
      if (small) slots[slot0 + key] = p;
    }
    pos += u64(__popcll(b));
  }
  __syncthreads();  // s_wc is reused by the caller's next unit
}

// Emit pass, one workgroup per slice (past the last: exit): a one-slice
// unit is aggregated (ranged: its records in LDS) and written at its
// offset; a slice of a larger unit leaves its packed aggregates in estage.
__global__ __launch_bounds__(BT_EBLOCK) void k_bt_emit(
    const BtSlice* __restrict__ desc, u64* __restrict__ slots, const uint32_t* __restrict__ rec,
    u64* __restrict__ estage, int64_t* __restrict__ ukeys, int64_t* __restrict__ counts,
    int64_t* __restrict__ sums, u64 cap, const BtMeta* __restrict__ meta) {
  __shared__ u64 s_acc[BT_RANGE];
  __shared__ uint32_t s_wc[BT_EBLOCK / WAVE];
  const uint32_t sl = blockIdx.x;
  if (sl >= meta->nslices) return;  // uniform
  const BtSlice d = desc[sl];
  if (d.slot0 != ~0ull) {
    // staged in LDS: the compaction then reads LDS only (a load that may be
    // global or LDS compiles to a flat load, whose wait also waits for the
    // group stores in flight)
    for (uint32_t k = threadIdx.x; k < d.nkeys; k += BT_EBLOCK) s_acc[k] = slots[d.slot0 + k];
    __syncthreads();
  } else {
    for (uint32_t k = threadIdx.x; k < d.nkeys; k += BT_EBLOCK) s_acc[k] = 0;
    __syncthreads();
    bt_for_records<BT_EBLOCK, 8>(d.a, d.e, rec, [&](uint32_t r) {
      atomicAdd(&s_acc[r >> 16], (u64(1) << BT_SHIFT) + u64(r & 0xFFFFu));
    });
    __syncthreads();
    if (d.split) {
      u64* dst = estage + size_t(d.stage) * BT_RANGE;
      for (uint32_t k = threadIdx.x; k < d.nkeys; k += BT_EBLOCK) dst[k] = s_acc[k];
      return;  // uniform
    }
  }
  bt_compact(
      d.slot0, d.key0, d.pos, [&](uint32_t k) -> u64 { return k < d.nkeys ? s_acc[k] : 0ull; },
      s_wc, slots, ukeys, counts, sums, cap);
}

// Units of several slices: the sum of their slices' packed aggregates,
// then the groups in key order.
__global__ __launch_bounds__(BT_EBLOCK) void k_bt_emerge(
    const BtUnit* __restrict__ units, const uint32_t* __restrict__ multi,
    const uint32_t* __restrict__ sstart, const uint32_t* __restrict__ pstart,
    const u64* __restrict__ uoff, const u64* __restrict__ estage, u64* __restrict__ slots,
    int64_t* __restrict__ ukeys, int64_t* __restrict__ counts, int64_t* __restrict__ sums,
    u64 cap, const BtMeta* __restrict__ meta) {
  __shared__ u64 s_acc[BT_RANGE];
  __shared__ uint32_t s_wc[BT_EBLOCK / WAVE];
  if (blockIdx.x < meta->nmulti) {  // uniform
    const uint32_t u = multi[blockIdx.x];
    const uint32_t S = sstart[u + 1] - sstart[u];
    const BtUnit un = units[u];
    const u64* src = estage + size_t(pstart[u]) * BT_RANGE;
    for (uint32_t k = threadIdx.x; k < un.nkeys; k += BT_EBLOCK) {
      u64 x = 0;
      for (uint32_t i = 0; i < S; ++i) x += src[size_t(i) * BT_RANGE + k];
      s_acc[k] = x;
    }
    __syncthreads();
    bt_compact(
        un.slot0, un.key0, uoff[u], [&](uint32_t k) -> u64 { return k < un.nkeys ? s_acc[k] : 0ull; },
        s_wc, slots, ukeys, counts, sums, cap);
  }
}

// Overflow rows: their true values join the compacted sums.  Small
// queries find the output position in the slot; ranged queries search
// their (ascending) output keys.
__global__ void k_bt_overflow(const BtOver* __restrict__ over, const BtMeta* __restrict__ meta,
                              const BtQuery* __restrict__ qs, const u64* __restrict__ slots,
                              const int64_t* __restrict__ group_off,
                              const int64_t* __restrict__ ukeys, int64_t* __restrict__ sums,
                              u64 cap) {
  const uint32_t n = meta->nover;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const BtOver e = over[i];
    const BtQuery qd = qs[e.q];
    u64 pos = ~0ull;
    if (qd.nr == 0) {
      pos = slots[qd.slot0 + e.key];
    } else {
      u64 lo = u64(group_off[e.q]), hi = std::min<u64>(u64(group_off[e.q + 1]), cap);
      while (lo < hi) {  // first output key >= e.key
        const u64 mid = (lo + hi) / 2;
        if (ukeys[mid] < int64_t(e.key))
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < std::min<u64>(u64(group_off[e.q + 1]), cap) && ukeys[lo] == int64_t(e.key))
        pos = lo;
    }
    if (pos < cap) atomicAdd(reinterpret_cast<u64*>(sums + pos), u64(int64_t(e.v)));
  }
}

// meta reset for a batched run: counters 0, bad_row none, base 0.
__global__ void k_bt_reset(BtMeta* meta) {
  meta->work = 0;
  meta->nover = 0;
  meta->bad_row = 0xFFFFFFFFu;
  meta->base = 0;
  meta->total = 0;
}
// per-query chain: work counter 0, base = previous query's total.
__global__ void k_bt_next(BtMeta* meta) {
  meta->work = 0;
  meta->base = meta->total;
}

// Mode 2's range check: a group key outside [0, card_q) of its query marks
// the run bad (the per-query group-by accepts any key).
__global__ void k_bt_check_keys(const int64_t* __restrict__ group_off,
                                const int64_t* __restrict__ ukeys,
                                const BtQuery* __restrict__ qs, uint32_t nq,
                                BtMeta* __restrict__ meta) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    const int64_t a = group_off[q], b = group_off[q + 1];
    if (b > a && (ukeys[a] < 0 || ukeys[b - 1] >= int64_t(qs[q].card)))
      atomicMin(&meta->bad_row, 0u);
  }
}

// ------------------------------------------------------------ plan
struct BatchPlan {
  const int32_t* keys = nullptr;
  const int32_t* vals = nullptr;
  int64_t* ukeys = nullptr;
  int64_t* counts = nullptr;
  int64_t* sums = nullptr;
  int64_t* group_off = nullptr;
  u64 cap = 0;
  uint32_t nq = 0;
  uint64_t nrows = 0, nslots = 0;
  // per query (+1): first work item, tile and unit
  std::vector<uint32_t> item_start, tile_start, unit_start;
  BtItem* items = nullptr;
  BtTile* tiles = nullptr;
  BtUnit* units = nullptr;
  BtQuery* qs = nullptr;
  u64* slots = nullptr;
  uint32_t* rec = nullptr;
  uint32_t* ecnt = nullptr;  // rows per (tile, unit) entry, unit-major per query
  uint32_t* eoff = nullptr;  // records before each entry (+1)
  uint32_t* csum = nullptr;  // entry-scan chunk sums / bases
  uint32_t* sstart = nullptr;  // per unit (+1): first slice
  uint32_t* pstart = nullptr;  // per unit: first staging slot (units of several slices)
  uint32_t* cstage = nullptr;  // presence bitmaps of slices
  u64* estage = nullptr;       // packed aggregates of slices
  BtSlice* desc = nullptr;         // descriptor of each slice
  uint32_t* multi = nullptr;       // units of several slices
  uint32_t* csl = nullptr;         // slice-scan chunk totals
  std::vector<uint32_t> p_start;  // per query (+1): first entry
  uint32_t* ucount = nullptr;
  u64* uoff = nullptr;
  BtOver* over = nullptr;
  BtMeta* meta = nullptr;
  uint32_t nitems = 0, ntiles = 0, nunits = 0;
  uint64_t nentries = 0;
  uint32_t rows_per_item = 0, grid = 0;
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> owned;
  std::vector<int64_t> roff;  // per query (+1): row offsets (host copy)
  std::vector<int32_t> card;  // per query: key range (host copy)
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : owned) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  template <class T>
  T* alloc(size_t count) {
    void* p = nullptr;
    NFH_CHECK(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256)));
    owned.push_back(p);
    return static_cast<T*>(p);
  }
};

static bool bt_small(int32_t card) { return card <= BT_LDS_SLOTS; }

// Work items of the small queries, R rows each.
static void build_items(BatchPlan& P, uint32_t R) {
  std::vector<BtItem> it;
  P.item_start.assign(P.nq + 1, 0);
  for (uint32_t q = 0; q < P.nq; ++q) {
    P.item_start[q] = uint32_t(it.size());
    if (!bt_small(P.card[q])) continue;
    const bool whole = P.roff[q + 1] - P.roff[q] <= int64_t(R);
    for (int64_t r = P.roff[q]; r < P.roff[q + 1]; r += R)
      it.push_back({q, uint32_t(r), uint32_t(std::min<int64_t>(r + R, P.roff[q + 1])),
                    whole ? 1u : 0u});
  }
  P.item_start[P.nq] = uint32_t(it.size());
  P.nitems = uint32_t(it.size());
  if (!P.items) P.items = P.alloc<BtItem>(P.nrows / 512 + P.nq + 16);  // R >= 512
  if (!it.empty())
    NFH_CHECK(hipMemcpy(P.items, it.data(), it.size() * sizeof(BtItem), hipMemcpyHostToDevice));
  P.rows_per_item = R;
}

// Histogram, entry scan (from base) and partition of tiles [t0, t1), whose
// entries are [p0, p1).
static void enqueue_ranged(BatchPlan& P, uint32_t t0, uint32_t t1, uint32_t p0, uint32_t p1,
                           uint32_t base) {
  if (t1 == t0) return;
  hipLaunchKernelGGL(k_bt_hist, dim3(t1 - t0), dim3(BT_PBLOCK), 0, P.s, P.keys, P.tiles, t0,
                     P.qs, P.ecnt, P.meta);
  NFH_LAUNCH_CHECK();
  const uint32_t np = p1 - p0, nc = (np + BT_ECHUNK - 1) / BT_ECHUNK;
  hipLaunchKernelGGL(k_bt_esum, dim3(nc), dim3(BT_SCAN), 0, P.s, P.ecnt + p0, np, P.csum);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_etop, dim3(1), dim3(BT_SCAN), 0, P.s, P.csum, nc, P.eoff + p0, np,
                     base);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_escan, dim3(nc), dim3(BT_SCAN), 0, P.s, P.ecnt + p0, np, P.csum,
                     P.eoff + p0);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_part, dim3(t1 - t0), dim3(BT_PBLOCK), 0, P.s, P.keys, P.vals,
                     P.tiles, t0, P.qs, P.eoff, P.rec, P.over, P.meta);
  NFH_LAUNCH_CHECK();
}

// Count, scan and emit of units [u0, u1) (their queries hold `rows` rows),
// group offsets up to group_off[nq_end].  Grids are the bounds on slices
// and split units; workgroups past the counts the slice pass leaves exit.
static void enqueue_units(BatchPlan& P, uint32_t u0, uint32_t u1, uint32_t nq_end,
                          uint64_t rows) {
  const uint32_t max_sl = uint32_t(uint64_t(u1 - u0) + rows / BT_SLICE + 1);
  const uint32_t max_multi = uint32_t(rows / BT_SLICE + 1);
  const uint32_t nch = (u1 - u0 + BT_SCAN - 1) / BT_SCAN;
  hipLaunchKernelGGL(k_bt_slice_sums, dim3(nch), dim3(BT_SCAN), 0, P.s, P.units, u0, u1, P.eoff,
                     P.csl);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_slice_scan, dim3(nch), dim3(BT_SCAN), 0, P.s, P.units, u0, u1, P.eoff,
                     P.csl, P.sstart, P.pstart, P.desc, P.multi, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_count, dim3(max_sl), dim3(BT_CBLOCK), 0, P.s, P.desc, P.slots, P.rec,
                     P.cstage, P.ucount, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_cmerge, dim3(max_multi), dim3(BT_CBLOCK), 0, P.s, P.multi, P.sstart,
                     P.pstart, P.cstage, P.ucount, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_usum, dim3(nch), dim3(BT_SCAN), 0, P.s, u0, u1, P.ucount, P.csl);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_uscan, dim3(nch), dim3(BT_SCAN), 0, P.s, P.units, u0, u1, P.ucount,
                     P.csl, P.sstart, P.desc, P.uoff, P.group_off, nq_end, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_emit, dim3(max_sl), dim3(BT_EBLOCK), 0, P.s, P.desc, P.slots, P.rec,
                     P.estage, P.ukeys, P.counts, P.sums, P.cap, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_emerge, dim3(max_multi), dim3(BT_EBLOCK), 0, P.s, P.units, P.multi,
                     P.sstart, P.pstart, P.uoff, P.estage, P.slots, P.ukeys, P.counts, P.sums,
                     P.cap, P.meta);
  NFH_LAUNCH_CHECK();
}

// The whole batched run on P.s: reset, zero the small queries' slots,
// accumulate (small), partition (ranged), count, scan, emit, overflow.  No
// host synchronisation (graph-capturable).
static void enqueue_batched(BatchPlan& P) {
  hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, P.s, P.meta);
  NFH_LAUNCH_CHECK();
  if (P.nslots) fill_u64(P.slots, 0, P.nslots, P.s);
  if (P.nitems) {
    hipLaunchKernelGGL(k_bt_accumulate, dim3(P.grid), dim3(BT_BLOCK), 0, P.s, P.keys, P.vals,
                       P.items, 0u, P.nitems, P.qs, P.slots, P.over, P.meta);
    NFH_LAUNCH_CHECK();
  }
  enqueue_ranged(P, 0, P.ntiles, 0, uint32_t(P.nentries), 0);
  enqueue_units(P, 0, P.nunits, P.nq, P.nrows);
  hipLaunchKernelGGL(k_bt_overflow, dim3(256), dim3(256), 0, P.s, P.over, P.meta, P.qs, P.slots,
                     P.group_off, P.ukeys, P.sums, P.cap);
  NFH_LAUNCH_CHECK();
}

// The plain form: per query its accumulate or partition launch, then its
// count, scan and emit launches, chained on the device.
static void enqueue_per_query(BatchPlan& P) {
  hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, P.s, P.meta);
  NFH_LAUNCH_CHECK();
  if (P.nslots) fill_u64(P.slots, 0, P.nslots, P.s);
  for (uint32_t q = 0; q < P.nq; ++q) {
    const uint32_t i0 = P.item_start[q], i1 = P.item_start[q + 1];
    const uint32_t t0 = P.tile_start[q], t1 = P.tile_start[q + 1];
    const uint32_t u0 = P.unit_start[q], u1 = P.unit_start[q + 1];
    if (q > 0) {
      hipLaunchKernelGGL(k_bt_next, dim3(1), dim3(1), 0, P.s, P.meta);
      NFH_LAUNCH_CHECK();
    }
    if (i1 > i0) {
      hipLaunchKernelGGL(k_bt_accumulate, dim3(std::min<uint32_t>(i1 - i0, P.grid)),
                         dim3(BT_BLOCK), 0, P.s, P.keys, P.vals, P.items, i0, i1, P.qs, P.slots,
                         P.over, P.meta);
      NFH_LAUNCH_CHECK();
    }
    // this query's records start at its first row
    enqueue_ranged(P, t0, t1, P.p_start[q], P.p_start[q + 1], uint32_t(P.roff[q]));
    enqueue_units(P, u0, u1, q + 1, uint64_t(P.roff[q + 1] - P.roff[q]));
  }
  hipLaunchKernelGGL(k_bt_overflow, dim3(256), dim3(256), 0, P.s, P.over, P.meta, P.qs, P.slots,
                     P.group_off, P.ukeys, P.sums, P.cap);
  NFH_LAUNCH_CHECK();
}

// The plain library path: each query through the library's own group-by
// (nfh_group_agg_i32's device path, one call per query on this thread's
// stream, each returning its group count), outputs at the running group
// offset; then the key-range check of the batch contract.
static int64_t run_library_per_query(BatchPlan& P) {
  hipStream_t cs = stream();
  std::vector<int64_t> goff(P.nq + 1, 0);
  int64_t off = 0;
  for (uint32_t q = 0; q < P.nq; ++q) {
    goff[q] = off;
    const int64_t r0 = P.roff[q], nrow = P.roff[q + 1] - P.roff[q];
    if (nrow == 0) continue;
    if (uint64_t(off) + uint64_t(std::min<int64_t>(nrow, P.card[q])) > P.cap) return TOO_SMALL;
    GroupAggOut o{};
    o.ukeys = P.ukeys + off;
    o.counts = P.counts + off;
    o.sums = P.sums + off;
    off += group_agg_device(P.keys + r0, P.vals + r0, nullptr, nullptr, size_t(nrow), o, cs);
  }
  goff[P.nq] = off;
  NFH_CHECK(hipMemcpyAsync(P.group_off, goff.data(), goff.size() * 8, hipMemcpyHostToDevice,
                           cs));
  hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, cs, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_check_keys, dim3(grid_for(P.nq, 256, 256)), dim3(256), 0, cs,
                     P.group_off, P.ukeys, P.qs, P.nq, P.meta);
  NFH_LAUNCH_CHECK();
  BtMeta m;
  NFH_CHECK(hipMemcpyAsync(&m, P.meta, sizeof(m), hipMemcpyDeviceToHost, cs));
  NFH_CHECK(hipStreamSynchronize(cs));
  return m.bad_row != 0xFFFFFFFFu ? int64_t(BAD_RANGE) : off;
}

static float time_accumulate(BatchPlan& P) {
  hipEvent_t a, b;
  NFH_CHECK(hipEventCreate(&a));
  NFH_CHECK(hipEventCreate(&b));
  float best = 1e30f;
  for (int rep = 0; rep < 2; ++rep) {
    hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, P.s, P.meta);
    if (P.nslots) fill_u64(P.slots, 0, P.nslots, P.s);
    NFH_CHECK(hipEventRecord(a, P.s));
    hipLaunchKernelGGL(k_bt_accumulate, dim3(P.grid), dim3(BT_BLOCK), 0, P.s, P.keys, P.vals,
                       P.items, 0u, P.nitems, P.qs, P.slots, P.over, P.meta);
    NFH_LAUNCH_CHECK();
    NFH_CHECK(hipEventRecord(b, P.s));
    NFH_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    NFH_CHECK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

}  // namespace nfh

using namespace nfh;

// Additive (no reference counterpart): build a batched group-by plan over
// device buffers.  keys / vals: all queries' rows, concatenated; query q
// owns rows [row_off[q], row_off[q+1]) and key range [0, card[q]).  Outputs
// (device, cap lanes): ukeys / counts / sums of the present groups, query
// by query, keys ascending; group_off[nq + 1] the per-query group offsets.
// row_off / card may be host or device arrays (read once here).  The plan
// picks rows per work item and the persistent grid of the small-query pass
// by timing candidates on these buffers (autotuning), then captures the
// whole run into one hipGraph.  Returns a handle or NULL (*rc_out: -1 null,
// -3 a query over 2^24 rows or keys, bad offsets, or buffers beyond the
// free device memory, -6 device).
NF_EXPORT void* nfh_group_batch_plan(const int32_t* keys, const int32_t* vals,
                                     const int64_t* row_off, const int32_t* card, size_t nq,
                                     int64_t* ukeys, int64_t* counts, int64_t* sums,
                                     int64_t* group_off, size_t cap, int32_t* rc_out) {
  auto fail = [&](int32_t rc) -> void* {
    if (rc_out) *rc_out = rc;
    return nullptr;
  };
  if (!keys || !vals || !row_off || !card || !ukeys || !counts || !sums || !group_off || nq == 0)
    return fail(NULL_OR_ABORT);
  try {
    ensure_device();
    (void)lds_rank_supported();
    std::vector<int64_t> roff(nq + 1);
    std::vector<int32_t> cd(nq);
    NFH_CHECK(hipMemcpy(roff.data(), row_off, (nq + 1) * 8, hipMemcpyDefault));
    NFH_CHECK(hipMemcpy(cd.data(), card, nq * 4, hipMemcpyDefault));
    if (roff[0] != 0 || nq >= size_t(INT32_MAX)) return fail(MALFORMED);
    for (size_t q = 0; q < nq; ++q)
      if (roff[q + 1] < roff[q] || roff[q + 1] - roff[q] >= int64_t(BT_MAX_ROWS) || cd[q] < 0 ||
          uint32_t(cd[q]) > BT_MAX_RANGES * BT_RANGE)
        return fail(MALFORMED);
    if (roff[nq] > int64_t(INT32_MAX)) return fail(MALFORMED);
    // queries, slots (small), tiles and entries (ranged), units
    std::vector<BtQuery> qs(nq);
    std::vector<BtTile> tl;
    std::vector<BtUnit> un;
    std::vector<uint32_t> tile_start(nq + 1), unit_start(nq + 1), p_start(nq + 1);
    u64 slot = 0, ent = 0;
    for (size_t q = 0; q < nq; ++q) {
      const uint32_t c = uint32_t(cd[q]);
      tile_start[q] = uint32_t(tl.size());
      unit_start[q] = uint32_t(un.size());
      p_start[q] = uint32_t(std::min<u64>(ent, UINT32_MAX));
      if (bt_small(cd[q])) {
        qs[q] = {slot, c, 0};
        un.push_back({uint32_t(q), 0, 0, c, 0, 0, slot});
        slot += c;
        continue;
      }
      const uint32_t nr = (c + BT_RANGE - 1) / BT_RANGE;
      qs[q] = {0, c, nr};
      const uint64_t nrow = uint64_t(roff[q + 1] - roff[q]);
      const uint32_t nt = uint32_t((nrow + BT_PTILE - 1) / BT_PTILE);
      for (uint32_t t = 0; t < nt; ++t) {
        const int64_t r = roff[q] + int64_t(t) * BT_PTILE;
        tl.push_back({uint32_t(q), uint32_t(r),
                      uint32_t(std::min<int64_t>(r + BT_PTILE, roff[q + 1])),
                      uint32_t(std::min<u64>(ent + t, UINT32_MAX)), nt, 0});
      }
      for (uint32_t j = 0; j < nr; ++j)
        un.push_back({uint32_t(q), j, j * BT_RANGE, std::min(BT_RANGE, c - j * BT_RANGE),
                      uint32_t(std::min<u64>(ent + u64(j) * nt, UINT32_MAX)), nt, ~0ull});
      ent += u64(nt) * nr;
    }
    tile_start[nq] = uint32_t(tl.size());
    unit_start[nq] = uint32_t(un.size());
    p_start[nq] = uint32_t(std::min<u64>(ent, UINT32_MAX));
    if (ent >= uint64_t(UINT32_MAX) || un.size() >= size_t(INT32_MAX)) return fail(MALFORMED);
    // slots (8 B per key of each small query), records (4 B per row), the
    // overflow list (16 B per row) and the entries must fit the free memory
    size_t free_b = 0, total_b = 0;
    NFH_CHECK(hipMemGetInfo(&free_b, &total_b));
    const uint64_t maxp = 2 * uint64_t(roff[nq]) / BT_SLICE + 2;  // slices of split units
    const uint64_t need = slot * 8 + uint64_t(roff[nq]) * 24 + ent * 8 + un.size() * 104 +
                          tl.size() * 24 + maxp * (BT_RANGE * 8 + BT_RANGE / 8);
    if (need > uint64_t(free_b)) return fail(MALFORMED);
    std::unique_ptr<BatchPlan> P(new BatchPlan());
    P->roff = roff;
    P->card = cd;
    P->keys = keys;
    P->vals = vals;
    P->ukeys = ukeys;
    P->counts = counts;
    P->sums = sums;
    P->group_off = group_off;
    P->cap = cap;
    P->nq = uint32_t(nq);
    P->nrows = uint64_t(roff[nq]);
    P->nslots = slot;
    P->ntiles = uint32_t(tl.size());
    P->nunits = uint32_t(un.size());
    P->nentries = ent;
    P->tile_start = tile_start;
    P->unit_start = unit_start;
    P->p_start = p_start;
    NFH_CHECK(hipStreamCreate(&P->s));
    P->qs = P->alloc<BtQuery>(nq);
    P->tiles = P->alloc<BtTile>(tl.size());
    P->units = P->alloc<BtUnit>(un.size());
    P->slots = P->alloc<u64>(P->nslots);
    P->rec = P->alloc<uint32_t>(P->nrows);
    P->ecnt = P->alloc<uint32_t>(ent);
    P->eoff = P->alloc<uint32_t>(ent + 1);
    P->csum = P->alloc<uint32_t>(ent / BT_ECHUNK + 2);
    P->sstart = P->alloc<uint32_t>(un.size() + 1);
    P->pstart = P->alloc<uint32_t>(un.size());
    P->cstage = P->alloc<uint32_t>(maxp * (BT_RANGE / 32));
    P->estage = P->alloc<u64>(maxp * BT_RANGE);
    P->desc = P->alloc<BtSlice>(un.size() + uint64_t(roff[nq]) / BT_SLICE + 2);
    P->multi = P->alloc<uint32_t>(uint64_t(roff[nq]) / BT_SLICE + 2);
    P->csl = P->alloc<uint32_t>(3 * (un.size() / BT_SCAN + 2));
    P->ucount = P->alloc<uint32_t>(un.size());
    P->uoff = P->alloc<u64>(un.size());
    P->over = P->alloc<BtOver>(P->nrows + 1);
    P->meta = P->alloc<BtMeta>(1);
    NFH_CHECK(hipMemcpy(P->qs, qs.data(), nq * sizeof(BtQuery), hipMemcpyHostToDevice));
    if (!tl.empty())
      NFH_CHECK(hipMemcpy(P->tiles, tl.data(), tl.size() * sizeof(BtTile), hipMemcpyHostToDevice));
    NFH_CHECK(hipMemcpy(P->units, un.data(), un.size() * sizeof(BtUnit), hipMemcpyHostToDevice));
    // autotune rows per work item and the persistent grid of the small pass
    uint32_t bestR = 8192, bestG = 4;
    const uint32_t cus = uint32_t(num_cus());
    build_items(*P, bestR);
    if (P->nitems) {
      float best = 1e30f;
      for (uint32_t R : {8192u, 32768u, 131072u, 524288u}) {
        build_items(*P, R);
        for (uint32_t g : {2u, 4u, 8u}) {
          P->grid = cus * g;
          const float ms = time_accumulate(*P);
          if (ms < best) {
            best = ms;
            bestR = R;
            bestG = g;
          }
        }
      }
      build_items(*P, bestR);
    }
    P->grid = cus * bestG;
    // capture the batched run (an error inside ends the capture first)
    NFH_CHECK(hipStreamBeginCapture(P->s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_batched(*P);
    } catch (...) {
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(P->s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      throw;
    }
    NFH_CHECK(hipStreamEndCapture(P->s, &P->graph));
    NFH_CHECK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    if (rc_out) *rc_out = OK;
    return P.release();
  } catch (...) {
    return fail(DEVICE_ERROR);
  }
}

static int64_t batch_result(BatchPlan* P) {
  BtMeta m;
  NFH_CHECK(hipMemcpyAsync(&m, P->meta, sizeof(m), hipMemcpyDeviceToHost, P->s));
  NFH_CHECK(hipStreamSynchronize(P->s));
  if (m.bad_row != 0xFFFFFFFFu) return BAD_RANGE;
  if (m.total > P->cap) return TOO_SMALL;
  return int64_t(m.total);
}

// Run a plan: mode 0 replays the captured graph (one persistent launch of
// the batch), mode 1 launches the same kernels query by query, mode 2
// runs each query through the library's plain group-by (one call per
// query, as a caller without the batch would).  Returns the number of
// groups; -2 when a key is outside its query's range (outputs undefined),
// -3 for an unknown mode, -7 when cap is too small.  Synchronous.
NF_EXPORT int64_t nfh_group_batch_run(void* plan, int32_t mode) {
  if (!plan) return NULL_OR_ABORT;
  if (mode < 0 || mode > 2) return MALFORMED;
  auto* P = static_cast<BatchPlan*>(plan);
  try {
    if (mode == 2) {
      NFH_CHECK(hipStreamSynchronize(P->s));
      return run_library_per_query(*P);
    }
    if (mode == 0)
      NFH_CHECK(hipGraphLaunch(P->exec, P->s));
    else
      enqueue_per_query(*P);
    return batch_result(P);
  } catch (...) {
    return DEVICE_ERROR;
  }
}

// Plan details for reports: rows per work item and workgroups of the
// small-query pass, its work items, the small queries' slots.
NF_EXPORT int32_t nfh_group_batch_info(void* plan, uint64_t* out4) {
  if (!plan || !out4) return NULL_OR_ABORT;
  auto* P = static_cast<BatchPlan*>(plan);
  out4[0] = P->rows_per_item;
  out4[1] = P->grid;
  out4[2] = P->nitems;
  out4[3] = P->nslots;
  return OK;
}

// Ranged-path sizes for reports: partition tiles, units, (tile, unit)
// entries.
NF_EXPORT int32_t nfh_group_batch_info2(void* plan, uint64_t* out3) {
  if (!plan || !out3) return NULL_OR_ABORT;
  auto* P = static_cast<BatchPlan*>(plan);
  out3[0] = P->ntiles;
  out3[1] = P->nunits;
  out3[2] = P->nentries;
  return OK;
}

NF_EXPORT void nfh_group_batch_free(void* plan) {
  auto* P = static_cast<BatchPlan*>(plan);
  if (P) {
    (void)hipStreamSynchronize(P->s);
    delete P;
  }
}
