// batch.hip -- batched group-by: many independent (keys, values) queries,
// each a dense sum + count over its own key range [0, card_q), in one
// persistent launch that a plan captures into a hipGraph.
//
// Per query the semantics are those of the dense group-by lane
// (numfast-native/src/groupby/dense.rs:18-37, lib.rs:85-103: sums and
// counts by key code, -2 on a key outside [0, g)), with exact int64 sums
// of int32 values, and the result compacted like the carry build
// (core/compact.rs:16-37): the present groups, ascending, as (ukey, count,
// sum), with per-query group offsets.
//
// Device layout:
//   slots: one packed u64 per (query, key): bits [40, 64) the row count,
//     [0, 40) the sum of (v + 2^15).  A row adds (1 << 40) + (v + 2^15) --
//     one atomic for count and sum, exact while a query has fewer than 2^24
//     rows (checked when the plan is built).  A value outside
//     [-2^15, 2^15) adds (1 << 40) + 2^15 (as if 0) and goes to an overflow
//     list (slot, value) that is added to the compacted sums at the end.
//   work items (query, row range) pulled from an atomic counter by
//     persistent workgroups; queries with card <= BT_LDS_SLOTS aggregate in
//     LDS and flush once per item, larger ones add to their slots directly.
//   compaction: one stable scan over per-query slot chunks (decoupled
//     look-back); chunk 0 of a query writes its group offset.
// The plain form launches the same kernels per query (one fill, one
// accumulate, one compaction per query), chaining the group offset on the
// device.
#include <algorithm>
#include <mutex>
#include <vector>

#include <memory>

#include "device.hpp"
#include "groupagg.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

typedef unsigned long long u64;

constexpr int BT_BLOCK = 256;
constexpr int BT_LDS_SLOTS = 4096;  // queries up to this card aggregate in LDS
constexpr int BT_CHUNK = 4096;      // slots per compaction chunk
constexpr int BT_CIPT = BT_CHUNK / BT_BLOCK;
constexpr int BT_SHIFT = 40;
constexpr u64 BT_MASK = (u64(1) << BT_SHIFT) - 1;
constexpr uint32_t BT_MAX_ROWS = 1u << 24;  // rows per query for the packed slots

struct BtItem {
  uint32_t q;
  uint32_t r0, r1;  // row range (global row indices)
};
struct BtChunk {
  uint32_t q;
  uint32_t pad;
  u64 s0, s1;  // slot range (global slot indices)
};
struct BtQuery {
  u64 slot0;     // first slot of the query
  uint32_t card;
  uint32_t pad;
};
struct BtOver {  // a row whose value is outside [-2^15, 2^15)
  u64 slot;
  int32_t v;
  uint32_t pad;
};
// device counters
struct BtMeta {
  uint32_t work;      // next work item
  uint32_t chunk;     // next compaction chunk
  uint32_t nover;     // overflow entries
  uint32_t bad_row;   // first row with a key out of range (0xFFFFFFFF: none)
  u64 base;           // group offset before the first chunk (0 when batched)
  u64 total;          // groups written
};

__device__ __forceinline__ u64 bt_packed(int32_t v, bool* over) {
  const uint32_t b = uint32_t(v) + 32768u;
  *over = b >= 65536u;
  return (u64(1) << BT_SHIFT) + (*over ? u64(32768) : u64(b));
}

// Persistent accumulation over work items [item0, item1).
__global__ __launch_bounds__(BT_BLOCK) void k_bt_accumulate(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const BtItem* __restrict__ items, uint32_t item0, uint32_t item1,
    const BtQuery* __restrict__ qs, u64* __restrict__ slots, BtOver* __restrict__ over,
    BtMeta* __restrict__ meta) {
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
    u64* qslots = slots + qd.slot0;
    if (card <= uint32_t(BT_LDS_SLOTS)) {
      for (uint32_t s = threadIdx.x; s < card; s += BT_BLOCK) s_acc[s] = 0;
      __syncthreads();
      for (uint32_t r = w.r0 + threadIdx.x; r < w.r1; r += BT_BLOCK) {
        const int32_t k = keys[r];
        if (uint32_t(k) >= card) {
          atomicMin(&meta->bad_row, r);
          continue;
        }
        bool ov;
        const u64 add = bt_packed(vals[r], &ov);
        atomicAdd(&s_acc[k], add);
        if (ov) over[atomicAdd(&meta->nover, 1u)] = BtOver{qd.slot0 + u64(k), vals[r], 0u};
      }
      __syncthreads();
      for (uint32_t s = threadIdx.x; s < card; s += BT_BLOCK) {
        const u64 a = s_acc[s];
        if (a) atomicAdd(&qslots[s], a);
      }
      __syncthreads();
    } else {
      for (uint32_t r = w.r0 + threadIdx.x; r < w.r1; r += BT_BLOCK) {
        const int32_t k = keys[r];
        if (uint32_t(k) >= card) {
          atomicMin(&meta->bad_row, r);
          continue;
        }
        bool ov;
        const u64 add = bt_packed(vals[r], &ov);
        atomicAdd(&qslots[k], add);
        if (ov) over[atomicAdd(&meta->nover, 1u)] = BtOver{qd.slot0 + u64(k), vals[r], 0u};
      }
    }
  }
}

// Stable compaction of present slots over chunks [c0, c1) (chunk order =
// query order, then key order).  Present slot s of query q becomes output
// group pos = base + prefix: ukeys[pos] = s - slot0(q), counts / sums from
// the packed word; slots[s] is overwritten with pos (for the overflow
// list).  Chunk 0 of each query writes group_off[q]; the last chunk writes
// meta->total and group_off[nq].
__global__ __launch_bounds__(BT_BLOCK) void k_bt_compact(
    const BtChunk* __restrict__ chunks, uint32_t c0, uint32_t c1,
    const BtQuery* __restrict__ qs, u64* __restrict__ slots, int64_t* __restrict__ ukeys,
    int64_t* __restrict__ counts, int64_t* __restrict__ sums, int64_t* __restrict__ group_off,
    uint32_t nq_end, u64 cap, uint64_t* __restrict__ lookback, uint32_t epoch,
    BtMeta* __restrict__ meta) {
  __shared__ uint32_t s_cnt[BT_CIPT][BT_BLOCK / WAVE];
  __shared__ uint32_t s_tile;
  __shared__ u64 s_pre;
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  if (tid == 0) s_tile = atomicAdd(&meta->chunk, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;  // position in [0, c1 - c0)
  const BtChunk ch = chunks[c0 + tile];
  const u64 qslot0 = qs[ch.q].slot0;
  const uint64_t lt = lanemask_lt();
  u64 pk[BT_CIPT];
  uint32_t r0[BT_CIPT];
#pragma unroll
  for (int j = 0; j < BT_CIPT; ++j) {
    const u64 s = ch.s0 + u64(j * BT_BLOCK + tid);
    pk[j] = s < ch.s1 ? slots[s] : 0ull;
    const uint64_t b = __ballot((pk[j] >> BT_SHIFT) != 0);
    r0[j] = uint32_t(__popcll(b & lt));
    if (lane == 0) s_cnt[j][wave] = uint32_t(__popcll(b));
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int j = 0; j < BT_CIPT; ++j)
#pragma unroll
      for (int w = 0; w < BT_BLOCK / WAVE; ++w) {
        const uint32_t c = s_cnt[j][w];
        s_cnt[j][w] = t;
        t += c;
      }
    uint64_t* my = lookback + tile;
    uint32_t ex = 0;
    if (tile == 0) {
      lb_store(my, lb_pack(epoch, LB_PREFIX, t));
    } else {
      lb_store(my, lb_pack(epoch, LB_AGG, t));
      ex = lb_lookback(lookback, tile, 1, epoch);
      lb_store(my, lb_pack(epoch, LB_PREFIX, ex + t));
    }
    const u64 pre = meta->base + u64(ex);
    s_pre = pre;
    if (ch.s0 == qslot0) group_off[ch.q] = int64_t(pre);
    if (tile == c1 - c0 - 1) {
      meta->total = pre + t;
      group_off[nq_end] = int64_t(pre + t);
    }
  }
  __syncthreads();
  const u64 pre = s_pre;
#pragma unroll
  for (int j = 0; j < BT_CIPT; ++j) {
    const u64 c = pk[j] >> BT_SHIFT;
    if (c == 0) continue;
    const u64 pos = pre + s_cnt[j][wave] + r0[j];
    const u64 s = ch.s0 + u64(j * BT_BLOCK + tid);
    if (pos < cap) {
      ukeys[pos] = int64_t(s - qslot0);
      counts[pos] = int64_t(c);
      sums[pos] = int64_t(pk[j] & BT_MASK) - int64_t(c) * 32768;
    }
    slots[s] = pos;
  }
}

// Overflow rows: their true values join the compacted sums.
__global__ void k_bt_overflow(const BtOver* __restrict__ over, const BtMeta* __restrict__ meta,
                              const u64* __restrict__ slots, int64_t* __restrict__ sums, u64 cap) {
  const uint32_t n = meta->nover;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const BtOver e = over[i];
    const u64 pos = slots[e.slot];
    if (pos < cap) atomicAdd(reinterpret_cast<u64*>(sums + pos), u64(int64_t(e.v)));
  }
}

// meta reset for a batched run: counters 0, bad_row none, base 0.
__global__ void k_bt_reset(BtMeta* meta) {
  meta->work = 0;
  meta->chunk = 0;
  meta->nover = 0;
  meta->bad_row = 0xFFFFFFFFu;
  meta->base = 0;
  meta->total = 0;
}
// per-query chain: counters 0, base = previous query's total.
__global__ void k_bt_next(BtMeta* meta) {
  meta->work = 0;
  meta->chunk = 0;
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
  std::vector<uint32_t> item_start, chunk_start;  // per query (+1): ranges
  BtItem* items = nullptr;
  BtChunk* chunks = nullptr;
  BtQuery* qs = nullptr;
  u64* slots = nullptr;
  BtOver* over = nullptr;
  BtMeta* meta = nullptr;
  uint64_t* lookback = nullptr;
  uint32_t nitems = 0, nchunks = 0;
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

static void build_items(BatchPlan& P, const std::vector<int64_t>& roff, uint32_t R) {
  std::vector<BtItem> it;
  P.item_start.assign(P.nq + 1, 0);
  for (uint32_t q = 0; q < P.nq; ++q) {
    P.item_start[q] = uint32_t(it.size());
    for (int64_t r = roff[q]; r < roff[q + 1]; r += R)
      it.push_back({q, uint32_t(r), uint32_t(std::min<int64_t>(r + R, roff[q + 1]))});
  }
  P.item_start[P.nq] = uint32_t(it.size());
  P.nitems = uint32_t(it.size());
  if (!P.items) P.items = P.alloc<BtItem>(P.nrows / 512 + P.nq + 16);  // R >= 512
  NFH_CHECK(hipMemcpy(P.items, it.data(), it.size() * sizeof(BtItem), hipMemcpyHostToDevice));
  P.rows_per_item = R;
}

// The whole batched run on P.s: reset, zero slots, accumulate, compact,
// overflow.  No host synchronisation (graph-capturable).
static void enqueue_batched(BatchPlan& P) {
  hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, P.s, P.meta);
  NFH_LAUNCH_CHECK();
  fill_u64(P.slots, 0, P.nslots, P.s);
  fill_u64(P.lookback, 0, P.nchunks + 1, P.s);  // look-back words of the last run
  hipLaunchKernelGGL(k_bt_accumulate, dim3(P.grid), dim3(BT_BLOCK), 0, P.s, P.keys, P.vals,
                     P.items, 0u, P.nitems, P.qs, P.slots, P.over, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_compact, dim3(P.nchunks), dim3(BT_BLOCK), 0, P.s, P.chunks, 0u,
                     P.nchunks, P.qs, P.slots, P.ukeys, P.counts, P.sums, P.group_off, P.nq,
                     P.cap, P.lookback, 1u, P.meta);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bt_overflow, dim3(256), dim3(256), 0, P.s, P.over, P.meta, P.slots,
                     P.sums, P.cap);
  NFH_LAUNCH_CHECK();
}

// The plain form: per query a slot fill, an accumulate launch over its
// items and a compaction launch over its chunks, chained on the device.
static void enqueue_per_query(BatchPlan& P) {
  hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, P.s, P.meta);
  NFH_LAUNCH_CHECK();
  fill_u64(P.slots, 0, P.nslots, P.s);
  fill_u64(P.lookback, 0, P.nchunks + 1, P.s);  // epochs 2.. below are unique per query
  for (uint32_t q = 0; q < P.nq; ++q) {
    const uint32_t i0 = P.item_start[q], i1 = P.item_start[q + 1];
    const uint32_t c0 = P.chunk_start[q], c1 = P.chunk_start[q + 1];
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
    hipLaunchKernelGGL(k_bt_compact, dim3(c1 - c0), dim3(BT_BLOCK), 0, P.s, P.chunks, c0, c1,
                       P.qs, P.slots, P.ukeys, P.counts, P.sums, P.group_off, q + 1, P.cap,
                       P.lookback, 2u + (q & 0x3FFFFFFu), P.meta);
    NFH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_bt_overflow, dim3(256), dim3(256), 0, P.s, P.over, P.meta, P.slots,
                     P.sums, P.cap);
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
    fill_u64(P.slots, 0, P.nslots, P.s);
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
// picks rows per work item and the persistent grid by timing candidates on
// these buffers (autotuning), then captures reset, slot fill, accumulate,
// compaction and overflow into one hipGraph.  Returns a handle or NULL
// (*rc_out: -1 null, -3 a query over 2^24 rows or bad offsets, -6 device).
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
    if (roff[0] != 0) return fail(MALFORMED);
    for (size_t q = 0; q < nq; ++q)
      if (roff[q + 1] < roff[q] || roff[q + 1] - roff[q] >= int64_t(BT_MAX_ROWS) || cd[q] < 0)
        return fail(MALFORMED);
    if (roff[nq] > int64_t(INT32_MAX)) return fail(MALFORMED);
    // the dense slots (8 B per key of every query) and the overflow list
    // (16 B per row) must fit the free device memory
    uint64_t total_card = 0;
    for (size_t q = 0; q < nq; ++q) total_card += uint64_t(cd[q]);
    size_t free_b = 0, total_b = 0;
    NFH_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (total_card * 8 + uint64_t(roff[nq]) * 16 > uint64_t(free_b)) return fail(MALFORMED);
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
    NFH_CHECK(hipStreamCreate(&P->s));
    // queries, slots, chunks
    std::vector<BtQuery> qs(nq);
    std::vector<BtChunk> chs;
    P->chunk_start.assign(nq + 1, 0);
    u64 slot = 0;
    for (size_t q = 0; q < nq; ++q) {
      qs[q] = {slot, uint32_t(cd[q]), 0};
      P->chunk_start[q] = uint32_t(chs.size());
      u64 s0 = slot;
      do {
        const u64 s1 = std::min<u64>(s0 + BT_CHUNK, slot + u64(cd[q]));
        chs.push_back({uint32_t(q), 0, s0, s1});
        s0 = s1;
      } while (s0 < slot + u64(cd[q]));
      slot += u64(cd[q]);
    }
    P->chunk_start[nq] = uint32_t(chs.size());
    P->nslots = slot;
    P->nchunks = uint32_t(chs.size());
    P->qs = P->alloc<BtQuery>(nq);
    P->chunks = P->alloc<BtChunk>(chs.size());
    P->slots = P->alloc<u64>(P->nslots);
    P->over = P->alloc<BtOver>(P->nrows + 1);
    P->meta = P->alloc<BtMeta>(1);
    P->lookback = P->alloc<uint64_t>(chs.size() + 1);
    NFH_CHECK(hipMemcpy(P->qs, qs.data(), nq * sizeof(BtQuery), hipMemcpyHostToDevice));
    NFH_CHECK(hipMemcpy(P->chunks, chs.data(), chs.size() * sizeof(BtChunk),
                        hipMemcpyHostToDevice));
    fill_u64(P->lookback, 0, chs.size() + 1, P->s);
    // autotune rows per work item and the persistent grid
    float best = 1e30f;
    uint32_t bestR = 8192, bestG = 4;
    const uint32_t cus = uint32_t(num_cus());
    for (uint32_t R : {2048u, 8192u, 32768u}) {
      build_items(*P, roff, R);
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
    build_items(*P, roff, bestR);
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

// Plan details for reports: rows per work item, workgroups, work items,
// slots.
NF_EXPORT int32_t nfh_group_batch_info(void* plan, uint64_t* out4) {
  if (!plan || !out4) return NULL_OR_ABORT;
  auto* P = static_cast<BatchPlan*>(plan);
  out4[0] = P->rows_per_item;
  out4[1] = P->grid;
  out4[2] = P->nitems;
  out4[3] = P->nslots;
  return OK;
}

NF_EXPORT void nfh_group_batch_free(void* plan) {
  auto* P = static_cast<BatchPlan*>(plan);
  if (P) {
    (void)hipStreamSynchronize(P->s);
    delete P;
  }
}
