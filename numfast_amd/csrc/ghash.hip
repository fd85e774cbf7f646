// ghash.hip -- partitioned pair-hash for grouped COUNT DISTINCT over int64
// (key, value) pairs: the six nf_ghash_* entry points of
// numfast-native/src/lib.rs:2266-2470 (algorithm:
// numfast-native/src/groupby/hash_grouped.rs).
//
// The reference shards rows (fingerprint, count, scatter) or partitions
// (insert, occupancy) over `nthreads` CPU threads.  Its outputs depend on
// that sharding in three places, which this file reproduces exactly:
//   * count: chunk tt of nt = min(fanout(nthreads, n), t) equal row chunks
//     adds into histogram row tt;
//   * scatter: chunk tt writes its rows in row order from cursor row tt;
//   * occ_fill: shard tt of the partitions writes from ostart[first].
// The insert (pins) is sequential linear probing per partition in row
// order; its table layout is an output.  The GPU reproduces that layout
// without a sequential loop:
//   1. dedupe: every row inserts its pair into a scratch table of row ids
//      (compare-and-swap to claim, full pair compare, atomic min of the row
//      id on equal pairs), so each distinct pair ends with its first row;
//   2. place: only those first rows insert into the real geometry, each
//      claiming slots with an atomic min of its row id and carrying any
//      larger id it evicts onward.  At rest every slot between a pair's
//      home and its slot holds a smaller row id and its own slot does not,
//      which is exactly where sequential insertion in row order puts it
//      (by induction on the row id), so the layout is identical;
//   3. materialise used / tk / tv from the row ids.
// Refusals where the reference would index out of bounds (a panic, which
// aborts the process) or probe forever: -2 (partition id or reservation
// outside its buffer), -3 (a table with no free slot left for a new pair).
#include <algorithm>
#include <vector>

#include "device.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

namespace {

constexpr int GH_PART_MAX = 256;
constexpr int GH_THREADS_MAX = 64;
constexpr int GH_BLOCK = 256;
constexpr int GH_IPT = 16;
constexpr int GH_TILE = GH_BLOCK * GH_IPT;  // rows per scatter / count tile
constexpr uint32_t GH_EMPTY = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint64_t gh_pair_fp(int64_t k, int64_t v) {
  const uint64_t uk = uint64_t(k), uv = uint64_t(v);
  return uk * 0x9E3779B97F4A7C15ull + (uv ^ (uv >> 33)) * 0x00BF58476D1FF973ull;
}

size_t gh_fanout(size_t nthreads, size_t work) {
  const size_t t = std::min<size_t>(std::max<size_t>(nthreads, 1), GH_THREADS_MAX);
  return std::min(t, std::max<size_t>(work, 1));
}

bool gh_pow2(size_t p) { return p != 0 && p <= size_t(GH_PART_MAX) && (p & (p - 1)) == 0; }

// Row chunks of count / scatter cut into GH_TILE-row tiles that never
// straddle a chunk: tile_first[c] = first tile of chunk c.
struct ChunkTiles {
  uint32_t nt = 0;
  uint64_t cs = 0;
  std::vector<uint32_t> first;  // nt + 1
};

ChunkTiles chunk_tiles(size_t n, size_t nthreads, size_t t) {
  ChunkTiles c;
  c.nt = uint32_t(std::min(gh_fanout(nthreads, n), t));
  c.cs = (n + c.nt - 1) / c.nt;
  c.first.assign(c.nt + 1, 0);
  for (uint32_t tt = 0; tt < c.nt; ++tt) {
    const uint64_t s0 = std::min<uint64_t>(tt * c.cs, n), s1 = std::min<uint64_t>((tt + 1) * c.cs, n);
    c.first[tt + 1] = c.first[tt] + uint32_t((s1 - s0 + GH_TILE - 1) / GH_TILE);
  }
  return c;
}

struct TileGeom {
  const uint32_t* first;  // nt + 1 tile starts (device)
  uint32_t nt;
  uint64_t cs;
  uint64_t n;
  // chunk and row range of tile b
  __device__ void locate(uint32_t b, uint32_t* chunk, uint64_t* r0, uint64_t* r1) const {
    uint32_t c = 0;
    while (c + 1 < nt && first[c + 1] <= b) ++c;
    const uint64_t c0 = uint64_t(c) * cs;
    const uint64_t c1 = c0 + cs < n ? c0 + cs : n;
    *chunk = c;
    *r0 = c0 + uint64_t(b - first[c]) * GH_TILE;
    *r1 = *r0 + GH_TILE < c1 ? *r0 + GH_TILE : c1;
  }
};

// ------------------------------------------------------------ fingerprint
__global__ void k_gh_fp(const int64_t* __restrict__ k, const int64_t* __restrict__ v, uint64_t n,
                        int64_t* __restrict__ fp, uint64_t pmask) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    fp[i] = int64_t(gh_pair_fp(k[i], v[i]) & pmask);
}

// ------------------------------------------------------------ count
// One workgroup per tile: LDS histogram of its rows' partition ids, added
// into histogram row `chunk` (the caller's buffer accumulates).
__global__ __launch_bounds__(GH_BLOCK) void k_gh_count(const int64_t* __restrict__ fp,
                                                       TileGeom g, uint32_t p,
                                                       unsigned long long* __restrict__ tc,
                                                       uint32_t* __restrict__ err) {
  __shared__ uint32_t h[GH_PART_MAX];
  for (uint32_t q = threadIdx.x; q < p; q += GH_BLOCK) h[q] = 0;
  __syncthreads();
  uint32_t c;
  uint64_t r0, r1;
  g.locate(blockIdx.x, &c, &r0, &r1);
  bool bad = false;
  for (uint64_t i = r0 + threadIdx.x; i < r1; i += GH_BLOCK) {
    const int64_t f = fp[i];
    if (f < 0 || uint64_t(f) >= p)
      bad = true;
    else
      atomicAdd(&h[f], 1u);
  }
  if (bad) atomicOr(err, 1u);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < p; q += GH_BLOCK)
    if (h[q]) atomicAdd(&tc[size_t(c) * p + q], (unsigned long long)h[q]);
}

// ------------------------------------------------------------ scatter
// Per-tile partition counts, partition-major: th[q * tiles + tile].
__global__ __launch_bounds__(GH_BLOCK) void k_gh_tile_hist(const int64_t* __restrict__ fp,
                                                           TileGeom g, uint32_t p,
                                                           uint32_t tiles,
                                                           uint32_t* __restrict__ th,
                                                           uint32_t* __restrict__ err) {
  __shared__ uint32_t h[GH_PART_MAX];
  for (uint32_t q = threadIdx.x; q < p; q += GH_BLOCK) h[q] = 0;
  __syncthreads();
  uint32_t c;
  uint64_t r0, r1;
  g.locate(blockIdx.x, &c, &r0, &r1);
  bool bad = false;
  for (uint64_t i = r0 + threadIdx.x; i < r1; i += GH_BLOCK) {
    const int64_t f = fp[i];
    if (f < 0 || uint64_t(f) >= p)
      bad = true;
    else
      atomicAdd(&h[f], 1u);
  }
  if (bad) atomicOr(err, 1u);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < p; q += GH_BLOCK) th[size_t(q) * tiles + blockIdx.x] = h[q];
}

// Cursor of each (tile, partition): tpos[chunk * p + q] plus the counts
// of the chunk's earlier tiles.  One workgroup per (chunk, partition),
// scanning the chunk's tiles 256 at a time.
__global__ __launch_bounds__(GH_BLOCK) void k_gh_tile_base(const uint32_t* __restrict__ th,
                                                           const uint32_t* __restrict__ first,
                                                           uint32_t p, uint32_t tiles,
                                                           const int64_t* __restrict__ tpos,
                                                           int64_t* __restrict__ tb) {
  __shared__ uint32_t s_tmp[GH_BLOCK / WAVE];
  const uint32_t c = blockIdx.x / p, q = blockIdx.x % p;
  int64_t run = tpos[size_t(c) * p + q];
  const uint32_t* row = th + size_t(q) * tiles;
  int64_t* out = tb + size_t(q) * tiles;
  for (uint32_t b0 = first[c]; b0 < first[c + 1]; b0 += GH_BLOCK) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t x = b < first[c + 1] ? row[b] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<GH_BLOCK / WAVE>(x, s_tmp, &tot);
    if (b < first[c + 1]) out[b] = run + int64_t(ex);
    run += int64_t(tot);
    __syncthreads();
  }
}

// One workgroup per tile: stable in-tile ranks per partition (item j of
// lane l of wave w is row r0 + w*64*IPT + j*64 + l; per-wave LDS counters
// give ranks in that order), then the pairs go to their reserved lanes.
template <bool LR>
__global__ __launch_bounds__(GH_BLOCK) void k_gh_scatter(
    const int64_t* __restrict__ k, const int64_t* __restrict__ v, const int64_t* __restrict__ fp,
    TileGeom g, uint32_t p, uint32_t tiles, const int64_t* __restrict__ tb,
    int64_t* __restrict__ pk,
    int64_t* __restrict__ pv, uint32_t* __restrict__ err) {
  constexpr int NW = GH_BLOCK / WAVE;
  __shared__ uint32_t s_cnt[NW][GH_PART_MAX];
  const int lane = lane_id(), wave = wave_id();
  for (int i = threadIdx.x; i < NW * GH_PART_MAX; i += GH_BLOCK) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  uint32_t c;
  uint64_t r0, r1;
  g.locate(blockIdx.x, &c, &r0, &r1);
  const uint64_t wb = r0 + uint64_t(wave) * (WAVE * GH_IPT) + uint64_t(lane);
  uint32_t d[GH_IPT], rk[GH_IPT];
  uint32_t live = 0;
#pragma unroll
  for (int j = 0; j < GH_IPT; ++j) {
    const uint64_t i = wb + uint64_t(j * WAVE);
    const int64_t f = i < r1 ? fp[i] : -1;
    const bool ok = f >= 0 && uint64_t(f) < p;  // out-of-range ids: flagged by the histogram
    d[j] = ok ? uint32_t(f) : 0u;
    live |= uint32_t(ok) << j;
  }
#pragma unroll
  for (int j = 0; j < GH_IPT; ++j) {
    const bool valid = (live >> j) & 1u;
    if constexpr (LR) {
      rk[j] = valid ? atomicAdd(&s_cnt[wave][d[j]], 1u) : 0u;
    } else {
      const uint64_t peers = match_digit8(d[j], valid);
      const uint32_t r = __popcll(peers & lanemask_lt());
      const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : lane;
      uint32_t old = 0;
      if (valid && r == 0) {
        old = s_cnt[wave][d[j]];
        s_cnt[wave][d[j]] = old + uint32_t(__popcll(peers));
      }
      old = __shfl(old, leader);
      rk[j] = old + r;
    }
  }
  __syncthreads();
  // per partition: exclusive offsets of the waves' runs
  for (uint32_t q = threadIdx.x; q < p; q += GH_BLOCK) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t x = s_cnt[w][q];
      s_cnt[w][q] = run;
      run += x;
    }
  }
  __syncthreads();
  bool bad = false;
#pragma unroll
  for (int j = 0; j < GH_IPT; ++j) {
    if (!((live >> j) & 1u)) continue;
    const uint64_t i = wb + uint64_t(j * WAVE);
    const int64_t lanepos = tb[size_t(d[j]) * tiles + blockIdx.x] + int64_t(s_cnt[wave][d[j]] + rk[j]);
    if (lanepos < 0 || uint64_t(lanepos) >= g.n) {
      bad = true;
      continue;
    }
    pk[lanepos] = k[i];
    pv[lanepos] = v[i];
  }
  if (bad) atomicOr(err, 2u);
}

// ------------------------------------------------------------ insert
struct PinsGeom {
  const int64_t* starts;  // p + 1
  const int64_t* toff;    // p + 1
  const int64_t* tm;      // p
  uint32_t p;
  uint64_t n;
};

// Partition of row j (starts is non-decreasing; rows of empty partitions
// do not exist).
__device__ __forceinline__ uint32_t gh_part_of(const int64_t* s_starts, uint32_t p, uint64_t j) {
  uint32_t lo = 0, hi = p;  // last q with starts[q] <= j
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (uint64_t(s_starts[mid]) <= j)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t ld_holder(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 1. dedupe: each distinct pair of a partition ends in one slot of `hold`
// holding its smallest row id.  Any layout will do here, so the home is
// the fingerprint's top bits (well mixed), not the reference's low bits
// (which within a partition repeat its id and make long probe runs).
__global__ __launch_bounds__(GH_BLOCK) void k_gh_dedupe(const int64_t* __restrict__ pk,
                                                        const int64_t* __restrict__ pv,
                                                        PinsGeom g, uint32_t* __restrict__ hold,
                                                        uint32_t* __restrict__ err) {
  __shared__ int64_t s_starts[GH_PART_MAX + 1];
  for (uint32_t q = threadIdx.x; q <= g.p; q += GH_BLOCK) s_starts[q] = g.starts[q];
  __syncthreads();
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < g.n; j += stride) {
    const uint32_t q = gh_part_of(s_starts, g.p, j);
    const uint64_t mask = uint64_t(g.tm[q]);
    uint32_t* tab = hold + g.toff[q];
    const int64_t rk = pk[j], vv = pv[j];
    const int lg = 64 - __clzll((unsigned long long)mask);  // log2(slots)
    uint64_t h = lg == 0 ? 0 : (gh_pair_fp(rk, vv) >> (64 - lg));
    bool placed = false;
    for (uint64_t probes = 0; probes <= mask; ++probes) {
      uint32_t cur = ld_holder(tab + h);
      if (cur == GH_EMPTY) {
        cur = atomicCAS(tab + h, GH_EMPTY, uint32_t(j));
        if (cur == GH_EMPTY) {
          placed = true;
          break;
        }
      }
      if (pk[cur] == rk && pv[cur] == vv) {
        atomicMin(tab + h, uint32_t(j));
        placed = true;
        break;
      }
      h = (h + 1) & mask;
    }
    if (!placed) atomicOr(err, 4u);  // no free slot for a new pair
  }
}

__global__ void k_gh_first(const uint32_t* __restrict__ hold, uint64_t m,
                           uint8_t* __restrict__ first) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < m; s += stride) {
    const uint32_t h = hold[s];
    if (h != GH_EMPTY) first[h] = 1;
  }
}

// 2. place, in LDS windows.  A partition's table is cut into windows of
// min(slots, GH_WIN) slots, the grid offset by (q mod slots): the
// reference's homes in partition q all sit at q modulo P, so windows that
// start at a home rarely begin inside a cluster.  The first rows are
// sorted by the window of
// their reference home (fp & mask), stable in row order.  Each window runs
// the priority claim in LDS (atomic min of row ids, evicted ids carried
// on) over its slots plus GH_EXT extension slots; the ids that end in the
// extension spill into the next window (cyclically), where they start at
// the first slot: every slot they passed holds a smaller id.  A window of
// the whole table wraps inside LDS and never spills.  Windows are first
// solved alone (spill-ins empty); then one workgroup per partition walks
// its windows in cyclic order from one whose predecessor spilled nothing
// and re-solves each window whose predecessor spilled, with those spills.
// A cluster longer than the extension sets the overflow flag, and the
// partition-free global form (k_gh_place) solves the whole call instead.
constexpr int GH_WIN = 8192;
constexpr int GH_EXT = 1024;
constexpr int GH_WBLOCK = 256;

struct WinGeom {
  const int64_t* toff;       // p + 1
  const int64_t* tm;         // p
  const uint32_t* winbase;   // p + 1: first window of each partition
  uint32_t p;
  uint32_t nwin;
};

__device__ __forceinline__ uint32_t gh_win_part(const uint32_t* s_wb, uint32_t p, uint32_t w) {
  uint32_t lo = 0, hi = p;  // last q with winbase[q] <= w
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (s_wb[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Window id of each row's reference home; GH_EMPTY for rows that are not
// their pair's first row (sorted last).
__global__ void k_gh_winkey(const int64_t* __restrict__ pk, const int64_t* __restrict__ pv,
                            const uint8_t* __restrict__ first, PinsGeom g,
                            const uint32_t* __restrict__ winbase, int32_t* __restrict__ key) {
  __shared__ int64_t s_starts[GH_PART_MAX + 1];
  for (uint32_t q = threadIdx.x; q <= g.p; q += blockDim.x) s_starts[q] = g.starts[q];
  __syncthreads();
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < g.n; j += stride) {
    if (!first[j]) {
      key[j] = int32_t(GH_EMPTY ^ 0x80000000u);  // sorts after every window id
      continue;
    }
    const uint32_t q = gh_part_of(s_starts, g.p, j);
    const uint64_t mask = uint64_t(g.tm[q]);
    const uint64_t wsz = mask + 1 < uint64_t(GH_WIN) ? mask + 1 : uint64_t(GH_WIN);
    const uint64_t h = gh_pair_fp(pk[j], pv[j]) & mask;
    const uint64_t off = uint64_t(q) & mask;
    key[j] = int32_t((winbase[q] + uint32_t(((h - off) & mask) / wsz)) ^ 0x80000000u);
  }
}

// wstart[w] = first sorted position of window w (w <= nwin).
__global__ void k_gh_winstart(const int32_t* __restrict__ skey, uint64_t n, uint32_t nwin,
                              uint32_t* __restrict__ wstart) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > nwin) return;
  const int32_t target = int32_t(w ^ 0x80000000u);
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (skey[mid] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  wstart[w] = uint32_t(lo);
}

// LDS address of window slot i.  A partition's reference homes are
// congruent modulo P (a multiple of the 64 LDS banks), so unswizzled every
// claim of a window starts in one bank; XOR-ing the low 6 bits with bits
// 8..13 spreads homes 256 slots apart over the banks.  It permutes each
// aligned 64-slot block, so [0, span) maps onto itself (span is a multiple
// of 64) and the probe order, which follows slot numbers, is unchanged.
__device__ __forceinline__ uint32_t gh_swz(uint32_t i) { return i ^ ((i >> 8) & 63u); }

// Solve window w in LDS: its rows perm[wstart[w] .. wstart[w+1]) from
// their homes, plus `nin` spilled ids from slot 0.  Writes the window's
// slots to `hold` and its extension's ids (slot order) to spill[w];
// returns through *changed whether the spill list differs from before.
__device__ void gh_solve_window(uint32_t w, const WinGeom& g, const uint32_t* s_wb,
                                const int64_t* __restrict__ pk, const int64_t* __restrict__ pv,
                                const int32_t* __restrict__ perm,
                                const uint32_t* __restrict__ wstart,
                                const uint32_t* __restrict__ spill_in, uint32_t nin,
                                uint32_t* __restrict__ hold, uint32_t* __restrict__ spill,
                                uint32_t* __restrict__ nspill, uint32_t* s_hold,
                                uint32_t* s_misc, uint32_t* __restrict__ err) {
  const uint32_t q = gh_win_part(s_wb, g.p, w);
  const uint64_t mask = uint64_t(g.tm[q]);
  const uint32_t wsz = mask + 1 < uint64_t(GH_WIN) ? uint32_t(mask + 1) : uint32_t(GH_WIN);
  const bool whole = uint64_t(wsz) == mask + 1;
  const uint32_t ext = whole ? 0u : uint32_t(GH_EXT);
  // first slot, within the partition (cyclic)
  const uint64_t a = (uint64_t(q) + uint64_t(w - s_wb[q]) * wsz) & mask;
  const uint32_t span = wsz + ext;
  for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) s_hold[i] = GH_EMPTY;
  if (threadIdx.x == 0) {
    s_misc[0] = 0;
    s_misc[1] = 0;
  }
  __syncthreads();
  const uint32_t r0 = wstart[w], r1 = wstart[w + 1];
  const uint32_t total = (r1 - r0) + nin;
  bool over = false;
  for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
    uint32_t carried, c;
    if (t < nin) {
      carried = spill_in[t];
      c = 0;
    } else {
      carried = uint32_t(perm[r0 + (t - nin)]);
      c = uint32_t(((gh_pair_fp(pk[carried], pv[carried]) & mask) - a) & mask);
    }
    for (uint32_t it = 0;; ++it) {
      if (c >= span) {
        if (!whole) {
          over = true;
          break;
        }
        c &= wsz - 1;
      }
      if (it > 4 * span + total) {  // cannot happen while the table has room
        over = true;
        break;
      }
      const uint32_t old = atomicMin(&s_hold[gh_swz(c)], carried);
      if (old == GH_EMPTY) break;
      if (old > carried) carried = old;
      ++c;
    }
  }
  if (over) atomicOr(err, 32u);
  __syncthreads();
  uint32_t* tab = hold + g.toff[q];
  for (uint32_t i = threadIdx.x; i < wsz; i += blockDim.x) tab[(a + i) & mask] = s_hold[gh_swz(i)];
  if (ext == 0) {
    if (threadIdx.x == 0) {
      if (nspill[w] != 0) s_misc[1] = 1;
      nspill[w] = 0;
    }
    __syncthreads();
    return;
  }
  // extension ids in slot order (one wave per step of 64 slots)
  uint32_t* out = spill + size_t(w) * GH_EXT;
  const uint32_t prev = nspill[w];
  __syncthreads();
  if (threadIdx.x < WAVE) {
    uint32_t base = 0;
    bool diff = false;
    for (uint32_t s0 = 0; s0 < ext; s0 += WAVE) {
      const uint32_t v = s_hold[gh_swz(wsz + s0 + threadIdx.x)];
      const bool occ = v != GH_EMPTY;
      const uint64_t bal = __ballot(occ);
      if (occ) {
        const uint32_t pos = base + uint32_t(__popcll(bal & lanemask_lt()));
        if (pos >= prev || out[pos] != v) diff = true;
        out[pos] = v;
      }
      base += uint32_t(__popcll(bal));
    }
    diff = __ballot(diff) != 0 || base != prev;
    if (threadIdx.x == 0) {
      nspill[w] = base;
      s_misc[1] = diff ? 1u : 0u;
    }
  }
  __syncthreads();
}

// Phase A: every window alone.
__global__ __launch_bounds__(GH_WBLOCK) void k_gh_win(WinGeom g, const int64_t* __restrict__ pk,
                                                      const int64_t* __restrict__ pv,
                                                      const int32_t* __restrict__ perm,
                                                      const uint32_t* __restrict__ wstart,
                                                      uint32_t* __restrict__ hold,
                                                      uint32_t* __restrict__ spill,
                                                      uint32_t* __restrict__ nspill,
                                                      uint32_t* __restrict__ err) {
  __shared__ uint32_t s_hold[GH_WIN + GH_EXT];
  __shared__ uint32_t s_wb[GH_PART_MAX + 1];
  __shared__ uint32_t s_misc[2];
  for (uint32_t q = threadIdx.x; q <= g.p; q += blockDim.x) s_wb[q] = g.winbase[q];
  __syncthreads();
  gh_solve_window(blockIdx.x, g, s_wb, pk, pv, perm, wstart, nullptr, 0, hold, spill, nspill,
                  s_hold, s_misc, err);
}

// Phase B: one workgroup per partition of several windows walks them in
// cyclic order from a window whose predecessor spilled nothing (else from
// window 0, three laps), re-solving each window whose predecessor spills.
// A spill list still changing on the last lap sets the overflow flag.
__global__ __launch_bounds__(GH_WBLOCK) void k_gh_win_fix(WinGeom g, const int64_t* __restrict__ pk,
                                                          const int64_t* __restrict__ pv,
                                                          const int32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ wstart,
                                                          uint32_t* __restrict__ hold,
                                                          uint32_t* __restrict__ spill,
                                                          uint32_t* __restrict__ nspill,
                                                          uint32_t* __restrict__ tmp,
                                                          uint32_t* __restrict__ err) {
  __shared__ uint32_t s_hold[GH_WIN + GH_EXT];
  __shared__ uint32_t s_wb[GH_PART_MAX + 1];
  __shared__ uint32_t s_misc[2];
  __shared__ uint32_t s_start;
  for (uint32_t q = threadIdx.x; q <= g.p; q += blockDim.x) s_wb[q] = g.winbase[q];
  __syncthreads();
  const uint32_t q = blockIdx.x;
  const uint32_t w0 = s_wb[q], nw = s_wb[q + 1] - w0;
  if (nw <= 1) return;
  bool any = false;
  for (uint32_t i = 0; i < nw; ++i) any |= nspill[w0 + i] != 0;
  if (!any) return;
  if (threadIdx.x == 0) {
    uint32_t st = nw;  // a window whose predecessor spilled nothing
    for (uint32_t i = 0; i < nw && st == nw; ++i)
      if (nspill[w0 + (i + nw - 1) % nw] == 0) st = i;
    s_start = st;
  }
  __syncthreads();
  const uint32_t st = s_start == nw ? 0 : s_start;
  const uint32_t laps = s_start == nw ? 3 : 1;
  // the spill-ins of a window are copied out first: the window rewrites
  // its own list, and with one window per partition it is its own
  // predecessor
  uint32_t* in = tmp + size_t(q) * GH_EXT;
  for (uint32_t step = 0; step < laps * nw; ++step) {
    const uint32_t wi = (st + step) % nw, w = w0 + wi, pw = w0 + (wi + nw - 1) % nw;
    const uint32_t nin = nspill[pw];
    if (nin == 0 && step < nw && laps == 1) continue;  // its alone-solution stands
    for (uint32_t i = threadIdx.x; i < nin; i += blockDim.x) in[i] = spill[size_t(pw) * GH_EXT + i];
    __syncthreads();
    gh_solve_window(w, g, s_wb, pk, pv, perm, wstart, in, nin, hold, spill, nspill, s_hold,
                    s_misc, err);
    if (laps == 3 && step >= 2 * nw && s_misc[1] && threadIdx.x == 0) atomicOr(err, 32u);
    __syncthreads();
  }
}

// 2. place: first rows claim slots in the real geometry by atomic min of
// the row id; an evicted (larger) id is carried on from the next slot.
__global__ __launch_bounds__(GH_BLOCK) void k_gh_place(const int64_t* __restrict__ pk,
                                                       const int64_t* __restrict__ pv,
                                                       const uint8_t* __restrict__ first,
                                                       PinsGeom g, uint32_t* __restrict__ hold,
                                                       uint32_t* __restrict__ err) {
  __shared__ int64_t s_starts[GH_PART_MAX + 1];
  for (uint32_t q = threadIdx.x; q <= g.p; q += GH_BLOCK) s_starts[q] = g.starts[q];
  __syncthreads();
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < g.n; j += stride) {
    if (!first[j]) continue;
    const uint32_t q = gh_part_of(s_starts, g.p, j);
    const uint64_t mask = uint64_t(g.tm[q]);
    uint32_t* tab = hold + g.toff[q];
    uint32_t carried = uint32_t(j);
    uint64_t h = gh_pair_fp(pk[j], pv[j]) & mask;
    bool done = false;
    // every step moves the cursor one slot; a partition's pairs fit its
    // table (checked by the dedupe), so this bound is never reached
    const uint64_t limit = 4 * (mask + 1) + 64;
    for (uint64_t it = 0; it < limit; ++it) {
      const uint32_t old = atomicMin(tab + h, carried);
      if (old == GH_EMPTY) {
        done = true;
        break;
      }
      if (old > carried) carried = old;  // took the slot: carry the evicted id on
      h = (h + 1) & mask;
    }
    if (!done) atomicOr(err, 8u);
  }
}

// 3. used / tk / tv of the occupied slots (others untouched).
__global__ void k_gh_materialize(const uint32_t* __restrict__ hold, uint64_t m,
                                 const int64_t* __restrict__ pk, const int64_t* __restrict__ pv,
                                 int64_t* __restrict__ tk, int64_t* __restrict__ tv,
                                 uint8_t* __restrict__ used) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < m; s += stride) {
    const uint32_t h = hold[s];
    if (h == GH_EMPTY) continue;
    used[s] = 1;
    tk[s] = pk[h];
    tv[s] = pv[h];
  }
}

// ------------------------------------------------------------ occupancy
// Table windows cut into GH_CH-slot chunks: chunk c of partition q is
// cb[q] + c, covering [toff[q] + c*GH_CH, min(.., toff[q+1])).
constexpr int GH_CH = 4096;
constexpr int GH_CH_IPT = GH_CH / GH_BLOCK;

__device__ __forceinline__ uint32_t gh_chunk_part(const uint32_t* s_cb, uint32_t p, uint32_t c) {
  uint32_t lo = 0, hi = p;  // last q with cb[q] <= c
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (s_cb[mid] <= c)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Used slots per chunk.
__global__ __launch_bounds__(GH_BLOCK) void k_gh_chunk_count(const uint8_t* __restrict__ used,
                                                             const int64_t* __restrict__ toff,
                                                             const uint32_t* __restrict__ cb,
                                                             uint32_t p, uint32_t* __restrict__ cc) {
  __shared__ uint32_t s_cb[GH_PART_MAX + 1];
  __shared__ uint32_t s_w[GH_BLOCK / WAVE];
  for (uint32_t q = threadIdx.x; q <= p; q += GH_BLOCK) s_cb[q] = cb[q];
  __syncthreads();
  const uint32_t c = blockIdx.x, q = gh_chunk_part(s_cb, p, c);
  const int64_t s0 = toff[q] + int64_t(c - s_cb[q]) * GH_CH;
  const int64_t s1 = s0 + GH_CH < toff[q + 1] ? s0 + GH_CH : toff[q + 1];
  uint32_t x = 0;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += GH_BLOCK) x += used[s] != 0;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if (lane_id() == 0) s_w[wave_id()] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < GH_BLOCK / WAVE; ++w) t += s_w[w];
    cc[c] = t;
  }
}

// Per partition: the total of its chunks (cnt[q]) and, when cpos is set,
// each chunk's first output position base[q] + the earlier chunks' counts.
__global__ __launch_bounds__(GH_BLOCK) void k_gh_chunk_scan(const uint32_t* __restrict__ cc,
                                                            const uint32_t* __restrict__ cb,
                                                            const int64_t* __restrict__ base,
                                                            int64_t* __restrict__ cpos,
                                                            int64_t* __restrict__ cnt) {
  __shared__ uint32_t s_tmp[GH_BLOCK / WAVE];
  const uint32_t q = blockIdx.x;
  const uint32_t c0 = cb[q], c1 = cb[q + 1];
  int64_t run = base ? base[q] : 0;
  int64_t tot_all = 0;
  for (uint32_t b0 = c0; b0 < c1; b0 += GH_BLOCK) {
    const uint32_t c = b0 + threadIdx.x;
    const uint32_t x = c < c1 ? cc[c] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<GH_BLOCK / WAVE>(x, s_tmp, &tot);
    if (cpos && c < c1) cpos[c] = run + int64_t(ex);
    run += int64_t(tot);
    tot_all += int64_t(tot);
    __syncthreads();
  }
  if (cnt && threadIdx.x == 0) cnt[q] = tot_all;
}

// One workgroup per chunk: its occupied keys in slot order from
// ok[cpos[chunk]] (thread t owns 16 consecutive slots), bounded by the
// shard piece's end lim[q] (the reference panics past it).
__global__ __launch_bounds__(GH_BLOCK) void k_gh_occ_fill(const int64_t* __restrict__ tk,
                                                          const uint8_t* __restrict__ used,
                                                          const int64_t* __restrict__ toff,
                                                          const uint32_t* __restrict__ cb,
                                                          uint32_t p,
                                                          const int64_t* __restrict__ cpos,
                                                          const int64_t* __restrict__ lim,
                                                          int64_t* __restrict__ ok,
                                                          uint32_t* __restrict__ err) {
  __shared__ uint32_t s_cb[GH_PART_MAX + 1];
  __shared__ uint32_t s_tmp[GH_BLOCK / WAVE];
  for (uint32_t q = threadIdx.x; q <= p; q += GH_BLOCK) s_cb[q] = cb[q];
  __syncthreads();
  const uint32_t c = blockIdx.x, q = gh_chunk_part(s_cb, p, c);
  const int64_t s0 = toff[q] + int64_t(c - s_cb[q]) * GH_CH;
  const int64_t s1 = s0 + GH_CH < toff[q + 1] ? s0 + GH_CH : toff[q + 1];
  const int64_t my = s0 + int64_t(threadIdx.x) * GH_CH_IPT;
  uint32_t mine = 0;
#pragma unroll
  for (int j = 0; j < GH_CH_IPT; ++j) {
    const int64_t s = my + j;
    mine |= uint32_t(s < s1 && used[s] != 0) << j;
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan<GH_BLOCK / WAVE>(uint32_t(__popc(mine)), s_tmp, &tot);
  int64_t pos = cpos[c] + int64_t(ex);
  const int64_t end = lim[q];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < GH_CH_IPT; ++j) {
    if (!((mine >> j) & 1u)) continue;
    if (pos < 0 || pos >= end)
      bad = true;
    else
      ok[pos] = tk[my + j];
    ++pos;
  }
  if (bad) atomicOr(err, 16u);
}

template <class T>
std::vector<T> host_copy(const T* p, size_t count, hipStream_t s) {
  std::vector<T> h(count);
  if (count == 0) return h;
  if (is_device_ptr(p)) {
    NFH_CHECK(hipMemcpyAsync(h.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    NFH_CHECK(hipStreamSynchronize(s));
  } else {
    std::copy(p, p + count, h.begin());
  }
  return h;
}

int32_t err_rc(uint32_t e) {
  if (e & (1u | 2u | 16u)) return BAD_RANGE;
  if (e & (4u | 8u)) return MALFORMED;
  return 0;
}

}  // namespace
}  // namespace nfh

using namespace nfh;

// lib.rs:2266-2287: fp[i] = pair_fp(k[i], v[i]) & pmask.
NF_EXPORT int32_t nf_ghash_fp_i64(const int64_t* k, const int64_t* v, size_t n, int64_t* fp,
                                  int64_t pmask, size_t nthreads) {
  (void)nthreads;
  if (n == 0) return 0;
  if (!k || !v || !fp) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int64_t* dk = st.in(k, n, S_IN0);
  const int64_t* dv = st.in(v, n, S_IN1);
  int64_t* dfp = st.out(fp, n, S_OUT0);
  {
    ProfScope prof("ghash_fp", st.s, double(n) * 24.0);
    hipLaunchKernelGGL(k_gh_fp, dim3(grid_for(n, GH_BLOCK, 16 * num_cus())), dim3(GH_BLOCK), 0,
                       st.s, dk, dv, uint64_t(n), dfp, uint64_t(pmask));
    NFH_LAUNCH_CHECK();
  }
  st.finish();
  return 0;
  NFH_GUARD_END(int32_t)
}

// lib.rs:2299-2316: tc[tt * p + q] += rows of chunk tt with fp == q.
NF_EXPORT int32_t nf_ghash_count_i64(const int64_t* fp, size_t n, int64_t* tc, size_t t, size_t p,
                                     size_t nthreads) {
  if (n == 0) return 0;
  if (!fp || !tc) return NULL_OR_ABORT;
  if (!gh_pow2(p) || t == 0 || t > size_t(GH_THREADS_MAX)) return MALFORMED;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const ChunkTiles ct = chunk_tiles(n, nthreads, t);
  const int64_t* dfp = st.in(fp, n, S_IN0);
  int64_t* dtc = st.out(tc, t * p, S_OUT0, true);
  uint32_t* first = scratch_t<uint32_t>(S_T0, ct.nt + 1 + 4);
  uint32_t* err = first + ct.nt + 1;
  NFH_CHECK(hipMemcpyAsync(first, ct.first.data(), (ct.nt + 1) * 4, hipMemcpyHostToDevice, st.s));
  fill_u32(err, 0, 1, st.s);
  const uint32_t tiles = ct.first[ct.nt];
  if (tiles > 0) {
    ProfScope prof("ghash_count", st.s, double(n) * 8.0);
    hipLaunchKernelGGL(k_gh_count, dim3(tiles), dim3(GH_BLOCK), 0, st.s, dfp,
                       TileGeom{first, ct.nt, ct.cs, uint64_t(n)}, uint32_t(p),
                       reinterpret_cast<unsigned long long*>(dtc), err);
    NFH_LAUNCH_CHECK();
  }
  const uint32_t e = read_scalar(err, st.s);
  st.finish();
  return err_rc(e);
  NFH_GUARD_END(int32_t)
}

// lib.rs:2334-2360: chunk tt writes its rows, in row order, from
// tpos[tt * p + fp] on.
NF_EXPORT int32_t nf_ghash_scatter_i64(const int64_t* k, const int64_t* v, const int64_t* fp,
                                       size_t n, const int64_t* tpos, size_t t, size_t p,
                                       int64_t* pk, int64_t* pv, size_t nthreads) {
  if (n == 0) return 0;
  if (!k || !v || !fp || !tpos || !pk || !pv) return NULL_OR_ABORT;
  if (!gh_pow2(p) || t == 0 || t > size_t(GH_THREADS_MAX)) return MALFORMED;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const ChunkTiles ct = chunk_tiles(n, nthreads, t);
  const int64_t* dk = st.in(k, n, S_IN0);
  const int64_t* dv = st.in(v, n, S_IN1);
  const int64_t* dfp = st.in(fp, n, S_IN2);
  const int64_t* dtpos = st.in(tpos, t * p, S_IN3);
  int64_t* dpk = st.out(pk, n, S_OUT0, true);
  int64_t* dpv = st.out(pv, n, S_OUT1, true);
  const uint32_t tiles = ct.first[ct.nt];
  uint32_t* first = scratch_t<uint32_t>(S_T0, ct.nt + 1 + 4);
  uint32_t* err = first + ct.nt + 1;
  uint32_t* th = scratch_t<uint32_t>(S_T1, size_t(tiles) * p + 1);
  int64_t* tb = scratch_t<int64_t>(S_T2, size_t(tiles) * p + 1);
  NFH_CHECK(hipMemcpyAsync(first, ct.first.data(), (ct.nt + 1) * 4, hipMemcpyHostToDevice, st.s));
  fill_u32(err, 0, 1, st.s);
  const TileGeom g{first, ct.nt, ct.cs, uint64_t(n)};
  hipLaunchKernelGGL(k_gh_tile_hist, dim3(tiles), dim3(GH_BLOCK), 0, st.s, dfp, g, uint32_t(p),
                     tiles, th, err);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gh_tile_base, dim3(ct.nt * uint32_t(p)), dim3(GH_BLOCK), 0, st.s, th, first,
                     uint32_t(p), tiles, dtpos, tb);
  NFH_LAUNCH_CHECK();
  {
    ProfScope prof("ghash_scatter", st.s, double(n) * 40.0);
    if (lds_rank_supported())
      hipLaunchKernelGGL(k_gh_scatter<true>, dim3(tiles), dim3(GH_BLOCK), 0, st.s, dk, dv, dfp, g,
                         uint32_t(p), tiles, tb, dpk, dpv, err);
    else
      hipLaunchKernelGGL(k_gh_scatter<false>, dim3(tiles), dim3(GH_BLOCK), 0, st.s, dk, dv, dfp,
                         g, uint32_t(p), tiles, tb, dpk, dpv, err);
    NFH_LAUNCH_CHECK();
  }
  const uint32_t e = read_scalar(err, st.s);
  st.finish();
  return err_rc(e);
  NFH_GUARD_END(int32_t)
}

// lib.rs:2374-2409: each partition's rows inserted in row order into its
// linear-probing table (same layout as the sequential insert).
NF_EXPORT int32_t nf_ghash_pins_i64(const int64_t* pk, const int64_t* pv, const int64_t* starts,
                                    const int64_t* toff, const int64_t* tm, int64_t* tk,
                                    int64_t* tv, uint8_t* used, size_t p, size_t nthreads) {
  (void)nthreads;
  if (!pk || !pv || !starts || !toff || !tm || !tk || !tv || !used) return NULL_OR_ABORT;
  if (!gh_pow2(p)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const std::vector<int64_t> hs = host_copy(starts, p + 1, st.s);
  const std::vector<int64_t> ho = host_copy(toff, p + 1, st.s);
  const std::vector<int64_t> hm = host_copy(tm, p, st.s);
  for (size_t q = 0; q < p; ++q) {
    const uint64_t slots = uint64_t(hm[q]) + 1;
    if (hm[q] < 0 || (slots & (slots - 1)) != 0 || hs[q] < 0 || hs[q] > hs[q + 1] || ho[q] < 0 ||
        ho[q] + int64_t(slots) > ho[q + 1])
      return MALFORMED;
  }
  const uint64_t n = uint64_t(hs[p]), m = uint64_t(ho[p]);
  if (n >= uint64_t(INT32_MAX) || m > (uint64_t(1) << 34)) return MALFORMED;
  if (n == 0) return 0;  // no rows: the tables are left as they are
  const int64_t* dpk = st.in(pk, n, S_IN0);
  const int64_t* dpv = st.in(pv, n, S_IN1);
  int64_t* dgeo = scratch_t<int64_t>(S_T0, 3 * p + 2 + 2);
  NFH_CHECK(hipMemcpyAsync(dgeo, hs.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  NFH_CHECK(hipMemcpyAsync(dgeo + p + 1, ho.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  NFH_CHECK(hipMemcpyAsync(dgeo + 2 * p + 2, hm.data(), p * 8, hipMemcpyHostToDevice, st.s));
  uint32_t* err = reinterpret_cast<uint32_t*>(dgeo + 3 * p + 2);
  int64_t* dtk = st.out(tk, m, S_OUT0, true);
  int64_t* dtv = st.out(tv, m, S_OUT1, true);
  uint8_t* dused = st.out(used, m, S_OUT2, true);
  uint32_t* hold = scratch_t<uint32_t>(S_T1, m);
  uint8_t* first = scratch_t<uint8_t>(S_T2, n + 4);
  fill_u32(err, 0, 1, st.s);
  const PinsGeom g{dgeo, dgeo + p + 1, dgeo + 2 * p + 2, uint32_t(p), n};
  const unsigned grow = grid_for(n, GH_BLOCK, 16 * num_cus());
  const unsigned gslot = grid_for(m, GH_BLOCK, 16 * num_cus());
  ProfScope prof("ghash_pins", st.s, double(n) * 16.0 + double(m) * 17.0);
  {
    fill_u32(hold, GH_EMPTY, m, st.s);
    fill_u32(first, 0, (n + 3) / 4, st.s);
    hipLaunchKernelGGL(k_gh_dedupe, dim3(grow), dim3(GH_BLOCK), 0, st.s, dpk, dpv, g, hold, err);
    NFH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gh_first, dim3(gslot), dim3(GH_BLOCK), 0, st.s, hold, m, first);
    NFH_LAUNCH_CHECK();
    // a table with no room for a new pair would make the placement probe
    // forever (as the reference does): refuse before placing
    const uint32_t e0 = read_scalar(err, st.s);
    if (e0 != 0) {
      st.finish();
      return err_rc(e0);
    }
    fill_u32(hold, GH_EMPTY, m, st.s);
    // windows of min(slots, GH_WIN) slots per partition
    std::vector<uint32_t> wb(p + 1, 0);
    for (size_t q = 0; q < p; ++q) {
      const uint64_t slots = uint64_t(hm[q]) + 1;
      wb[q + 1] = wb[q] + uint32_t(slots / std::min<uint64_t>(slots, GH_WIN));
    }
    const uint32_t nwin = wb[p];
    bool windowed = uint64_t(nwin) * GH_EXT * 4 <= (uint64_t(1) << 30);
    if (windowed) {
      int32_t* key = scratch_t<int32_t>(S_T3, n);
      int32_t* perm = scratch_t<int32_t>(S_T4, n);
      uint32_t* t0 = scratch_t<uint32_t>(S_T5, n);
      uint32_t* t1 = scratch_t<uint32_t>(S_WORK0, n);
      int32_t* tp = scratch_t<int32_t>(S_WORK1, n);
      int32_t* skey = scratch_t<int32_t>(S_WORK2, n);
      uint32_t* small = scratch_t<uint32_t>(S_WORK3, (p + 1) + (nwin + 1) + nwin + p * GH_EXT);
      uint32_t* dwb = small;
      uint32_t* wstart = dwb + p + 1;
      uint32_t* nspill = wstart + nwin + 1;
      uint32_t* tmpin = nspill + nwin;
      uint32_t* spill = scratch_t<uint32_t>(S_MISC, size_t(nwin) * GH_EXT);
      NFH_CHECK(hipMemcpyAsync(dwb, wb.data(), (p + 1) * 4, hipMemcpyHostToDevice, st.s));
      fill_u32(nspill, 0, nwin, st.s);
      hipLaunchKernelGGL(k_gh_winkey, dim3(grow), dim3(GH_BLOCK), 0, st.s, dpk, dpv, first, g, dwb,
                         key);
      NFH_LAUNCH_CHECK();
      radix_argsort<int32_t>(key, n, false, perm, t0, t1, tp, skey, st.s);
      hipLaunchKernelGGL(k_gh_winstart, dim3((nwin + 1 + 255) / 256), dim3(256), 0, st.s, skey,
                         uint64_t(n), nwin, wstart);
      NFH_LAUNCH_CHECK();
      const WinGeom wg{g.toff, g.tm, dwb, uint32_t(p), nwin};
      hipLaunchKernelGGL(k_gh_win, dim3(nwin), dim3(GH_WBLOCK), 0, st.s, wg, dpk, dpv, perm,
                         wstart, hold, spill, nspill, err);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_win_fix, dim3(p), dim3(GH_WBLOCK), 0, st.s, wg, dpk, dpv, perm,
                         wstart, hold, spill, nspill, tmpin, err);
      NFH_LAUNCH_CHECK();
      const uint32_t e1 = read_scalar(err, st.s);
      if (e1 & 32u) {  // a cluster longer than the extension: global form
        fill_u32(err, e1 & ~32u, 1, st.s);
        fill_u32(hold, GH_EMPTY, m, st.s);
        windowed = false;
      }
    }
    if (!windowed) {
      hipLaunchKernelGGL(k_gh_place, dim3(grow), dim3(GH_BLOCK), 0, st.s, dpk, dpv, first, g,
                         hold, err);
      NFH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_gh_materialize, dim3(gslot), dim3(GH_BLOCK), 0, st.s, hold, m, dpk, dpv,
                       dtk, dtv, dused);
    NFH_LAUNCH_CHECK();
  }
  const uint32_t e = read_scalar(err, st.s);
  st.finish();
  return err_rc(e);
  NFH_GUARD_END(int32_t)
}

// lib.rs:2419-2438: cnt[q] = used slots of partition q's window.
NF_EXPORT int32_t nf_ghash_occ_count(const uint8_t* used, const int64_t* toff, size_t p,
                                     int64_t* cnt, size_t nthreads) {
  (void)nthreads;
  if (!used || !toff || !cnt) return NULL_OR_ABORT;
  if (!gh_pow2(p)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const std::vector<int64_t> ho = host_copy(toff, p + 1, st.s);
  for (size_t q = 0; q < p; ++q)
    if (ho[q] < 0 || ho[q] > ho[q + 1]) return MALFORMED;
  const uint64_t m = uint64_t(ho[p]);
  const uint8_t* du = m ? st.in(used, m, S_IN0) : nullptr;
  std::vector<uint32_t> hcb(p + 1, 0);
  for (size_t q = 0; q < p; ++q)
    hcb[q + 1] = hcb[q] + uint32_t((ho[q + 1] - ho[q] + GH_CH - 1) / GH_CH);
  const uint32_t nch = hcb[p];
  int64_t* dtoff = scratch_t<int64_t>(S_T0, p + 1);
  uint32_t* dcb = scratch_t<uint32_t>(S_T1, (p + 1) + size_t(nch) + 1);
  uint32_t* cc = dcb + p + 1;
  NFH_CHECK(hipMemcpyAsync(dtoff, ho.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  NFH_CHECK(hipMemcpyAsync(dcb, hcb.data(), (p + 1) * 4, hipMemcpyHostToDevice, st.s));
  int64_t* dcnt = st.out(cnt, p, S_OUT0);
  {
    ProfScope prof("ghash_occ_count", st.s, double(m));
    if (nch > 0) {
      hipLaunchKernelGGL(k_gh_chunk_count, dim3(nch), dim3(GH_BLOCK), 0, st.s, du, dtoff, dcb,
                         uint32_t(p), cc);
      NFH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_gh_chunk_scan, dim3(p), dim3(GH_BLOCK), 0, st.s, cc, dcb, nullptr,
                       nullptr, dcnt);
    NFH_LAUNCH_CHECK();
  }
  st.finish();
  return 0;
  NFH_GUARD_END(int32_t)
}

// lib.rs:2448-2470: shard tt of the partitions writes its occupied keys
// in slot order from ok[ostart[first partition of the shard]] on.
NF_EXPORT int32_t nf_ghash_occ_fill(const int64_t* tk, const uint8_t* used, const int64_t* toff,
                                    const int64_t* ostart, int64_t* ok, size_t p,
                                    size_t nthreads) {
  if (!tk || !used || !toff || !ostart || !ok) return NULL_OR_ABORT;
  if (!gh_pow2(p)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const std::vector<int64_t> ho = host_copy(toff, p + 1, st.s);
  const std::vector<int64_t> hos = host_copy(ostart, p + 1, st.s);
  for (size_t q = 0; q < p; ++q)
    if (ho[q] < 0 || ho[q] > ho[q + 1]) return MALFORMED;
  const uint64_t m = uint64_t(ho[p]);
  const int64_t nocc = hos[p];
  // the reference carves ok[0, nocc) into consecutive per-shard pieces of
  // ostart[hi] - ostart[lo]; a carve past the end panics there
  const size_t t = gh_fanout(nthreads, p), cs = (p + t - 1) / t;
  if (hos[0] < 0) return BAD_RANGE;
  for (size_t plo = 0; plo < p; plo += cs) {
    const size_t phi = std::min(plo + cs, p);
    if (hos[phi] < hos[plo]) return BAD_RANGE;
  }
  const uint8_t* du = m ? st.in(used, m, S_IN1) : nullptr;
  const int64_t* dtk = m ? st.in(tk, m, S_IN0) : nullptr;
  std::vector<uint32_t> hcb(p + 1, 0);
  for (size_t q = 0; q < p; ++q)
    hcb[q + 1] = hcb[q] + uint32_t((ho[q + 1] - ho[q] + GH_CH - 1) / GH_CH);
  const uint32_t nch = hcb[p];
  int64_t* dtoff = scratch_t<int64_t>(S_T0, 4 * p + 4 + 2);
  int64_t* dbase = dtoff + p + 1;
  int64_t* dlim = dbase + p + 1;
  int64_t* dcnt = dlim + p + 1;
  uint32_t* err = reinterpret_cast<uint32_t*>(dcnt + p + 1);
  uint32_t* dcb = scratch_t<uint32_t>(S_T1, (p + 1) + size_t(nch) + 1);
  uint32_t* cc = dcb + p + 1;
  int64_t* cpos = scratch_t<int64_t>(S_T2, size_t(nch) + 1);
  NFH_CHECK(hipMemcpyAsync(dtoff, ho.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  NFH_CHECK(hipMemcpyAsync(dcb, hcb.data(), (p + 1) * 4, hipMemcpyHostToDevice, st.s));
  // per-chunk and per-partition counts, then each shard's running bases
  if (nch > 0) {
    hipLaunchKernelGGL(k_gh_chunk_count, dim3(nch), dim3(GH_BLOCK), 0, st.s, du, dtoff, dcb,
                       uint32_t(p), cc);
    NFH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_gh_chunk_scan, dim3(p), dim3(GH_BLOCK), 0, st.s, cc, dcb, nullptr, nullptr,
                     dcnt);
  NFH_LAUNCH_CHECK();
  const std::vector<int64_t> hc = host_copy(dcnt, p, st.s);
  std::vector<int64_t> hb(p + 1, 0), hl(p + 1, 0);
  for (size_t q = 0; q < p; ++q) {
    const size_t plo = (q / cs) * cs, phi = std::min(plo + cs, p);
    hb[q] = q == plo ? hos[plo] - hos[0] : hb[q - 1] + hc[q - 1];
    hl[q] = hos[phi] - hos[0];
  }
  NFH_CHECK(hipMemcpyAsync(dbase, hb.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  NFH_CHECK(hipMemcpyAsync(dlim, hl.data(), (p + 1) * 8, hipMemcpyHostToDevice, st.s));
  hipLaunchKernelGGL(k_gh_chunk_scan, dim3(p), dim3(GH_BLOCK), 0, st.s, cc, dcb, dbase, cpos,
                     nullptr);
  NFH_LAUNCH_CHECK();
  int64_t* dok = st.out(ok, size_t(std::max<int64_t>(nocc, 1)), S_OUT0, nocc > 0);
  st.set_out_bytes(ok, size_t(std::max<int64_t>(nocc, 0)) * 8);
  fill_u32(err, 0, 1, st.s);
  if (nch > 0) {
    ProfScope prof("ghash_occ_fill", st.s, double(m) * 9.0);
    hipLaunchKernelGGL(k_gh_occ_fill, dim3(nch), dim3(GH_BLOCK), 0, st.s, dtk, du, dtoff, dcb,
                       uint32_t(p), cpos, dlim, dok, err);
    NFH_LAUNCH_CHECK();
  }
  const uint32_t e = read_scalar(err, st.s);
  st.finish();
  return err_rc(e);
  NFH_GUARD_END(int32_t)
}
