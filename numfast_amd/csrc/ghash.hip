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
// without a sequential loop, by a priority claim: each distinct pair's
// first row claims slots from its home with an atomic minimum of its row
// id, carrying any larger id it evicts onward.  At rest every slot between
// a pair's home and its slot holds a smaller row id and its own slot does
// not, which is exactly where sequential insertion in row order puts it
// (by induction on the row id), so the layout is identical.  The claim
// runs in LDS, one window of the table at a time, after the rows are
// scattered by window (section 2 below); a global form over the table in
// HBM (dedupe, first-row flags, place, materialise) covers what the
// windows cannot.
// Refusals where the reference would index out of bounds (a panic, which
// aborts the process) or probe forever: -2 (partition id or reservation
// outside its buffer), -3 (a table with no free slot left for a new pair).
#include <algorithm>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "runtime.hpp"

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

// One workgroup per tile (tiles in XCD order): stable in-tile ranks per
// partition (item j of lane l of wave w is row r0 + w*64*IPT + j*64 + l;
// per-wave LDS counters give ranks in that order), then the tile's pairs
// are staged in LDS in (partition, rank) order, keys first and then
// values, and stored from there, so consecutive lanes store consecutive
// rows of one partition's run.
constexpr int GH_SCAT_BLOCK = 512;
constexpr int GH_SCAT_IPT = GH_TILE / GH_SCAT_BLOCK;
static_assert(GH_SCAT_IPT * GH_SCAT_BLOCK == GH_TILE, "scatter tiles are count tiles");
template <bool LR>
__global__ __launch_bounds__(GH_SCAT_BLOCK) void k_gh_scatter(
    const int64_t* __restrict__ k, const int64_t* __restrict__ v, const int64_t* __restrict__ fp,
    TileGeom g, uint32_t p, uint32_t tiles, const int64_t* __restrict__ tb,
    int64_t* __restrict__ pk,
    int64_t* __restrict__ pv, uint32_t* __restrict__ err) {
  constexpr int NW = GH_SCAT_BLOCK / WAVE;
  constexpr int IPT = GH_SCAT_IPT;
  __shared__ uint32_t s_cnt[NW][GH_PART_MAX];
  __shared__ uint32_t s_off[GH_PART_MAX];  // first staged position of each partition
  __shared__ int64_t s_base[GH_PART_MAX];  // output row of each partition's run
  __shared__ uint32_t s_dst[GH_TILE];      // output row of each staged position
  __shared__ int64_t s_stage[GH_TILE];
  __shared__ uint32_t s_tmp[NW];
  const int lane = lane_id(), wave = wave_id();
  for (int i = threadIdx.x; i < NW * GH_PART_MAX; i += GH_SCAT_BLOCK) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  uint32_t c;
  uint64_t r0, r1;
  g.locate(tile, &c, &r0, &r1);
  const uint64_t wb = r0 + uint64_t(wave) * (WAVE * IPT) + uint64_t(lane);
  uint32_t d[IPT], rk[IPT];
  uint32_t live = 0;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint64_t i = wb + uint64_t(j * WAVE);
    const int64_t f = i < r1 ? fp[i] : -1;
    const bool ok = f >= 0 && uint64_t(f) < p;  // out-of-range ids: flagged by the histogram
    d[j] = ok ? uint32_t(f) : 0u;
    live |= uint32_t(ok) << j;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
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
  // per partition: exclusive offsets of the waves' runs, the tile's count,
  // its first staged position and its output row
  uint32_t tot = 0;
  if (threadIdx.x < p) {
    const uint32_t q = threadIdx.x;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t x = s_cnt[w][q];
      s_cnt[w][q] = tot;
      tot += x;
    }
    s_base[q] = tb[size_t(q) * tiles + tile];
  }
  uint32_t all;
  const uint32_t off = block_excl_scan<NW>(tot, s_tmp, &all);
  if (threadIdx.x < p) s_off[threadIdx.x] = off;
  __syncthreads();
  bool bad = false;
  uint32_t pos[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (!((live >> j) & 1u)) continue;
    const uint64_t i = wb + uint64_t(j * WAVE);
    const uint32_t r = s_cnt[wave][d[j]] + rk[j];
    pos[j] = s_off[d[j]] + r;
    const int64_t lanepos = s_base[d[j]] + int64_t(r);
    const bool in = lanepos >= 0 && uint64_t(lanepos) < g.n;
    bad |= !in;
    s_dst[pos[j]] = in ? uint32_t(lanepos) : GH_EMPTY;
    s_stage[pos[j]] = k[i];
  }
  if (bad) atomicOr(err, 2u);
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < all; j += GH_SCAT_BLOCK) {
    const uint32_t dst = s_dst[j];
    if (dst != GH_EMPTY) pk[dst] = s_stage[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if ((live >> j) & 1u) s_stage[pos[j]] = v[wb + uint64_t(j * WAVE)];
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < all; j += GH_SCAT_BLOCK) {
    const uint32_t dst = s_dst[j];
    if (dst != GH_EMPTY) pv[dst] = s_stage[j];
  }
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

// The global form.  1. dedupe: each distinct pair of a partition ends in
// one slot of `hold` holding its smallest row id.  Any layout will do here, so the home is
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
    if (j < uint64_t(s_starts[0])) continue;  // before the first partition
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

// 2. place, in LDS windows (the form used whenever it fits; the global
// dedupe above and the global place below are the fallback).  A
// partition's table of S slots is cut into windows of W = min(S, GH_WIN)
// slots, window i covering slots (q + i*W) mod S onward: the reference's
// homes in partition q all sit at q modulo P, so a window starts at a home
// and clusters rarely cross its end.  A pair's window follows from its
// home, so every row of a pair falls in the same window:
//   k_gh_whist counts each 8192-row tile's rows per window of its partition
//     into one entry per (window, tile), laid out window-major within the
//     partition; a chunked scan of the entries gives each its place;
//   k_gh_wscatter writes each row there as a (key, value, row) record,
//     staged in LDS in window order so that the stores are contiguous runs;
//   k_gh_wsolve, one workgroup per window, inserts the window's records
//     into an LDS hash table keyed by a 64-bit pair hash (compare-and-swap
//     to claim, atomic minimum of the row id), then checks every record
//     that joined a slot against the claimer's pair, so each distinct pair
//     keeps its first row; ranks each pair among its home's pairs in row
//     order; runs the priority claim over the window's slots plus GH_EXT
//     extension slots (atomic minimum of (row id, record), evicted ids
//     carried on), each pair starting `rank` slots past its home; and
//     writes used / tk / tv of the window's occupied slots from the
//     records.  Ids that end in the extension spill into the next window
//     (cyclically), where they start at the first slot: every slot they
//     passed holds a smaller id.  A window of the whole table wraps inside
//     LDS and never spills;
//   k_gh_wfix, one workgroup per partition of several windows, re-solves
//     in cyclic order each window whose predecessor's spill list changed.
// Spills only add ids to a window, so its occupied slots only grow and a
// re-solve overwrites every slot the first solve wrote.  A cluster longer
// than the extension, a window with more distinct pairs than its LDS
// table holds, or two pairs with one hash sets bit 32 of err and the
// caller runs the global form, which rewrites every occupied slot.  The
// windowed form runs only when no partition has more rows than slots, so
// it never meets a full table.
constexpr int GH_WIN = 2048;             // window slots
constexpr int GH_EXT = 1024;             // extension slots
constexpr int GH_WTHREADS = 512;         // threads of the tile kernels
constexpr int GH_WIPT = 16;
constexpr uint32_t GH_WTILE = uint32_t(GH_WTHREADS) * GH_WIPT;  // 8192 rows
constexpr uint32_t GH_WQ_MAX = 4096;     // windows per partition
constexpr int GH_SBLOCK = 512;           // threads of the window solve
constexpr uint32_t GH_SD = 2048;         // LDS hash table slots of the window solve
constexpr int GH_SU = 2;                 // records per thread per pass
constexpr unsigned long long GH_NONE = ~0ull;
constexpr uint32_t GH_ERR_WIN = 32u;     // fall back to the global form

// Mixed 64-bit hash of a pair for the LDS table (its low bits pick the
// slot, its high half is the tag).
__device__ __forceinline__ uint64_t gh_dedupe_hash(int64_t k, int64_t v) {
  uint64_t h = gh_pair_fp(k, v);
  h ^= h >> 31;
  h *= 0x7FB5D329728EA185ull;
  h ^= h >> 27;
  h *= 0x81DADEF4BC2DD44Dull;
  h ^= h >> 33;
  return h;
}

struct WPart {
  uint32_t tb;      // first tile
  uint32_t e0;      // entry of (window 0, tile 0); (window i, tile t) at e0 + i * nt + t
  uint32_t nt;      // tiles
  uint32_t nwq;     // windows
  uint32_t r0, r1;  // rows
  uint32_t lw;      // log2 of the window size
  uint32_t pad;
};

struct WTile {
  uint32_t r0, r1;  // rows
  uint32_t e0;      // entry of (window 0, this tile); window i at e0 + i * nt
  uint32_t nt;      // tiles of the partition
  uint32_t q;       // partition
  uint32_t nwq;     // windows of the partition
  uint32_t lw;      // log2 of the window size
  uint32_t pad;
};

// tiles[b] from the per-partition table: tile b belongs to the last
// partition whose first tile is at most b (partitions without rows own no
// tile and never are that last one).
__global__ void k_gh_wtiles(const WPart* __restrict__ parts, uint32_t p, uint32_t ntl,
                            WTile* __restrict__ tiles) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ntl) return;
  uint32_t lo = 0, hi = p;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (parts[mid].tb <= b)
      lo = mid;
    else
      hi = mid;
  }
  const WPart w = parts[lo];
  const uint32_t t = b - w.tb;
  const uint32_t r = w.r0 + t * GH_WTILE;
  tiles[b] = WTile{r, min(r + GH_WTILE, w.r1), w.e0 + t, w.nt, lo, w.nwq, w.lw, 0};
}

// Window (within its partition) of the reference home of pair (k, v).
__device__ __forceinline__ uint32_t gh_win_of(int64_t k, int64_t v, uint64_t mask, uint32_t q,
                                              uint32_t lw) {
  const uint64_t h = gh_pair_fp(k, v) & mask;
  return uint32_t(((h - q) & mask) >> lw);
}

__global__ __launch_bounds__(GH_WTHREADS) void k_gh_whist(const int64_t* __restrict__ pk,
                                                          const int64_t* __restrict__ pv,
                                                          const WTile* __restrict__ tiles,
                                                          const int64_t* __restrict__ tm,
                                                          uint32_t* __restrict__ ecnt) {
  __shared__ uint32_t s_cnt[GH_WQ_MAX];
  const WTile t = tiles[blockIdx.x];
  const uint64_t mask = uint64_t(tm[t.q]);
  for (uint32_t b = threadIdx.x; b < t.nwq; b += GH_WTHREADS) s_cnt[b] = 0;
  __syncthreads();
  int64_t k[GH_WIPT], v[GH_WIPT];
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    k[u] = r < t.r1 ? pk[r] : 0;
    v[u] = r < t.r1 ? pv[r] : 0;
  }
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    if (r < t.r1) atomicAdd(&s_cnt[gh_win_of(k[u], v[u], mask, t.q, t.lw)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < t.nwq; b += GH_WTHREADS) ecnt[t.e0 + b * t.nt] = s_cnt[b];
}

// Entry scan in three steps over chunks of GH_ECHUNK entries (off[np] =
// all records).
constexpr int GH_SCAN = 1024;
constexpr uint32_t GH_ECHUNK = uint32_t(GH_SCAN) * 4;
__global__ __launch_bounds__(GH_SCAN) void k_gh_esum(const uint32_t* __restrict__ ecnt,
                                                     uint32_t np, uint32_t* __restrict__ csum) {
  __shared__ uint32_t s_tmp[GH_SCAN / WAVE];
  const uint32_t b = blockIdx.x * GH_ECHUNK;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t e = b + threadIdx.x + uint32_t(i) * GH_SCAN;
    c += e < np ? ecnt[e] : 0u;
  }
  uint32_t tot;
  (void)block_excl_scan<GH_SCAN / WAVE>(c, s_tmp, &tot);
  if (threadIdx.x == 0) csum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(GH_SCAN) void k_gh_etop(uint32_t* __restrict__ csum, uint32_t nc,
                                                     uint32_t* __restrict__ off, uint32_t np) {
  __shared__ uint32_t s_tmp[GH_SCAN / WAVE];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < nc; b += GH_SCAN) {
    const uint32_t c = b + threadIdx.x;
    const uint32_t x = c < nc ? csum[c] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<GH_SCAN / WAVE>(x, s_tmp, &tot);
    if (c < nc) csum[c] = s_carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) off[np] = s_carry;
}
__global__ __launch_bounds__(GH_SCAN) void k_gh_escan(const uint32_t* __restrict__ ecnt,
                                                      uint32_t np,
                                                      const uint32_t* __restrict__ cbase,
                                                      uint32_t* __restrict__ off) {
  __shared__ uint32_t s_tmp[GH_SCAN / WAVE];
  const uint32_t b = blockIdx.x * GH_ECHUNK + threadIdx.x * 4;
  uint32_t c[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c[i] = b + i < np ? ecnt[b + i] : 0u;
    sum += c[i];
  }
  uint32_t tot;
  uint32_t run = cbase[blockIdx.x] + block_excl_scan<GH_SCAN / WAVE>(sum, s_tmp, &tot);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (b + i < np) off[b + i] = run;
    run += c[i];
  }
}

// Records of a tile: each row's rank within (its window, this tile) comes
// from an LDS counter; the tile's rows are staged in LDS in window order,
// one array at a time, and stored from there, so consecutive lanes store
// consecutive records of one (window, tile) run.
__global__ __launch_bounds__(GH_WTHREADS) void k_gh_wscatter(
    const int64_t* __restrict__ pk, const int64_t* __restrict__ pv,
    const WTile* __restrict__ tiles, const int64_t* __restrict__ tm,
    const uint32_t* __restrict__ off, int64_t* __restrict__ rk, int64_t* __restrict__ rv,
    uint32_t* __restrict__ rr) {
  constexpr int PER = GH_WQ_MAX / GH_WTHREADS;
  __shared__ uint32_t s_cnt[GH_WQ_MAX];  // counts, then the tile's window offsets
  __shared__ uint32_t s_gb[GH_WQ_MAX];   // first record of (window, this tile)
  __shared__ uint32_t s_dst[GH_WTILE];   // record of each staged position
  __shared__ unsigned long long s_stage[GH_WTILE];
  __shared__ uint32_t s_tmp[GH_WTHREADS / WAVE];
  const WTile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const uint64_t mask = uint64_t(tm[t.q]);
  for (uint32_t b = threadIdx.x; b < t.nwq; b += GH_WTHREADS) s_cnt[b] = 0;
  __syncthreads();
  int64_t k[GH_WIPT], v[GH_WIPT];
  uint32_t wi[GH_WIPT], pos[GH_WIPT];
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    k[u] = r < t.r1 ? pk[r] : 0;
    v[u] = r < t.r1 ? pv[r] : 0;
  }
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    wi[u] = gh_win_of(k[u], v[u], mask, t.q, t.lw);
    pos[u] = r < t.r1 ? atomicAdd(&s_cnt[wi[u]], 1u) : 0u;
  }
  __syncthreads();
  // exclusive scan of the counts: thread i owns windows [i*PER, i*PER+PER)
  uint32_t c[PER], sum = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t b = threadIdx.x * PER + j;
    c[j] = b < t.nwq ? s_cnt[b] : 0u;
    sum += c[j];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<GH_WTHREADS / WAVE>(sum, s_tmp, &tot);
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t b = threadIdx.x * PER + j;
    if (b < t.nwq) {
      s_cnt[b] = run;
      s_gb[b] = off[t.e0 + b * t.nt];
    }
    run += c[j];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    if (r < t.r1) {
      const uint32_t loc = pos[u];
      pos[u] = s_cnt[wi[u]] + loc;
      s_dst[pos[u]] = s_gb[wi[u]] + loc;
      s_stage[pos[u]] = (unsigned long long)k[u];
    }
  }
  __syncthreads();
  const uint32_t nrow = t.r1 - t.r0;
  for (uint32_t j = threadIdx.x; j < nrow; j += GH_WTHREADS) rk[s_dst[j]] = int64_t(s_stage[j]);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    if (r < t.r1) s_stage[pos[u]] = (unsigned long long)v[u];
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nrow; j += GH_WTHREADS) rv[s_dst[j]] = int64_t(s_stage[j]);
  __syncthreads();
  uint32_t* s_row = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
  for (int u = 0; u < GH_WIPT; ++u) {
    const uint32_t r = t.r0 + uint32_t(u) * GH_WTHREADS + threadIdx.x;
    if (r < t.r1) s_row[pos[u]] = r;
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nrow; j += GH_WTHREADS) rr[s_dst[j]] = s_row[j];
}

struct WinGeom {
  const int64_t* toff;       // p + 1
  const int64_t* tm;         // p
  const uint32_t* winbase;   // p + 1: first window of each partition
  const uint4* desc;         // 2 per window: (first record, end record, partition, window
                             // in it), (table mask, table base) as 32-bit halves
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

// desc[w] of every window (records from the scanned entries).
__global__ void k_gh_wdesc(const uint32_t* __restrict__ winbase, uint32_t p, uint32_t nwin,
                           const WPart* __restrict__ parts, const uint32_t* __restrict__ off,
                           const int64_t* __restrict__ tm, const int64_t* __restrict__ toff,
                           uint4* __restrict__ desc) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwin) return;
  const uint32_t q = gh_win_part(winbase, p, w);  // every partition has a window
  const uint32_t wi = w - winbase[q];
  const uint32_t e0 = parts[q].e0, nt = parts[q].nt;
  const uint64_t mask = uint64_t(tm[q]), base = uint64_t(toff[q]);
  desc[2 * w] = make_uint4(off[e0 + wi * nt], off[e0 + (wi + 1) * nt], q, wi);
  desc[2 * w + 1] = make_uint4(uint32_t(mask), uint32_t(mask >> 32), uint32_t(base),
                               uint32_t(base >> 32));
}

// LDS address of window slot i.  A partition's reference homes are
// congruent modulo P (a multiple of the 32 two-bank slots of 64-bit
// entries), so unswizzled every claim of a window would start in one bank
// pair; XOR-ing the low 6 bits with bits 8..13 spreads homes 256 slots
// apart.  It permutes each aligned 64-slot block, so [0, span) maps onto
// itself (span is a multiple of 64, or below 256 where it is the identity)
// and the probe order, which follows slot numbers, is unchanged.
__device__ __forceinline__ uint32_t gh_swz(uint32_t i) { return i ^ ((i >> 8) & 63u); }

union WinTable {
  struct {
    unsigned long long key[GH_SD];  // pair hash; GH_NONE if free
    uint32_t rec[GH_SD];            // record of the slot's claimer
    uint32_t first[GH_SD];          // smallest row id of the pair
  } d;
  unsigned long long hold[GH_WIN + GH_EXT];  // (row id << 32) | record; GH_NONE if free
  uint32_t rank[GH_WIN + 2 * GH_SD];          // home counters, grouped rows, ranks
};
static_assert(sizeof(WinTable) <= 32 * 1024, "four window solves per CU");

struct WinLds {
  WinTable u;
  uint16_t home[GH_SD];  // home slot within the window, set by the claimer
  uint32_t bad, over, changed, start;
  uint32_t tmp[GH_SBLOCK / WAVE];
};

// Solve window w from its records, plus `nin` spilled ids from slot 0.
// Writes used / tk / tv of the window's occupied slots and its
// extension's ids (slot order) to spill[w]; L.changed = whether that list
// differs from before.  False (uniform) when the window cannot be solved
// in LDS (bit 32 of err set).
__device__ __forceinline__ bool gh_wsolve(uint32_t w, const WinGeom& g, const int64_t* __restrict__ rk,
                          const int64_t* __restrict__ rv, const uint32_t* __restrict__ rr,
                          const unsigned long long* __restrict__ spill_in, uint32_t nin,
                          int64_t* __restrict__ tk, int64_t* __restrict__ tv,
                          uint8_t* __restrict__ used, unsigned long long* __restrict__ spill,
                          uint32_t* __restrict__ nspill, WinLds& L,
                          uint32_t* __restrict__ err) {
  constexpr int PER = GH_SD / GH_SBLOCK;
  const uint4 dsc = g.desc[2 * w], dsg = g.desc[2 * w + 1];
  const uint32_t r0 = dsc.x, r1 = dsc.y, q = dsc.z, wi = dsc.w;
  const uint64_t mask = uint64_t(dsg.x) | (uint64_t(dsg.y) << 32);
  const uint64_t base = uint64_t(dsg.z) | (uint64_t(dsg.w) << 32);
  const uint32_t wsz = mask + 1 < uint64_t(GH_WIN) ? uint32_t(mask + 1) : uint32_t(GH_WIN);
  const bool whole = uint64_t(wsz) == mask + 1;
  const uint32_t ext = whole ? 0u : uint32_t(GH_EXT);
  const uint32_t span = wsz + ext;
  const uint64_t a = (uint64_t(q) + uint64_t(wi) * wsz) & mask;  // first slot (cyclic)
  const uint32_t nrec = r1 - r0;
  uint32_t sd = 64;
  while (sd < GH_SD && sd < 2 * nrec) sd *= 2;
  for (uint32_t i = threadIdx.x; i < sd; i += GH_SBLOCK) {
    L.u.d.key[i] = GH_NONE;
    L.u.d.first[i] = GH_EMPTY;
  }
  if (threadIdx.x == 0) {
    L.bad = 0;
    L.over = 0;
  }
  __syncthreads();
  // 1. distinct pairs of the window's records: a record claims a free slot
  // with its pair's hash or joins the slot holding that hash; after each
  // pass every record that joined compares its pair with the claimer's
  // (which never changes), so a hash shared by two pairs is always found
  for (uint32_t i0 = r0; i0 < r1; i0 += GH_SBLOCK * GH_SU) {  // uniform trip count
    int64_t k[GH_SU], v[GH_SU];
    uint32_t row[GH_SU], slot[GH_SU];
#pragma unroll
    for (int u = 0; u < GH_SU; ++u) {
      const uint32_t i = i0 + uint32_t(u) * GH_SBLOCK + threadIdx.x;
      k[u] = i < r1 ? rk[i] : 0;
      v[u] = i < r1 ? rv[i] : 0;
      row[u] = i < r1 ? rr[i] : GH_EMPTY;
    }
#pragma unroll
    for (int u = 0; u < GH_SU; ++u) {
      const uint32_t i = i0 + uint32_t(u) * GH_SBLOCK + threadIdx.x;
      slot[u] = GH_EMPTY;
      if (i >= r1) continue;
      uint64_t h = gh_dedupe_hash(k[u], v[u]);
      if (h == GH_NONE) h = GH_NONE - 1;
      // the probe loop only finds the slot; the claimer's stores and the
      // row minimum follow it
      uint32_t sl = uint32_t(h) & (sd - 1);
      bool claimed = false;
      for (uint32_t probe = 0; probe < sd; ++probe) {
        const unsigned long long old = atomicCAS(&L.u.d.key[sl], GH_NONE, h);
        if (old == GH_NONE || old == h) {
          claimed = old == GH_NONE;
          slot[u] = sl;
          break;
        }
        sl = (sl + 1) & (sd - 1);
      }
      if (slot[u] == GH_EMPTY) {
        L.bad = 1u;  // more distinct pairs than table slots
        continue;
      }
      if (claimed) {
        L.u.d.rec[slot[u]] = i;
        L.home[slot[u]] = uint16_t(((gh_pair_fp(k[u], v[u]) & mask) - a) & mask);
      }
      atomicMin(&L.u.d.first[slot[u]], row[u]);
    }
    __syncthreads();  // every claimer has stored its record
#pragma unroll
    for (int u = 0; u < GH_SU; ++u) {
      const uint32_t i = i0 + uint32_t(u) * GH_SBLOCK + threadIdx.x;
      if (slot[u] != GH_EMPTY) {
        const uint32_t c = L.u.d.rec[slot[u]];
        if (c != i && (rk[c] != k[u] || rv[c] != v[u])) L.bad = 1u;  // two pairs, one hash
      }
    }
  }
  __syncthreads();
  if (L.bad) {  // uniform
    if (threadIdx.x == 0) atomicOr(err, GH_ERR_WIN);
    return false;
  }
  // 2. the distinct pairs to registers
  unsigned long long id[PER];
  uint32_t c0[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const uint32_t j = threadIdx.x + uint32_t(s) * GH_SBLOCK;
    id[s] = GH_NONE;
    c0[s] = 0;
    if (j < sd) {
      const unsigned long long key = L.u.d.key[j];
      if (key != GH_NONE) {
        id[s] = (uint64_t(L.u.d.first[j]) << 32) | L.u.d.rec[j];
        c0[s] = L.home[j];
      }
    }
  }
  __syncthreads();
  // 3. each pair's rank among the window's pairs of the same home, in row
  // order.  In the sequential layout the pairs of one home with smaller
  // rows sit in distinct slots between that home and the pair's slot, so
  // the pair's claim may start `rank` slots past its home: every slot it
  // skips ends with a smaller id.  Without this, all pairs of a home (about
  // a hundred at half load and 256 partitions) would start at one slot and
  // each would pass every smaller one.  Pairs are grouped by home (LDS
  // counters, a scan, a scatter of their rows), then each position of the
  // grouped list counts the smaller rows of its group (consecutive lanes
  // read one group, so the reads are broadcasts).
  {
    uint32_t* s_hc = L.u.rank;                           // per home: count, then first position
    uint32_t* s_rows = s_hc + GH_WIN;                    // rows, grouped by home
    uint32_t* s_rk = s_rows + GH_SD;                     // home, then rank, per position
    for (uint32_t i = threadIdx.x; i < wsz; i += GH_SBLOCK) s_hc[i] = 0;
    __syncthreads();
    uint32_t loc[PER];
#pragma unroll
    for (int s = 0; s < PER; ++s) loc[s] = id[s] != GH_NONE ? atomicAdd(&s_hc[c0[s]], 1u) : 0u;
    __syncthreads();
    constexpr int HPT = GH_WIN / GH_SBLOCK;  // counters per thread
    uint32_t hc[HPT], sum = 0;
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const uint32_t b = threadIdx.x * HPT + j;
      hc[j] = b < wsz ? s_hc[b] : 0u;
      sum += hc[j];
    }
    uint32_t total;
    uint32_t run = block_excl_scan<GH_SBLOCK / WAVE>(sum, L.tmp, &total);
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const uint32_t b = threadIdx.x * HPT + j;
      if (b < wsz) s_hc[b] = run;
      run += hc[j];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PER; ++s)
      if (id[s] != GH_NONE) {
        loc[s] += s_hc[c0[s]];
        s_rows[loc[s]] = uint32_t(id[s] >> 32);
        s_rk[loc[s]] = c0[s];
      }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += GH_SBLOCK) {
      const uint32_t c = s_rk[j], row = s_rows[j];
      const uint32_t b = s_hc[c], e = c + 1 < wsz ? s_hc[c + 1] : total;
      uint32_t r = 0, i = b;
      for (; i + 8 <= e; i += 8) {  // eight independent reads in flight
        uint32_t x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = s_rows[i + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) r += x[k] < row ? 1u : 0u;
      }
      for (; i < e; ++i) r += s_rows[i] < row ? 1u : 0u;
      s_rk[j] = r;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PER; ++s)
      if (id[s] != GH_NONE) c0[s] += s_rk[loc[s]];
    __syncthreads();
  }
  // the LDS becomes the hold
  for (uint32_t i = threadIdx.x; i < span; i += GH_SBLOCK) L.u.hold[i] = GH_NONE;
  __syncthreads();
  // 3. priority claim: an id that meets a smaller one moves on, one that
  // meets a larger one takes the slot and carries the evicted id on
  const uint32_t limit = 4 * span + sd + nin;
  bool over = false;
  auto claim = [&](unsigned long long carried, uint32_t c) {
    for (uint32_t it = 0;; ++it) {
      if (c >= span) {
        if (!whole) {
          over = true;
          break;
        }
        c &= wsz - 1;
      }
      if (it > limit) {  // cannot happen while the table has room
        over = true;
        break;
      }
      const unsigned long long old = atomicMin(&L.u.hold[gh_swz(c)], carried);
      if (old == GH_NONE) break;
      if (old > carried) carried = old;
      ++c;
    }
  };
#pragma unroll
  for (int s = 0; s < PER; ++s)
    if (id[s] != GH_NONE) claim(id[s], c0[s]);
  for (uint32_t t = threadIdx.x; t < nin; t += GH_SBLOCK) claim(spill_in[t], 0);
  if (over) L.over = 1u;
  __syncthreads();
  if (L.over) {  // uniform
    if (threadIdx.x == 0) atomicOr(err, GH_ERR_WIN);
    return false;
  }
  // 4. the window's slots
  constexpr int OPT = GH_WIN / GH_SBLOCK;
#pragma unroll
  for (int j = 0; j < OPT; ++j) {
    const uint32_t i = threadIdx.x + uint32_t(j) * GH_SBLOCK;
    const unsigned long long x = i < wsz ? L.u.hold[gh_swz(i)] : GH_NONE;
    if (x != GH_NONE) {
      const uint32_t o = uint32_t(x);
      const uint64_t s = base + ((a + i) & mask);
      tk[s] = rk[o];
      tv[s] = rv[o];
      used[s] = 1;
    }
  }
  // an extension holding nothing spills nothing (the usual case)
  bool occ = false;
  for (uint32_t i = threadIdx.x; i < ext; i += GH_SBLOCK) occ |= L.u.hold[gh_swz(wsz + i)] != GH_NONE;
  if (!__syncthreads_or(occ)) {
    if (threadIdx.x == 0) {
      L.changed = nspill[w] != 0 ? 1u : 0u;
      nspill[w] = 0;
    }
    __syncthreads();
    return true;
  }
  // extension ids in slot order (one wave per step of 64 slots)
  unsigned long long* out = spill + size_t(w) * GH_EXT;
  const uint32_t prev = nspill[w];
  __syncthreads();
  if (threadIdx.x < WAVE) {
    uint32_t cnt = 0;
    bool diff = false;
    for (uint32_t s0 = 0; s0 < ext; s0 += WAVE) {
      const unsigned long long x = L.u.hold[gh_swz(wsz + s0 + threadIdx.x)];
      const bool occ = x != GH_NONE;
      const uint64_t bal = __ballot(occ);
      if (occ) {
        const uint32_t p = cnt + uint32_t(__popcll(bal & lanemask_lt()));
        if (p >= prev || out[p] != x) diff = true;
        out[p] = x;
      }
      cnt += uint32_t(__popcll(bal));
    }
    diff = __ballot(diff) != 0 || cnt != prev;
    if (threadIdx.x == 0) {
      nspill[w] = cnt;
      L.changed = diff ? 1u : 0u;
    }
  }
  __syncthreads();
  return true;
}

// Phase A: every window alone.
__global__ __launch_bounds__(GH_SBLOCK) void k_gh_wsolve(
    WinGeom g, const int64_t* __restrict__ rk, const int64_t* __restrict__ rv,
    const uint32_t* __restrict__ rr, int64_t* __restrict__ tk, int64_t* __restrict__ tv,
    uint8_t* __restrict__ used, unsigned long long* __restrict__ spill,
    uint32_t* __restrict__ nspill, uint32_t* __restrict__ err) {
  __shared__ WinLds L;
  (void)gh_wsolve(blockIdx.x, g, rk, rv, rr, nullptr, 0, tk, tv, used, spill, nspill, L, err);
}

// Phase B: one workgroup per partition of several windows walks them in
// cyclic order and re-solves each window whose predecessor's spill list
// changed since the window was last solved: first those whose predecessor
// spilled at all, then those behind a re-solve that changed its list.  It
// starts at a window whose predecessor spilled nothing (if any) and stops
// after a full lap with nothing to re-solve; work left after three laps
// sets the overflow flag.
__global__ __launch_bounds__(GH_SBLOCK) void k_gh_wfix(
    WinGeom g, const int64_t* __restrict__ rk, const int64_t* __restrict__ rv,
    const uint32_t* __restrict__ rr, int64_t* __restrict__ tk, int64_t* __restrict__ tv,
    uint8_t* __restrict__ used, unsigned long long* __restrict__ spill,
    uint32_t* __restrict__ nspill, uint32_t* __restrict__ err) {
  __shared__ WinLds L;
  __shared__ uint8_t s_dirty[GH_WQ_MAX];
  const uint32_t q = blockIdx.x;
  const uint32_t w0 = g.winbase[q], nw = g.winbase[q + 1] - w0;
  if (nw <= 1) return;  // one window holds the whole table and never spills
  for (uint32_t i = threadIdx.x; i < nw; i += GH_SBLOCK)
    s_dirty[i] = nspill[w0 + (i + nw - 1) % nw] != 0 ? 1 : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t st = 0;
    while (st < nw && s_dirty[st]) ++st;
    L.start = st == nw ? 0 : st;
  }
  __syncthreads();
  const uint32_t st = L.start;
  uint32_t clean = 0;  // windows passed in a row with nothing to do
  for (uint32_t step = 0; step < 3 * nw && clean < nw; ++step) {
    const uint32_t wi = (st + step) % nw;
    if (!s_dirty[wi]) {  // uniform
      ++clean;
      continue;
    }
    clean = 0;
    const uint32_t pw = w0 + (wi + nw - 1) % nw;
    __syncthreads();  // every thread has read s_dirty[wi]
    if (threadIdx.x == 0) s_dirty[wi] = 0;
    // with two or more windows the predecessor's list is not the one the
    // solve rewrites
    if (!gh_wsolve(w0 + wi, g, rk, rv, rr, spill + size_t(pw) * GH_EXT, nspill[pw], tk, tv,
                   used, spill, nspill, L, err))
      return;
    if (L.changed && threadIdx.x == 0) s_dirty[(wi + 1) % nw] = 1;
    __syncthreads();
  }
  if (clean < nw && threadIdx.x == 0) {  // work left: a spill chain that does not settle
    bool left = false;
    for (uint32_t i = 0; i < nw; ++i) left |= s_dirty[i] != 0;
    if (left) atomicOr(err, GH_ERR_WIN);
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

// Used slots per chunk, GH_CPW chunks per workgroup (slot s0 + j*GH_BLOCK
// + t of a chunk is thread t's j-th, so every load is coalesced).
constexpr int GH_CPW = 4;
__global__ __launch_bounds__(GH_BLOCK) void k_gh_chunk_count(const uint8_t* __restrict__ used,
                                                             const int64_t* __restrict__ toff,
                                                             const uint32_t* __restrict__ cb,
                                                             uint32_t p, uint32_t nch,
                                                             uint32_t* __restrict__ cc) {
  __shared__ uint32_t s_cb[GH_PART_MAX + 1];
  __shared__ uint32_t s_w[GH_CPW][GH_BLOCK / WAVE];
  for (uint32_t q = threadIdx.x; q <= p; q += GH_BLOCK) s_cb[q] = cb[q];
  __syncthreads();
  uint32_t x[GH_CPW];
#pragma unroll
  for (int i = 0; i < GH_CPW; ++i) {
    const uint32_t c = blockIdx.x * GH_CPW + i;
    x[i] = 0;
    if (c >= nch) continue;  // uniform
    const uint32_t q = gh_chunk_part(s_cb, p, c);
    const int64_t s0 = toff[q] + int64_t(c - s_cb[q]) * GH_CH;
    const int64_t s1 = s0 + GH_CH < toff[q + 1] ? s0 + GH_CH : toff[q + 1];
#pragma unroll
    for (int j = 0; j < GH_CH_IPT; ++j) {
      const int64_t s = s0 + j * GH_BLOCK + threadIdx.x;
      x[i] += (s < s1 && used[s] != 0) ? 1u : 0u;
    }
  }
#pragma unroll
  for (int i = 0; i < GH_CPW; ++i) {
    for (int off = 32; off > 0; off >>= 1) x[i] += __shfl_xor(x[i], off);
    if (lane_id() == 0) s_w[i][wave_id()] = x[i];
  }
  __syncthreads();
  if (threadIdx.x < GH_CPW) {
    const uint32_t c = blockIdx.x * GH_CPW + threadIdx.x;
    uint32_t t = 0;
    for (int w = 0; w < GH_BLOCK / WAVE; ++w) t += s_w[threadIdx.x][w];
    if (c < nch) cc[c] = t;
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
// ok[cpos[chunk]], bounded by the shard piece's end lim[q] (the reference
// panics past it).  Slot s0 + j*GH_BLOCK + w*64 + l is lane l of wave w's
// j-th, so slot order is (j, w, l): per (j, w) ballot counts in LDS give
// each group's first position, the ballot its lanes' ranks.
__global__ __launch_bounds__(GH_BLOCK) void k_gh_occ_fill(const int64_t* __restrict__ tk,
                                                          const uint8_t* __restrict__ used,
                                                          const int64_t* __restrict__ toff,
                                                          const uint32_t* __restrict__ cb,
                                                          uint32_t p,
                                                          const int64_t* __restrict__ cpos,
                                                          const int64_t* __restrict__ lim,
                                                          int64_t* __restrict__ ok,
                                                          uint32_t* __restrict__ err) {
  constexpr int NW = GH_BLOCK / WAVE;
  __shared__ uint32_t s_cb[GH_PART_MAX + 1];
  __shared__ uint32_t s_c[GH_CH_IPT * NW];
  for (uint32_t q = threadIdx.x; q <= p; q += GH_BLOCK) s_cb[q] = cb[q];
  __syncthreads();
  const uint32_t c = blockIdx.x, q = gh_chunk_part(s_cb, p, c);
  const int64_t s0 = toff[q] + int64_t(c - s_cb[q]) * GH_CH;
  const int64_t s1 = s0 + GH_CH < toff[q + 1] ? s0 + GH_CH : toff[q + 1];
  const int w = wave_id();
  uint8_t u[GH_CH_IPT];
  int64_t key[GH_CH_IPT];
#pragma unroll
  for (int j = 0; j < GH_CH_IPT; ++j) {
    const int64_t s = s0 + j * GH_BLOCK + threadIdx.x;
    u[j] = s < s1 ? used[s] : 0;
    key[j] = s < s1 ? tk[s] : 0;
  }
  uint64_t bal[GH_CH_IPT];
#pragma unroll
  for (int j = 0; j < GH_CH_IPT; ++j) {
    bal[j] = __ballot(u[j] != 0);
    if (lane_id() == 0) s_c[j * NW + w] = uint32_t(__popcll(bal[j]));
  }
  __syncthreads();
  const int64_t end = lim[q];
  int64_t run = cpos[c];  // first position of group (j, w), walked in slot order
  bool bad = false;
#pragma unroll
  for (int j = 0; j < GH_CH_IPT; ++j) {
#pragma unroll
    for (int v = 0; v < NW; ++v)
      if (v < w) run += s_c[j * NW + v];
    if (u[j] != 0) {
      const int64_t pos = run + int64_t(__popcll(bal[j] & lanemask_lt()));
      if (pos < 0 || pos >= end)
        bad = true;
      else
        ok[pos] = key[j];
    }
#pragma unroll
    for (int v = 0; v < NW; ++v)
      if (v >= w) run += s_c[j * NW + v];
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
      hipLaunchKernelGGL(k_gh_scatter<true>, dim3(tiles), dim3(GH_SCAT_BLOCK), 0, st.s, dk, dv,
                         dfp, g,
                         uint32_t(p), tiles, tb, dpk, dpv, err);
    else
      hipLaunchKernelGGL(k_gh_scatter<false>, dim3(tiles), dim3(GH_SCAT_BLOCK), 0, st.s, dk, dv,
                         dfp,
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
  // windows of min(slots, GH_WIN) slots per partition; the windowed form
  // needs every partition to have no more rows than slots (so no table can
  // run full) and at most GH_WQ_MAX windows
  std::vector<uint32_t> wb(p + 1, 0);
  std::vector<WPart> parts(p);
  bool windowed = true;
  uint64_t ent = 0, ntiles = 0;
  for (size_t q = 0; q < p; ++q) {
    const uint64_t r0 = uint64_t(hs[q]), r1 = uint64_t(hs[q + 1]);
    const uint64_t slots = uint64_t(hm[q]) + 1;
    const uint32_t wsz = uint32_t(std::min<uint64_t>(slots, GH_WIN));
    const uint64_t nwq = std::min<uint64_t>(slots / wsz, GH_WQ_MAX);
    if (r1 - r0 > slots || slots / wsz > GH_WQ_MAX) windowed = false;
    wb[q + 1] = wb[q] + uint32_t(nwq);
    uint32_t lw = 0;
    while ((1u << lw) < wsz) ++lw;
    const uint32_t nt = uint32_t((r1 - r0 + GH_WTILE - 1) / GH_WTILE);
    parts[q] = WPart{uint32_t(ntiles), uint32_t(ent), nt, uint32_t(nwq), uint32_t(r0),
                     uint32_t(r1), lw, 0};
    ntiles += nt;
    ent += uint64_t(nt) * nwq;
  }
  const uint32_t nwin = wb[p];
  windowed = windowed && uint64_t(nwin) * GH_EXT * 8 <= (uint64_t(1) << 31) &&
             ent < (uint64_t(1) << 28) && ntiles != 0;
  // one copy: starts, toff, tm, the error word, the partition table, wb
  const size_t geo_b = (3 * p + 3) * 8, part_b = p * sizeof(WPart), wb_b = (p + 1) * 4;
  std::vector<unsigned char> blob(geo_b + part_b + wb_b, 0);
  std::memcpy(blob.data(), hs.data(), (p + 1) * 8);
  std::memcpy(blob.data() + (p + 1) * 8, ho.data(), (p + 1) * 8);
  std::memcpy(blob.data() + (2 * p + 2) * 8, hm.data(), p * 8);
  std::memcpy(blob.data() + geo_b, parts.data(), part_b);
  std::memcpy(blob.data() + geo_b + part_b, wb.data(), wb_b);
  unsigned char* dblob = static_cast<unsigned char*>(scratch(S_T0, blob.size()));
  NFH_CHECK(hipMemcpyAsync(dblob, blob.data(), blob.size(), hipMemcpyHostToDevice, st.s));
  int64_t* dgeo = reinterpret_cast<int64_t*>(dblob);
  uint32_t* err = reinterpret_cast<uint32_t*>(dgeo + 3 * p + 2);  // zero from the copy
  const WPart* dparts = reinterpret_cast<const WPart*>(dblob + geo_b);
  const uint32_t* dwb = reinterpret_cast<const uint32_t*>(dblob + geo_b + part_b);
  int64_t* dtk = st.out(tk, m, S_OUT0, true);
  int64_t* dtv = st.out(tv, m, S_OUT1, true);
  uint8_t* dused = st.out(used, m, S_OUT2, true);
  const PinsGeom g{dgeo, dgeo + p + 1, dgeo + 2 * p + 2, uint32_t(p), n};
  ProfScope prof("ghash_pins", st.s, double(n) * 16.0 + double(m) * 17.0);
  {
    if (windowed) {
      const uint32_t np = uint32_t(ent), nc = (np + GH_ECHUNK - 1) / GH_ECHUNK;
      const uint32_t ntl = uint32_t(ntiles);
      int64_t* rk = scratch_t<int64_t>(S_T3, n);
      int64_t* rv = scratch_t<int64_t>(S_T4, n);
      uint32_t* rr = scratch_t<uint32_t>(S_T5, n);
      uint32_t* ecnt = scratch_t<uint32_t>(S_WORK0, np);
      uint32_t* eoff = scratch_t<uint32_t>(S_WORK1, size_t(np) + 1);
      unsigned long long* spill = scratch_t<unsigned long long>(S_MISC, size_t(nwin) * GH_EXT);
      unsigned char* aux = static_cast<unsigned char*>(
          scratch(S_WORK2, size_t(ntl) * sizeof(WTile) + size_t(nwin) * 36 +
                               size_t(nc + 1) * 4 + 256));
      uint4* ddesc = reinterpret_cast<uint4*>(aux);
      WTile* dtl = reinterpret_cast<WTile*>(ddesc + 2 * size_t(nwin));
      uint32_t* nspill = reinterpret_cast<uint32_t*>(dtl + ntl);
      uint32_t* csum = nspill + nwin;
      fill_u32(nspill, 0, nwin, st.s);
      hipLaunchKernelGGL(k_gh_wtiles, dim3((ntl + 255) / 256), dim3(256), 0, st.s, dparts,
                         uint32_t(p), ntl, dtl);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_whist, dim3(ntl), dim3(GH_WTHREADS), 0, st.s, dpk, dpv, dtl, g.tm,
                         ecnt);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_esum, dim3(nc), dim3(GH_SCAN), 0, st.s, ecnt, np, csum);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_etop, dim3(1), dim3(GH_SCAN), 0, st.s, csum, nc, eoff, np);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_escan, dim3(nc), dim3(GH_SCAN), 0, st.s, ecnt, np, csum, eoff);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_wscatter, dim3(ntl), dim3(GH_WTHREADS), 0, st.s, dpk, dpv, dtl,
                         g.tm, eoff, rk, rv, rr);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_wdesc, dim3((nwin + 255) / 256), dim3(256), 0, st.s, dwb,
                         uint32_t(p), nwin, dparts, eoff, g.tm, g.toff, ddesc);
      NFH_LAUNCH_CHECK();
      const WinGeom wg{g.toff, g.tm, dwb, ddesc, uint32_t(p), nwin};
      hipLaunchKernelGGL(k_gh_wsolve, dim3(nwin), dim3(GH_SBLOCK), 0, st.s, wg, rk, rv, rr, dtk,
                         dtv, dused, spill, nspill, err);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_wfix, dim3(uint32_t(p)), dim3(GH_SBLOCK), 0, st.s, wg, rk, rv, rr,
                         dtk, dtv, dused, spill, nspill, err);
      NFH_LAUNCH_CHECK();
      if (read_scalar(err, st.s) & GH_ERR_WIN) {  // not solvable in LDS: global form
        fill_u32(err, 0, 1, st.s);
        windowed = false;
      }
    }
    if (!windowed) {
      uint32_t* hold = scratch_t<uint32_t>(S_T1, m);
      uint8_t* first = scratch_t<uint8_t>(S_T2, n + 4);
      const unsigned grow = grid_for(n, GH_BLOCK, 16 * num_cus());
      const unsigned gslot = grid_for(m, GH_BLOCK, 16 * num_cus());
      fill_u32(first, 0, (n + 3) / 4, st.s);
      fill_u32(hold, GH_EMPTY, m, st.s);
      hipLaunchKernelGGL(k_gh_dedupe, dim3(grow), dim3(GH_BLOCK), 0, st.s, dpk, dpv, g, hold,
                         err);
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
      hipLaunchKernelGGL(k_gh_place, dim3(grow), dim3(GH_BLOCK), 0, st.s, dpk, dpv, first, g,
                         hold, err);
      NFH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gh_materialize, dim3(gslot), dim3(GH_BLOCK), 0, st.s, hold, m, dpk,
                         dpv, dtk, dtv, dused);
      NFH_LAUNCH_CHECK();
    }
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
      hipLaunchKernelGGL(k_gh_chunk_count, dim3((nch + GH_CPW - 1) / GH_CPW), dim3(GH_BLOCK), 0,
                         st.s, du, dtoff, dcb, uint32_t(p), nch, cc);
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
    hipLaunchKernelGGL(k_gh_chunk_count, dim3((nch + GH_CPW - 1) / GH_CPW), dim3(GH_BLOCK), 0,
                       st.s, du, dtoff, dcb, uint32_t(p), nch, cc);
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
