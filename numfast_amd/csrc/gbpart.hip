// gbpart.hip -- partitioned dense group-by (sum / count / min / max) for
// large key spans, with no per-row global atomics.
//
// Device-scope atomics on MI355X execute beside memory at a fixed rate
// (about 25 G/s whatever the footprint), so a row-per-atomic scatter
// cannot approach HBM bandwidth.  Instead:
//   1. hot keys (sampled, see groupagg.hip) are aggregated per workgroup in
//      LDS during the bucket-histogram pass and merged once at the end;
//   2. every other valid row becomes a record (o = key - kmin, value) and
//      one or two onesweep passes (decoupled look-back, 7-bit digits of
//      o >> 11) order the records by 2048-key bucket (LSD, stable);
//   3. one workgroup per bucket counts its present keys (published at once
//      for the buckets after it), aggregates its records in LDS, takes its
//      hot keys' aggregates, and writes its groups in key order at its
//      offset among all buckets' groups (decoupled look-back).
// Histogram of buckets first, so every pass knows its global bin starts.
#include <algorithm>
#include <mutex>

#include "device.hpp"
#include "gbpart.hpp"
#include "groupagg.hpp"
#include "hot.hpp"
#include "onesweep.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

typedef unsigned long long u64;

constexpr int GP_BLOCK = 256;
// the histogram pass: 1024 threads, one workgroup per CU (about 58 KB of
// LDS with a 2^24-key span; two of 512 threads measured 1.5% slower)
constexpr int GP_HBLOCK = 1024;
// Partition passes use 8192-row tiles (512 threads, ~70 KB of LDS) so a
// digit's run of records fills whole 128-byte lines when written out; 512
// threads leave the kernel a 256-VGPR budget (1024 would cap it at 128 and
// spill).
constexpr int GP_PBLOCK = 512;
constexpr int GP_WAVES = GP_PBLOCK / WAVE;
constexpr int GP_IPT = 16;
constexpr int GP_TILE = GP_PBLOCK * GP_IPT;
constexpr size_t GP_PART_LDS = size_t(GP_TILE) * 8 + GP_WAVES * 128 * 4 + 128 * 4 * 2 +
                               GP_WAVES * 4 + 16 + size_t(NCHAIN) * 128 * 4 + 128;
// Regions of the histogram pass's record mode (one per workgroup of that
// pass, a multiple of NCHAIN)
constexpr int GP_REG_MAX = 256;
static_assert(GP_REG_MAX <= GP_PBLOCK, "one region per thread in the region scan");
static_assert(2 * (GP_PART_LDS + 4 * (GP_REG_MAX + 1)) <= 160 * 1024,
              "two partition workgroups per CU with the region table");
static_assert(HOT_MAX == 512 && HOT_TBL == 1024, "LDS budget of the passes below");

// ------------------------------------------------------------ histogram
struct HotAcc {
  u64 sum;
  uint32_t cnt;
  int32_t mn;
  int32_t mx;
  uint32_t pad;
};

// One persistent pass over the rows: hot rows aggregate in LDS (flushed
// once per workgroup), every other valid row is counted in its 2048-key
// bucket.
//
// Record mode (rec_o set and the sampled hot keys carrying at least
// reg_min of the HOT_SAMPLE sampled rows, so that few rows are records):
// workgroup g takes the contiguous rows [g * rsize, (g + 1) * rsize) and
// also writes each record (key - kmin, value) of them, compacted, from
// rec_o / rec_v + g * rsize on; reg_cnt[g] = its record count.  Region g
// belongs to look-back chain g / (gridDim.x / NCHAIN) of the first
// partition pass, which then reads the records instead of the rows.  The
// records of one wave's batch are stored in (quad, row, lane) order, so
// each store instruction writes one contiguous run.  mode bit 2 says the
// mode was taken.
__global__ __launch_bounds__(GP_HBLOCK) void k_gp_hist(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const uint8_t* __restrict__ vm, uint32_t n, int32_t kmin, uint32_t nb,
    const int2* __restrict__ table, uint32_t* __restrict__ hist, HotAcc* __restrict__ hot_acc,
    uint32_t* __restrict__ chist, uint32_t ptiles, bool vec, int bbits,
    uint32_t* __restrict__ mode, const uint32_t* __restrict__ hot_n, uint32_t reg_min,
    uint32_t* __restrict__ rec_o, int32_t* __restrict__ rec_v, uint32_t* __restrict__ reg_cnt,
    uint32_t rsize) {
  extern __shared__ uint32_t s_h[];  // nb counters
  __shared__ uint32_t s_c0[NCHAIN][128];  // digit-0 counts per chain of the first pass
  __shared__ uint32_t s_cur;              // records of this workgroup (record mode)
  __shared__ int32_t t_key[HOT_TBL], t_idx[HOT_TBL];
  // hot aggregates: values in [-2^15, 2^15) add (1 << 40) | (x + 2^15) to
  // h_pk (count in bits 40..63, biased sum below: exact while a workgroup
  // sees fewer than 2^24 rows), others go to h_sum / h_cnt
  __shared__ u64 h_pk[HOT_MAX];
  __shared__ u64 h_sum[HOT_MAX];
  __shared__ uint32_t h_cnt[HOT_MAX];
  __shared__ int32_t h_mn[HOT_MAX], h_mx[HOT_MAX];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) s_h[b] = 0;
  for (uint32_t b = threadIdx.x; b < NCHAIN * 128; b += blockDim.x) (&s_c0[0][0])[b] = 0;
  hot_load(table, t_key, t_idx);
  for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) {
    h_pk[b] = 0;
    h_sum[b] = 0;
    h_cnt[b] = 0;
    h_mn[b] = INT32_MAX;
    h_mx[b] = INT32_MIN;
  }
  if (threadIdx.x == 0) s_cur = 0;
  __syncthreads();
  bool oor = false;  // a record value outside [-2^15, 2^15) (mode bit 0)
  // chain of row i's tile in the first partition pass
  auto chain_of_row = [&](uint32_t i) { return chain_of_tile(ptiles, i / uint32_t(GP_TILE)); };
  const bool reg = rec_o != nullptr && table != nullptr && hot_n[1] >= reg_min;
  // true when the row is a record (valid and not hot); in record mode the
  // per-chain digit counts come from the bucket counts at the end (one
  // chain per workgroup)
  auto row = [&](int32_t k, int32_t x, bool ok, uint32_t ch) -> bool {
    if (!ok) return false;
    const int hi = table ? hot_index(t_key, t_idx, k) : -1;
    if (hi >= 0) {
      if (x >= -32768 && x < 32768) {
        atomicAdd(&h_pk[hi], (u64(1) << 40) | u64(uint32_t(x + 32768)));
      } else {
        atomicAdd(&h_sum[hi], (u64)(long long)x);
        atomicAdd(&h_cnt[hi], 1u);
      }
      // min / max only when x would change them (see k_gp_bucket_emit)
      if (x < h_mn[hi]) atomicMin(&h_mn[hi], x);
      if (x > h_mx[hi]) atomicMax(&h_mx[hi], x);
      return false;
    }
    const uint32_t b = uint32_t(k - kmin) >> bbits;
    oor |= x < -32768 || x >= 32768;
    atomicAdd(&s_h[b], 1u);
    if (!reg) atomicAdd(&s_c0[ch][b & 127u], 1u);
    return true;
  };
  if (reg) {
    // record mode: one region per workgroup (the host guarantees
    // gridDim.x * rsize >= n, rsize % 4 == 0 and aligned quads)
    const uint32_t ch = blockIdx.x / (gridDim.x / NCHAIN);
    const uint32_t r0 = blockIdx.x * rsize;
    const uint32_t r1 = r0 < n ? min(n, r0 + rsize) : r0;
    const uint32_t q0 = r0 / 4, q1 = r1 / 4;  // whole quads of the region
    const uint64_t lt = lanemask_lt();
    const int lane = lane_id();
    // Emits the records whose bit is set in `cold` (bit 4u + r: row r of
    // quad u), in (u, r, lane) order, at the wave's share of the region.
    auto emit = [&](uint32_t cold, const int4* ka, const int4* va, int nq) {
      uint64_t bal[4 * 4];
      uint32_t tot = 0;
#pragma unroll
      for (int sl = 0; sl < 16; ++sl) {
        bal[sl] = sl < 4 * nq ? __ballot((cold >> sl) & 1u) : 0ull;
        tot += uint32_t(__popcll(bal[sl]));
      }
      if (tot == 0) return;  // uniform over the wave
      uint32_t wb = 0;
      if (lane == 0) wb = atomicAdd(&s_cur, tot);
      wb = r0 + uint32_t(__shfl(int(wb), 0));
#pragma unroll
      for (int sl = 0; sl < 16; ++sl) {
        if (sl < 4 * nq && ((cold >> sl) & 1u)) {
          const int4 k4 = ka[sl >> 2], v4 = va[sl >> 2];
          const int r = sl & 3;
          const int32_t k = r == 0 ? k4.x : r == 1 ? k4.y : r == 2 ? k4.z : k4.w;
          const int32_t x = r == 0 ? v4.x : r == 1 ? v4.y : r == 2 ? v4.z : v4.w;
          const uint32_t at = wb + uint32_t(__popcll(bal[sl] & lt));
          rec_o[at] = uint32_t(k - kmin);
          rec_v[at] = x;
        }
        wb += uint32_t(__popcll(bal[sl]));
      }
    };
    constexpr int RQ = 4;  // quads per thread per batch, all loaded first
    for (uint32_t qb = q0; qb < q1; qb += RQ * GP_HBLOCK) {  // uniform trip count
      int4 ka[RQ], va[RQ];
      uint32_t ma[RQ];
      bool in[RQ];
#pragma unroll
      for (int u = 0; u < RQ; ++u) {
        const uint32_t qq = qb + uint32_t(u * GP_HBLOCK) + threadIdx.x;
        in[u] = qq < q1;
        const uint32_t qc = in[u] ? qq : q1 - 1u;
        ka[u] = ld_stream4(keys, qc);
        va[u] = ld_stream4(vals, qc);
        ma[u] = vm ? reinterpret_cast<const uint32_t*>(vm)[qc] : 0x01010101u;
      }
      asm volatile("" ::: "memory");  // keep every load above the adds
      uint32_t cold = 0;
      // two quads at a time: their eight hot lookups, then the current
      // min / max of the hot rows' keys, are read in one batch each
#pragma unroll
      for (int u2 = 0; u2 < RQ; u2 += 2) {
        int32_t k8[8], x8[8];
        int h8[8];
        bool ok8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int u = u2 + (e >> 2), r = e & 3;
          k8[e] = r == 0 ? ka[u].x : r == 1 ? ka[u].y : r == 2 ? ka[u].z : ka[u].w;
          x8[e] = r == 0 ? va[u].x : r == 1 ? va[u].y : r == 2 ? va[u].z : va[u].w;
          ok8[e] = in[u] && ((ma[u] >> (8 * r)) & 0xFFu) != 0u;
        }
        hot_index_batch<8>(t_key, t_idx, k8, h8);
        int32_t mn8[8], mx8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int h = h8[e] >= 0 ? h8[e] : 0;
          mn8[e] = h_mn[h];
          mx8[e] = h_mx[h];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          if (!ok8[e]) continue;
          const int32_t k = k8[e], x = x8[e];
          const int hi = h8[e];
          if (hi >= 0) {
            if (x >= -32768 && x < 32768) {
              atomicAdd(&h_pk[hi], (u64(1) << 40) | u64(uint32_t(x + 32768)));
            } else {
              atomicAdd(&h_sum[hi], (u64)(long long)x);
              atomicAdd(&h_cnt[hi], 1u);
            }
            // min / max only when x would change them (see k_gp_bucket_emit)
            if (x < mn8[e]) atomicMin(&h_mn[hi], x);
            if (x > mx8[e]) atomicMax(&h_mx[hi], x);
          } else {
            const uint32_t b = uint32_t(k - kmin) >> bbits;
            oor |= x < -32768 || x >= 32768;
            atomicAdd(&s_h[b], 1u);
            cold |= 1u << (4 * u2 + e);
          }
        }
      }
      emit(cold, ka, va, RQ);
    }
    // the last 0-3 rows of the region (only the last region ends off a quad)
    {
      const uint32_t i = q1 * 4u + threadIdx.x;
      int4 k1 = make_int4(0, 0, 0, 0), v1 = make_int4(0, 0, 0, 0);
      uint32_t cold = 0;
      if (i < r1) {
        k1.x = keys[i];
        v1.x = vals[i];
        cold = uint32_t(row(k1.x, v1.x, !vm || vm[i], ch));
      }
      emit(cold, &k1, &v1, 1);
    }
  }
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t tail = reg ? n : 0u;
  if (vec && !reg) {
    // quads: 128-bit key / value loads, a 32-bit load of four mask bytes,
    // HQ quads in flight per thread (all loaded before any row is added)
    constexpr int HQ = 4;
    const uint32_t nq = n / 4;
    uint32_t q = tid;
    for (; q + (HQ - 1) * stride < nq; q += HQ * stride) {
      int4 ka[HQ], va[HQ];
      uint32_t ma[HQ];
#pragma unroll
      for (int u = 0; u < HQ; ++u) {
        const uint32_t qq = q + uint32_t(u) * stride;
        ka[u] = ld_stream4(keys, qq);
        va[u] = ld_stream4(vals, qq);
        ma[u] = vm ? reinterpret_cast<const uint32_t*>(vm)[qq] : 0x01010101u;
      }
      asm volatile("" ::: "memory");  // keep every load above the adds
#pragma unroll
      for (int u = 0; u < HQ; ++u) {
        const uint32_t ca = chain_of_row(4 * (q + uint32_t(u) * stride));
        row(ka[u].x, va[u].x, ma[u] & 0xFFu, ca);
        row(ka[u].y, va[u].y, ma[u] & 0xFF00u, ca);
        row(ka[u].z, va[u].z, ma[u] & 0xFF0000u, ca);
        row(ka[u].w, va[u].w, ma[u] & 0xFF000000u, ca);
      }
    }
    for (; q < nq; q += stride) {
      const int4 ka = reinterpret_cast<const int4*>(keys)[q];
      const int4 va = reinterpret_cast<const int4*>(vals)[q];
      const uint32_t ma = vm ? reinterpret_cast<const uint32_t*>(vm)[q] : 0x01010101u;
      const uint32_t ca = chain_of_row(4 * q);
      row(ka.x, va.x, ma & 0xFFu, ca);
      row(ka.y, va.y, ma & 0xFF00u, ca);
      row(ka.z, va.z, ma & 0xFF0000u, ca);
      row(ka.w, va.w, ma & 0xFF000000u, ca);
    }
    tail = nq * 4;
  }
  for (uint32_t i = tail + tid; i < n; i += stride)
    row(keys[i], vals[i], !vm || vm[i], chain_of_row(i));
  if (__ballot(oor) != 0 && lane_id() == 0) atomicOr(mode, 1u);
  __syncthreads();
  if (reg && threadIdx.x == 0) {
    reg_cnt[blockIdx.x] = s_cur;
    if (blockIdx.x == 0) atomicOr(mode, 4u);
  }
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    if (s_h[b]) atomicAdd(&hist[b], s_h[b]);
  if (reg) {
    // the first partition pass's digit is the bucket id's low 7 bits
    const uint32_t ch = blockIdx.x / (gridDim.x / NCHAIN);
    for (uint32_t d = threadIdx.x; d < 128; d += blockDim.x) {
      uint32_t c = 0;
      for (uint32_t b = d; b < nb; b += 128) c += s_h[b];
      if (c) atomicAdd(&chist[ch * 128 + d], c);
    }
  } else {
    for (uint32_t b = threadIdx.x; b < NCHAIN * 128; b += blockDim.x) {
      const uint32_t c = (&s_c0[0][0])[b];
      if (c) atomicAdd(&chist[b], c);
    }
  }
  if (table) {
    for (int b = threadIdx.x; b < HOT_MAX; b += blockDim.x) {
      const u64 pk = h_pk[b];
      const uint32_t pc = uint32_t(pk >> 40);
      const uint32_t c = h_cnt[b] + pc;
      if (c == 0) continue;
      // biased sum minus 2^15 per packed row (wrapping u64 arithmetic)
      const u64 sum = h_sum[b] + (pk & ((u64(1) << 40) - 1)) - u64(pc) * 32768u;
      atomicAdd(&hot_acc[b].sum, sum);
      atomicAdd(&hot_acc[b].cnt, c);
      atomicMin(&hot_acc[b].mn, h_mn[b]);
      atomicMax(&hot_acc[b].mx, h_mx[b]);
    }
  }
}

// Bucket starts: exclusive scan of hist (bstart[nb] = number of records).
// mode |= 2 when a bucket holds 2^24 records or more (no packed counts).
__global__ __launch_bounds__(1024) void k_gp_bins(const uint32_t* __restrict__ hist, uint32_t nb,
                                                  uint32_t* __restrict__ bstart,
                                                  uint32_t* __restrict__ mode) {
  __shared__ uint32_t s_tmp[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += blockDim.x) {
    const uint32_t b = base + threadIdx.x;
    const uint32_t c = b < nb ? hist[b] : 0u;
    if (c >= (1u << 24)) atomicOr(mode, 2u);
    uint32_t tot;
    const uint32_t ex = block_excl_scan<16>(c, s_tmp, &tot);
    if (b < nb) bstart[b] = s_carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) bstart[nb] = s_carry;
}

// ------------------------------------------------------------ partition pass

// One onesweep pass over records ordered by digit (o >> (11 + 7p)) & 127.
// FIRST reads raw rows: NULL rows and hot rows (already aggregated by
// k_gp_hist) are dropped, the others become records.  The pass runs in
// NCHAIN look-back chains of tiles (one per XCD, as in onesweep.hpp):
// dstart = [NCHAIN][128] global starts from per-chain digit counts,
// tile_ctr = NCHAIN zeroed counters; persistent workgroups loop over their
// chain's tiles.  NEXT: count each record's next digit per chain of the
// next pass (next_tiles tiles of the same size) into next_hist[NCHAIN][128].
// PO (second pass only, every value in [-2^15, 2^15)): the pass writes one
// 32-bit record per row, the key's slot in its bucket (the low `sbits`
// bits of o) below the biased value (x + 2^15); the bucket itself is
// implied by the record's position after the last pass.
// REG (first pass over the histogram pass's records): the input is nreg
// regions of rsize slots, region g holding reg_cnt[g] records from slot
// g * rsize on; each region is cut into whole tiles and one partial tile,
// and chain c takes the tiles of regions [c, c + 1) * nreg / NCHAIN, in
// order.
template <bool FIRST, bool LR, bool NEXT, bool PO = false, bool REG = false>
__global__ __launch_bounds__(GP_PBLOCK) void k_gp_part(
    const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
    const uint8_t* __restrict__ vm, uint32_t n_rows, int32_t kmin,
    const uint32_t* __restrict__ oin, const int32_t* __restrict__ vin, uint32_t n_rec,
    uint32_t* __restrict__ oout, int32_t* __restrict__ vout, int shift,
    const uint32_t* __restrict__ dstart, uint64_t* __restrict__ lookback,
    uint32_t* __restrict__ tile_ctr, uint32_t epoch, const int2* __restrict__ table,
    uint32_t next_tiles, uint32_t* __restrict__ next_hist, int sbits,
    const uint32_t* __restrict__ reg_cnt, uint32_t nreg, uint32_t rsize) {
  static_assert(!(REG && FIRST), "record regions are an input of a record pass");
  extern __shared__ __align__(16) unsigned char s_raw[];
  uint32_t* s_o = reinterpret_cast<uint32_t*>(s_raw);
  int32_t* s_v = reinterpret_cast<int32_t*>(s_o + GP_TILE);
  uint32_t(*s_wh)[128] = reinterpret_cast<uint32_t(*)[128]>(s_v + GP_TILE);
  uint32_t* s_tstart = &s_wh[GP_WAVES][0];
  uint32_t* s_gbase = s_tstart + 128;
  uint32_t* s_tmp = s_gbase + 128;
  uint32_t* s_tile = s_tmp + GP_WAVES;
  uint32_t* s_nh = s_tile + 4;  // [NCHAIN][128]
  // per digit: the next pass's chain of this tile's whole run, or 0xFF
  // when the run crosses a chain boundary
  unsigned char* s_cc = reinterpret_cast<unsigned char*>(s_nh + NCHAIN * 128);
  __shared__ int32_t t_key[FIRST ? HOT_TBL : 1];
  // REG: first tile of each region (s_rf[nreg] = all tiles)
  __shared__ uint32_t s_rf[REG ? GP_REG_MAX + 1 : 1];

  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  const uint32_t n = FIRST ? n_rows : n_rec;
  const uint32_t chain = blockIdx.x % NCHAIN;
  uint32_t cfirst, ctiles;
  if constexpr (REG) {
    // the host guarantees nreg <= GP_REG_MAX <= GP_PBLOCK, nreg % NCHAIN == 0
    const uint32_t rt =
        uint32_t(tid) < nreg ? (reg_cnt[tid] + uint32_t(GP_TILE) - 1u) / uint32_t(GP_TILE) : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<GP_WAVES>(rt, s_tmp, &tot);
    if (uint32_t(tid) < nreg) s_rf[tid] = ex;
    if (tid == 0) s_rf[nreg] = tot;
    __syncthreads();
    const uint32_t rpc = nreg / NCHAIN;
    cfirst = s_rf[chain * rpc];
    ctiles = s_rf[(chain + 1) * rpc] - cfirst;
  } else {
    const uint32_t T = (n + GP_TILE - 1) / GP_TILE;
    cfirst = chain_first(T, chain);
    ctiles = chain_tiles(T, chain);
  }
  const uint32_t* bins = dstart + chain * 128;
  uint32_t nb[NCHAIN];
#pragma unroll
  for (int k = 0; k < NCHAIN; ++k) nb[k] = chain_first(next_tiles, k) * uint32_t(GP_TILE);
  if constexpr (NEXT)
    for (int i = tid; i < NCHAIN * 128; i += GP_PBLOCK) s_nh[i] = 0;
  if constexpr (FIRST) hot_load(table, t_key, nullptr);

  for (;;) {
    if (tid == 0) *s_tile = atomicAdd(&tile_ctr[chain], 1u);
    for (int i = tid; i < GP_WAVES * 128; i += GP_PBLOCK) s_wh[i >> 7][i & 127] = 0;
    __syncthreads();
    const uint32_t local = *s_tile;
    if (local >= ctiles) break;  // uniform over the workgroup
    const uint32_t tile = cfirst + local;
    const bool head = local == 0;
    uint32_t base = tile * uint32_t(GP_TILE);
    uint32_t lim = n;  // items of this tile are those below lim
    if constexpr (REG) {
      // the region holding this tile: the last of the chain's regions whose
      // first tile is at or before it (regions without tiles share their
      // first tile with the next one)
      const uint32_t rpc = nreg / NCHAIN;
      uint32_t lo = chain * rpc, hi = lo + rpc;
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (s_rf[mid] <= tile) lo = mid; else hi = mid;
      }
      base = lo * rsize + (tile - s_rf[lo]) * uint32_t(GP_TILE);
      lim = lo * rsize + reg_cnt[lo];
    }
    const uint32_t wbase = base + uint32_t(wave) * (WAVE * GP_IPT);

    uint32_t ov[GP_IPT];
    int32_t vv[GP_IPT];
    uint32_t live = 0;  // bit j: item j is a record of this pass
    if constexpr (FIRST) {
      // Every lane loads from an in-range row (past the end: the last
      // row), so all loads are in flight at once; values are loaded for
      // every row (a 64-byte segment with no record at all is rare, so
      // this fetches the same lines as loading records only).
      uint32_t mb[GP_IPT];
#pragma unroll
      for (int j = 0; j < GP_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        const uint32_t pc = pos < n ? pos : n - 1u;
        ov[j] = uint32_t(keys[pc]);
        vv[j] = vals[pc];
        mb[j] = 1u;
      }
      if (vm != nullptr) {
#pragma unroll
        for (int j = 0; j < GP_IPT; ++j) {
          const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
          mb[j] = vm[pos < n ? pos : n - 1u];
        }
      }
#pragma unroll
      for (int j = 0; j < GP_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        if (pos < n && mb[j] != 0u) {
          const int32_t k = int32_t(ov[j]);
          if (!table || hot_slot_of(t_key, k) < 0) {
            ov[j] = uint32_t(k - kmin);
            live |= 1u << j;
          }
        }
        if (!((live >> j) & 1u)) vv[j] = 0;
      }
    }
    if constexpr (!FIRST) {
#pragma unroll
      for (int j = 0; j < GP_IPT; ++j) {
        const uint32_t pos = wbase + uint32_t(j * WAVE + lane);
        const uint32_t pc = pos < lim ? pos : lim - 1u;
        // streaming loads here and in the histogram pass (read once):
        // the step 4.06 -> 4.03 ms; in the radix passes they measured
        // slower (0.83 -> 0.90 ms)
        ov[j] = __builtin_nontemporal_load(oin + pc);
        vv[j] = __builtin_nontemporal_load(vin + pc);
        live |= uint32_t(pos < lim) << j;
      }
    }

    // Stable in-wave ranks (position order: item j of lane l, then lane
    // l+1, then item j+1).  LR: one lane-ordered LDS atomic per record,
    // else the ballot digit match.  Dropped rows are not a tail of the
    // wave, so the atomic form is used directly.
    uint32_t rk[GP_IPT];
#pragma unroll
    for (int j = 0; j < GP_IPT; ++j) {
      const bool valid = (live >> j) & 1u;
      const uint32_t d = (ov[j] >> shift) & 127u;
      if constexpr (LR) {
        rk[j] = valid ? atomicAdd(&s_wh[wave][d], 1u) : 0u;
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

    uint32_t cnt = 0;
    if (tid < 128) {
#pragma unroll
      for (int w = 0; w < GP_WAVES; ++w) {
        const uint32_t c = s_wh[w][tid];
        s_wh[w][tid] = cnt;
        cnt += c;
      }
      lb_store(lookback + size_t(tile) * 128 + tid,
               lb_pack(epoch, head ? LB_PREFIX : LB_AGG, cnt));
    }
    uint32_t tile_total;
    const uint32_t tstart = block_excl_scan<GP_WAVES>(tid < 128 ? cnt : 0u, s_tmp, &tile_total);
    if (tid < 128) s_tstart[tid] = tstart;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < GP_IPT; ++j) {
      if ((live >> j) & 1u) {
        const uint32_t d = (ov[j] >> shift) & 127u;
        const uint32_t p = s_tstart[d] + s_wh[wave][d] + rk[j];
        s_o[p] = ov[j];
        s_v[p] = vv[j];
      }
    }
    if (tid < 128) {
      uint32_t excl = 0;
      if (!head) {
        excl = lb_lookback(lookback + tid, tile, 128, epoch);
        lb_store(lookback + size_t(tile) * 128 + tid, lb_pack(epoch, LB_PREFIX, excl + cnt));
      }
      s_gbase[tid] = bins[tid] + excl - tstart;
      if constexpr (NEXT) {
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
    // GWU records per thread per step, their LDS reads issued together
    // (0.89 -> 0.88 ms for the first pass at 2^28 rows; four measured no
    // better)
    constexpr int GWU = 2;
    for (uint32_t i0 = 0; i0 < tile_total; i0 += GWU * GP_PBLOCK) {  // uniform trip count
      uint32_t ob[GWU], xv[GWU], gb[GWU], cc[GWU];
      bool ok[GWU];
#pragma unroll
      for (int u = 0; u < GWU; ++u) {
        const uint32_t i = i0 + uint32_t(u * GP_PBLOCK + tid);
        ok[u] = i < tile_total;
        const uint32_t ic = ok[u] ? i : 0u;
        ob[u] = s_o[ic];
        xv[u] = uint32_t(s_v[ic]);
      }
#pragma unroll
      for (int u = 0; u < GWU; ++u) {
        if (!ok[u]) ob[u] = 0u;
        const uint32_t d = (ob[u] >> shift) & 127u;
        gb[u] = s_gbase[d];
        if constexpr (NEXT) cc[u] = s_cc[d];
      }
#pragma unroll
      for (int u = 0; u < GWU; ++u) {
        const uint32_t i = i0 + uint32_t(u * GP_PBLOCK + tid);
        const uint32_t g = gb[u] + i;
        if (ok[u]) {
          if constexpr (PO) {
            oout[g] = (ob[u] & ((1u << sbits) - 1u)) | ((xv[u] + 32768u) << sbits);
          } else {
            oout[g] = ob[u];
            vout[g] = int32_t(xv[u]);
          }
        }
        if constexpr (NEXT) {
          uint32_t c = cc[u];
          if (c == 0xFFu) {
            c = 0;
#pragma unroll
            for (int k = 1; k < NCHAIN; ++k) c += g >= nb[k] ? 1u : 0u;
          }
          hist_add(s_nh, c * 128u + ((ob[u] >> (shift + 7)) & 127u), ok[u]);
        }
      }
    }
    __syncthreads();  // the tile buffers are reused by the next tile
  }
  if constexpr (NEXT) {
    for (int i = tid; i < NCHAIN * 128; i += GP_PBLOCK)
      if (s_nh[i]) atomicAdd(&next_hist[i], s_nh[i]);
  }
}

// ------------------------------------------------------------ bucket aggregation
// One workgroup per 2^BITS-key bucket, taken in order from a ticket:
//   1. a presence bitmap of the bucket's record keys and of its hot keys
//      with rows (hot keys have no records: their rows were aggregated by
//      the histogram pass) gives the bucket's group count, published at
//      once, so a later bucket's look-back seldom waits for an earlier
//      bucket's aggregation (buckets take unequal times);
//   2. LDS aggregation of the records;
//   3. the present groups are written in key order at the bucket's offset
//      (decoupled look-back over buckets), each wave W / waves consecutive
//      slots, 64 per ballot.
// No slot array: writing span slots and reading them back for a separate
// compaction (2 x 512 MB at a 2^24-key span) took 0.61 ms against 0.47.
// PACKED (every record value in [-2^15, 2^15) and every bucket under 2^24
// records, known from the histogram pass): one 64-bit LDS add per record
// carries the count (bits 40..63) and the biased sum (x + 2^15, below
// 2^40), so a record costs three LDS atomics instead of four and the
// bucket 16 B of LDS per key instead of 20.
template <int BITS, bool PACKED>
constexpr size_t gp_emit_lds() {
  return size_t(1 << BITS) * ((PACKED ? 16 : 20) + 2);
}

// PR: the records are the second pass's packed 32-bit form (slot below
// the biased value; v unused).
template <int BITS, int BLOCK, bool PACKED, bool PR = false>
__global__ __launch_bounds__(BLOCK) void k_gp_bucket_emit(
    const uint32_t* __restrict__ o, const int32_t* __restrict__ v,
    const uint32_t* __restrict__ bstart, uint32_t nb, uint64_t span, int32_t kmin,
    const HotAcc* __restrict__ hot_acc, const int32_t* __restrict__ hot_keys,
    const uint32_t* __restrict__ hot_n, GroupAggOut out, uint64_t* __restrict__ lookback,
    uint32_t* __restrict__ ticket, uint32_t epoch, int64_t* __restrict__ ng_out) {
  constexpr int W = 1 << BITS;
  constexpr int NWV = BLOCK / WAVE;
  constexpr int PER = W / NWV;     // slots per wave
  constexpr int CHK = PER / WAVE;  // ballots per wave
  static_assert(PER % WAVE == 0, "whole 64-slot chunks per wave");
  static_assert(!PR || PACKED, "packed records carry values in [-2^15, 2^15)");
  extern __shared__ __align__(16) unsigned char s_raw[];
  u64* a_sum = reinterpret_cast<u64*>(s_raw);  // PACKED: count << 40 | biased sum
  uint32_t* a_cnt = reinterpret_cast<uint32_t*>(a_sum + W);  // unused when PACKED
  int32_t* a_mn = PACKED ? reinterpret_cast<int32_t*>(a_sum + W)
                         : reinterpret_cast<int32_t*>(a_cnt + W);
  int32_t* a_mx = a_mn + W;
  uint16_t* s_hot = reinterpret_cast<uint16_t*>(a_mx + W);  // hot index + 1, or 0
  __shared__ uint32_t s_b, s_prefix, s_wtot[NWV], s_cnt;
  __shared__ uint32_t s_pres[W / 32];  // presence bitmap of the bucket's keys
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  if (tid == 0) {
    s_b = atomicAdd(ticket, 1u);
    s_cnt = 0;
  }
  for (int i = tid; i < W; i += BLOCK) {
    a_sum[i] = 0;
    if constexpr (!PACKED) a_cnt[i] = 0;
    a_mn[i] = INT32_MAX;
    a_mx[i] = INT32_MIN;
    s_hot[i] = 0;
  }
  for (int i = tid; i < W / 32; i += BLOCK) s_pres[i] = 0;
  __syncthreads();
  const uint32_t b = s_b;
  const uint32_t lo = bstart[b], hi = bstart[b + 1];
  const uint64_t first = uint64_t(b) << BITS;
  // 1. the bucket's group count first, from a presence bitmap of its record
  // keys and its hot keys with rows, published at once: buckets take
  // different times to aggregate, and a later bucket's look-back then
  // finds every earlier count without waiting for those aggregations
  const uint32_t nh = hot_n ? *hot_n : 0u;
  for (uint32_t h = tid; h < nh; h += BLOCK) {
    // (offsets of keys below kmin wrap, so they never match)
    const uint32_t off = uint32_t(hot_keys[h]) - uint32_t(kmin);
    if ((off >> BITS) == b && hot_acc[h].cnt != 0u) {
      s_hot[off & (W - 1)] = uint16_t(h + 1u);
      atomicOr(&s_pres[(off & (W - 1)) >> 5], 1u << (off & 31u));
    }
  }
  for (uint32_t i0 = lo; i0 < hi; i0 += 8 * BLOCK) {
    uint32_t os[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t i = i0 + uint32_t(r * BLOCK) + tid;
      os[r] = o[i < hi ? i : hi - 1u];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (i0 + uint32_t(r * BLOCK) + tid >= hi) continue;
      const uint32_t sl = os[r] & (W - 1);
      const uint32_t bit = 1u << (sl & 31u);
      if (!(s_pres[sl >> 5] & bit)) atomicOr(&s_pres[sl >> 5], bit);
    }
  }
  __syncthreads();
  {
    uint32_t c = 0;
    for (int i = tid; i < W / 32; i += BLOCK) {
      uint32_t wd = s_pres[i];
      const uint64_t g0 = first + uint64_t(i) * 32u;  // slots past the span hold no key
      if (g0 + 32u > span) wd &= g0 >= span ? 0u : ((1u << uint32_t(span - g0)) - 1u);
      c += uint32_t(__popc(wd));
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0 && c) atomicAdd(&s_cnt, c);
  }
  __syncthreads();
  const uint32_t ttot = s_cnt;
  if (tid == 0) lb_store(lookback + b, lb_pack(epoch, b == 0 ? LB_PREFIX : LB_AGG, ttot));
  // 2. aggregation of the records
  constexpr int RB = 8;
  for (uint32_t i0 = lo; i0 < hi; i0 += RB * BLOCK) {
    uint32_t os[RB];
    int32_t xs[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const uint32_t i = i0 + uint32_t(r * BLOCK) + tid;
      const uint32_t ic = i < hi ? i : hi - 1u;
      os[r] = o[ic];
      if constexpr (!PR) xs[r] = v[ic];
    }
    asm volatile("" ::: "memory");  // keep every load above the adds
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (i0 + uint32_t(r * BLOCK) + tid >= hi) continue;
      const uint32_t sl = os[r] & (W - 1);
      const int32_t x = PR ? int32_t(os[r] >> BITS) - 32768 : xs[r];
      if constexpr (PACKED) {
        atomicAdd(&a_sum[sl], (u64(1) << 40) | u64(uint32_t(x + 32768)));
      } else {
        atomicAdd(&a_sum[sl], (u64)(long long)x);
        atomicAdd(&a_cnt[sl], 1u);
      }
      if (x < a_mn[sl]) atomicMin(&a_mn[sl], x);
      if (x > a_mx[sl]) atomicMax(&a_mx[sl], x);
    }
  }
  // 3. emit: 64 slots per ballot, each wave its own run of slots
  const int wslot = wave * PER;
  uint64_t bal[CHK];
  uint32_t wtot = 0;
#pragma unroll
  for (int j = 0; j < CHK; ++j) {
    const int i = wslot + j * WAVE + lane;
    const bool present =
        (first + uint64_t(i) < span) && ((s_pres[i >> 5] >> (uint32_t(i) & 31u)) & 1u);
    bal[j] = __ballot(present);
    wtot += uint32_t(__popcll(bal[j]));
  }
  if (lane == 0) s_wtot[wave] = wtot;
  if (tid == 0) {
    if (b == 0) {
      s_prefix = 0;
    } else {
      const uint32_t ex = lb_lookback(lookback, b, 1, epoch);
      lb_store(lookback + b, lb_pack(epoch, LB_PREFIX, ex + ttot));
      s_prefix = ex;
    }
  }
  __syncthreads();
  uint32_t woff = 0;
#pragma unroll
  for (int w = 0; w < NWV; ++w)
    if (w < wave) woff += s_wtot[w];
  uint32_t pos = s_prefix + woff;
#pragma unroll
  for (int j = 0; j < CHK; ++j) {
    const int i = wslot + j * WAVE + lane;
    if ((bal[j] >> lane) & 1ull) {
      const uint32_t ex = pos + uint32_t(__popcll(bal[j] & lanemask_lt()));
      u64 sum;
      uint32_t c;
      int32_t mn, mx;
      const uint32_t hh = s_hot[i];
      if (hh != 0) {
        const HotAcc a = hot_acc[hh - 1u];
        sum = a.sum;
        c = a.cnt;
        mn = a.mn;
        mx = a.mx;
      } else {
        if constexpr (PACKED) {
          const u64 pk = a_sum[i];
          c = uint32_t(pk >> 40);
          sum = (pk & ((u64(1) << 40) - 1)) - u64(c) * 32768u;  // wrapping u64
        } else {
          sum = a_sum[i];
          c = a_cnt[i];
        }
        mn = a_mn[i];
        mx = a_mx[i];
      }
      const int64_t ssum = (int64_t)sum;
      if (out.ukeys) out.ukeys[ex] = int64_t(kmin) + int64_t(first + uint64_t(i));
      if (out.counts) out.counts[ex] = int64_t(c);
      if (out.sums) out.sums[ex] = ssum;
      if (out.mins) out.mins[ex] = mn;
      if (out.maxs) out.maxs[ex] = mx;
      if (out.means) out.means[ex] = double(ssum) / double(c);
    }
    pos += uint32_t(__popcll(bal[j]));
  }
  if (b == nb - 1 && tid == 0) *ng_out = int64_t(s_prefix + ttot);
}

__global__ void k_gp_hotinit(HotAcc* __restrict__ h, uint32_t m) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    HotAcc z;
    z.sum = 0;
    z.cnt = 0;
    z.mn = INT32_MAX;
    z.mx = INT32_MIN;
    z.pad = 0;
    h[i] = z;
  }
}

bool gp_supported(uint64_t span, int bucket_bits) {
  return (bucket_bits == GP_BUCKET_BITS || bucket_bits == GP_BUCKET_BITS_WIDE) &&
         span <= (uint64_t(1) << (bucket_bits + 13));
}

int64_t gp_aggregate(const int32_t* keys, const int32_t* vals, const uint8_t* vm, size_t n_,
                     int32_t kmin, uint64_t span, const int2* table, const int32_t* hot_keys,
                     const uint32_t* hot_n, const GroupAggOut& out, hipStream_t s,
                     int bucket_bits) {
  const uint32_t n = uint32_t(n_);
  const int bb = bucket_bits == GP_BUCKET_BITS_WIDE ? GP_BUCKET_BITS_WIDE : GP_BUCKET_BITS;
  const uint32_t nb = uint32_t((span + (1u << bb) - 1) >> bb);
  const int passes = nb <= 1 ? 0 : (nb <= 128 ? 1 : 2);
  constexpr size_t CH = size_t(NCHAIN) * 128;
  // scratch: hist[nb] | bstart[nb+1] | mode | per-chain digit counts
  // [2][CH] | per-chain digit starts [2][CH] | tile counters [2][NCHAIN] |
  // hot_acc
  const size_t meta_words = size_t(nb) * 2 + 2 + 4 * CH + 2 * NCHAIN;
  uint32_t* meta = scratch_t<uint32_t>(S_META, meta_words + 4 + HOT_MAX * 6 + 8 + GP_REG_MAX);
  uint32_t* hist = meta;
  uint32_t* bstart = hist + nb;
  uint32_t* mode = bstart + nb + 1;
  uint32_t* chist = mode + 1;
  uint32_t* dstart = chist + 2 * CH;
  uint32_t* ctr = dstart + 2 * CH;
  HotAcc* hot_acc = reinterpret_cast<HotAcc*>(meta + ((meta_words + 3) & ~size_t(3)));
  // bucket ticket, then the group count (8-byte aligned), then the record
  // counts of the histogram pass's regions
  uint32_t* ticket = reinterpret_cast<uint32_t*>(hot_acc + HOT_MAX);
  int64_t* ng_d = reinterpret_cast<int64_t*>(ticket + 2);
  uint32_t* reg_cnt = reinterpret_cast<uint32_t*>(ng_d + 1);
  fill_u32(meta, 0, meta_words, s);
  hipLaunchKernelGGL(k_gp_hotinit, dim3(2), dim3(256), 0, s, hot_acc, uint32_t(HOT_MAX));
  NFH_LAUNCH_CHECK();
  // record buffers (ping-pong): at most n records
  uint32_t* o0 = scratch_t<uint32_t>(S_T0, n);
  int32_t* v0 = scratch_t<int32_t>(S_T1, n);
  uint32_t* o1 = scratch_t<uint32_t>(S_T2, n);
  int32_t* v1 = scratch_t<int32_t>(S_T3, n);
  const double mask_b = vm ? 1.0 : 0.0;
  const uint32_t tiles = (n + GP_TILE - 1) / GP_TILE;
  const bool vec = (reinterpret_cast<uintptr_t>(keys) % 16) == 0 &&
                   (reinterpret_cast<uintptr_t>(vals) % 16) == 0 &&
                   (vm == nullptr || (reinterpret_cast<uintptr_t>(vm) % 4) == 0);
  // Record mode of the histogram pass: one region of rows per workgroup
  // (one workgroup per CU, a multiple of NCHAIN), offered when the sampled
  // hot keys exist; the pass takes it when they carry at least 40% of the
  // sampled rows (the first partition pass then reads 8 B per record
  // instead of 9 B per row and skips the hot-key lookups: at 2^28 Zipf(1.1)
  // rows, first partition pass + histogram pass 0.87 + 0.57 ms became
  // 0.40 + 0.70 ms).
  unsigned hgrid = grid_for(n / 8 + 1, GP_HBLOCK, num_cus());
  const uint32_t nreg = uint32_t(std::min(GP_REG_MAX, num_cus() / NCHAIN * NCHAIN));
  const bool offer_reg = table != nullptr && vec && n >= (1u << 20) && n <= (1u << 31) &&
                         nreg >= uint32_t(NCHAIN);
  uint32_t rsize = 0;
  if (offer_reg) {
    hgrid = nreg;
    rsize = ((n + nreg - 1) / nreg + 63u) & ~63u;  // nreg * rsize >= n, whole quads
  }
  constexpr uint32_t REG_MIN = uint32_t(HOT_SAMPLE) * 2u / 5u;
  {
    // algorithmic bytes: key + value + validity per row (+ 8 B per record
    // in record mode, added below once the count is known)
    ProfScope prof("gp_hist", s, double(n) * (8.0 + mask_b));
    const size_t lds = size_t(nb) * 4;
    hipLaunchKernelGGL(k_gp_hist, dim3(hgrid), dim3(GP_HBLOCK), lds, s, keys, vals, vm, n, kmin,
                       nb, table, hist, hot_acc, chist, tiles, vec, bb, mode, hot_n, REG_MIN,
                       offer_reg ? o1 : nullptr, v1, reg_cnt, rsize);
    NFH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_gp_bins, dim3(1), dim3(1024), 0, s, hist, nb, bstart, mode);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_chain_bins<128>, dim3(1), dim3(128), 0, s, chist, dstart);
  NFH_LAUNCH_CHECK();
  // records = valid non-hot rows (bstart[nb]), then the mode word
  uint32_t* h_m = static_cast<uint32_t*>(pinned(8));
  NFH_CHECK(hipMemcpyAsync(h_m, bstart + nb, 8, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  const uint32_t m = h_m[0];
  const bool packed = (h_m[1] & 3u) == 0u;
  const bool regm = (h_m[1] & 4u) != 0u;
  if (regm) prof_add_bytes("gp_hist", double(m) * 8.0);
  // at most one partial tile per region beyond the whole tiles
  uint64_t* lb = lookback_words((size_t(tiles) + nreg) * 128, s);
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    const void* fns[14] = {reinterpret_cast<const void*>(&k_gp_part<true, true, true>),
                           reinterpret_cast<const void*>(&k_gp_part<true, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, true, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<true, true, false>),
                           reinterpret_cast<const void*>(&k_gp_part<true, false, false>),
                           reinterpret_cast<const void*>(&k_gp_part<false, true, false>),
                           reinterpret_cast<const void*>(&k_gp_part<false, false, false>),
                           reinterpret_cast<const void*>(&k_gp_part<false, true, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, false, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, true, true, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, false, true, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, true, false, false, true>),
                           reinterpret_cast<const void*>(&k_gp_part<false, false, false, false, true>)};
    for (const void* f : fns)
      NFH_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(GP_PART_LDS)));
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(
            &k_gp_bucket_emit<GP_BUCKET_BITS_WIDE, 2 * GP_BLOCK, false>),
        hipFuncAttributeMaxDynamicSharedMemorySize,
        int(gp_emit_lds<GP_BUCKET_BITS_WIDE, false>())));
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_gp_bucket_emit<GP_BUCKET_BITS_WIDE, 2 * GP_BLOCK, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize,
        int(gp_emit_lds<GP_BUCKET_BITS_WIDE, true>())));
    NFH_CHECK(hipFuncSetAttribute(
        reinterpret_cast<const void*>(
            &k_gp_bucket_emit<GP_BUCKET_BITS_WIDE, 2 * GP_BLOCK, true, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize,
        int(gp_emit_lds<GP_BUCKET_BITS_WIDE, true>())));
  });
  const bool lr = lds_rank_supported();
  const bool two = passes == 2 && m > 0;
  const uint32_t t2 = (m + GP_TILE - 1) / GP_TILE;
  // persistent: two workgroups per CU (about 75 KB of LDS each), a
  // multiple of NCHAIN
  const unsigned grid = unsigned(std::max<int>(NCHAIN, 2 * num_cus() / NCHAIN * NCHAIN));
  // pass 1 always runs: over the rows it drops NULL and hot rows (digit 0
  // when nb <= 1); in record mode it reads the histogram pass's records
  if (regm) {
    // reads and writes (o, value) per record
    ProfScope prof("gp_part_first", s, double(m) * 16.0);
    const uint32_t epoch = lookback_epoch(s);
    const auto fn = two ? (lr ? &k_gp_part<false, true, true, false, true>
                              : &k_gp_part<false, false, true, false, true>)
                        : (lr ? &k_gp_part<false, true, false, false, true>
                              : &k_gp_part<false, false, false, false, true>);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(GP_PBLOCK), GP_PART_LDS, s, nullptr, nullptr,
                       nullptr, 0u, kmin, o1, v1, m, o0, v0, bb, dstart, lb, ctr, epoch,
                       nullptr, t2, chist + CH, bb, reg_cnt, nreg, rsize);
    NFH_LAUNCH_CHECK();
  } else {
    // reads key + validity of every row and the value of each record,
    // writes (o, value) per record
    ProfScope prof("gp_part_first", s, double(n) * (4.0 + mask_b) + double(m) * 12.0);
    const uint32_t epoch = lookback_epoch(s);
    const auto fn = two ? (lr ? &k_gp_part<true, true, true> : &k_gp_part<true, false, true>)
                        : (lr ? &k_gp_part<true, true, false> : &k_gp_part<true, false, false>);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(GP_PBLOCK), GP_PART_LDS, s, keys, vals, vm, n, kmin,
                       nullptr, nullptr, 0u, o0, v0, bb, dstart, lb, ctr, epoch, table,
                       t2, chist + CH, bb, nullptr, 0u, 0u);
    NFH_LAUNCH_CHECK();
  }
  const uint32_t* fo = o0;
  const int32_t* fv = v0;
  // packed records out of the second pass when every value fits 16 bits
  // (the bucket pass then reads 4 B per record instead of 8, twice)
  const bool pr = two && packed;
  if (passes == 2) {
    if (m > 0) {
      hipLaunchKernelGGL(k_chain_bins<128>, dim3(1), dim3(128), 0, s, chist + CH, dstart + CH);
      NFH_LAUNCH_CHECK();
      ProfScope prof("gp_part", s, double(m) * (pr ? 12.0 : 16.0));
      const uint32_t epoch = lookback_epoch(s);
      const auto fn = pr ? (lr ? &k_gp_part<false, true, false, true>
                               : &k_gp_part<false, false, false, true>)
                         : (lr ? &k_gp_part<false, true, false> : &k_gp_part<false, false, false>);
      hipLaunchKernelGGL(fn, dim3(grid), dim3(GP_PBLOCK), GP_PART_LDS, s, nullptr, nullptr,
                         nullptr, 0u, kmin, o0, v0, m, o1, v1, bb + 7, dstart + CH,
                         lb, ctr + NCHAIN, epoch, nullptr, 0u, nullptr, bb, nullptr, 0u, 0u);
      NFH_LAUNCH_CHECK();
    }
    fo = o1;
    fv = v1;
  }
  fill_u32(ticket, 0, 1, s);
  uint64_t* lbb = lookback_words(nb, s);
  const uint32_t epoch = lookback_epoch(s);
  const uint32_t* hn = table ? hot_n : nullptr;
  {
    // reads the records (twice: the presence count, then the aggregation)
    // and writes 40 B per group; the byte model counts the first read
    ProfScope prof("gp_bucket_emit", s, double(m) * (pr ? 4.0 : 8.0));
    // 512 threads per bucket for both bucket widths (2048-key buckets:
    // 0.57 ms with 256, 0.47 with 512, 0.56 with 1024 at 2^28 Zipf rows)
#define NFH_GP_EMIT(BITS, PK, PRV)                                                           \
  hipLaunchKernelGGL((k_gp_bucket_emit<BITS, 2 * GP_BLOCK, PK, PRV>), dim3(nb),                \
                     dim3(2 * GP_BLOCK), (gp_emit_lds<BITS, PK>()), s, fo, fv, bstart, nb, span, \
                     kmin, hot_acc, hot_keys, hn, out, lbb, ticket, epoch, ng_d)
    if (bb == GP_BUCKET_BITS_WIDE) {
      if (pr)
        NFH_GP_EMIT(GP_BUCKET_BITS_WIDE, true, true);
      else if (packed)
        NFH_GP_EMIT(GP_BUCKET_BITS_WIDE, true, false);
      else
        NFH_GP_EMIT(GP_BUCKET_BITS_WIDE, false, false);
    } else {
      if (pr)
        NFH_GP_EMIT(GP_BUCKET_BITS, true, true);
      else if (packed)
        NFH_GP_EMIT(GP_BUCKET_BITS, true, false);
      else
        NFH_GP_EMIT(GP_BUCKET_BITS, false, false);
    }
#undef NFH_GP_EMIT
    NFH_LAUNCH_CHECK();
  }
  return read_scalar(ng_d, s);
}

}  // namespace nfh
