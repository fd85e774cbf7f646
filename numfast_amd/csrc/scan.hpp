// scan.hpp -- single-pass tiled exclusive scan with decoupled look-back.
//
// One kernel serves every "count, prefix, emit" pass of the library:
// compaction of present groups (carry), run heads (unique, sorted runs),
// group start offsets.  The functor `Op` supplies
//   __device__ uint32_t value(uint32_t i) const;             // item count
//   __device__ void emit(uint32_t i, uint32_t excl) const;    // per item
//   __device__ void total(uint32_t t) const;                  // once
// Items are owned in blocked order (thread t owns SC_IPT consecutive items),
// so emit() sees each item with the exclusive prefix of all earlier items.
#pragma once
#include "device.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

constexpr int SC_BLOCK = 256;
constexpr int SC_IPT = 16;
constexpr int SC_TILE = SC_BLOCK * SC_IPT;

template <class Op>
__global__ __launch_bounds__(SC_BLOCK) void k_scan_tiles(Op op, uint32_t n,
                                                         uint64_t* __restrict__ lookback,
                                                         uint32_t* __restrict__ tile_ctr,
                                                         uint32_t epoch) {
  __shared__ uint32_t s_tmp[SC_BLOCK / WAVE];
  __shared__ uint32_t s_tile, s_prefix;
  const int tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * uint32_t(SC_TILE) + uint32_t(tid) * SC_IPT;
  uint32_t v[SC_IPT];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < SC_IPT; ++j) {
    const uint32_t i = base + j;
    v[j] = i < n ? op.value(i) : 0u;
    c += v[j];
  }
  uint32_t tile_total;
  const uint32_t tex = block_excl_scan<SC_BLOCK / WAVE>(c, s_tmp, &tile_total);
  if (tid == 0) {
    uint64_t* my = lookback + tile;
    if (tile == 0) {
      lb_store(my, lb_pack(epoch, LB_PREFIX, tile_total));
      s_prefix = 0;
    } else {
      lb_store(my, lb_pack(epoch, LB_AGG, tile_total));
      const uint32_t ex = lb_lookback(lookback, tile, 1, epoch);
      lb_store(my, lb_pack(epoch, LB_PREFIX, ex + tile_total));
      s_prefix = ex;
    }
  }
  __syncthreads();
  uint32_t run = s_prefix + tex;
#pragma unroll
  for (int j = 0; j < SC_IPT; ++j) {
    const uint32_t i = base + j;
    if (i < n) op.emit(i, run);
    run += v[j];
  }
  const uint32_t ntiles = (n + SC_TILE - 1) / SC_TILE;
  if (tile == ntiles - 1 && tid == SC_BLOCK - 1) op.total(s_prefix + tile_total);
}

// Launch one scan over n items (n >= 1).  Asynchronous on s.
template <class Op>
void scan_tiles(const Op& op, size_t n, hipStream_t s) {
  if (n == 0) return;
  const uint32_t tiles = uint32_t((n + SC_TILE - 1) / SC_TILE);
  uint64_t* lb = lookback_words(tiles, s);
  const uint32_t epoch = lookback_epoch(s);
  uint32_t* ctr = scratch_t<uint32_t>(S_META2, 8) + 4;
  fill_u32(ctr, 0, 1, s);
  ProfScope prof("scan_tiles", s);
  hipLaunchKernelGGL(k_scan_tiles<Op>, dim3(tiles), dim3(SC_BLOCK), 0, s, op, uint32_t(n), lb,
                     ctr, epoch);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh
