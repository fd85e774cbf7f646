// unique.hip -- sorted-order unique + inverse (nf_unique_inverse_i32/_i64).
//
// Semantics: numfast-native/src/unique.rs:3-23, :30-91 -- identical
// labelling to np.unique(keys, return_inverse=True): uniq ascending,
// inv[row] = sorted position of keys[row], uniq[inv] == keys.
//
// Design: the stable radix argsort (sort.hip) runs ascending with its last
// pass also writing the keys in sorted order.  One run-encode pass then
// reads (sorted key, perm) tile by tile, flags run heads, scans the flags
// with a decoupled look-back across tiles, writes the distinct values
// compacted and scatters the code of each sorted position to inv[perm[r]].
#include "device.hpp"
#include "runtime.hpp"
#include "sort.hpp"
#include "unique.hpp"

namespace nfh {

constexpr int RE_BLOCK = 256;
constexpr int RE_WAVES = RE_BLOCK / WAVE;
constexpr int RE_IPT = 16;
constexpr int RE_TILE = RE_BLOCK * RE_IPT;

template <class K>
__global__ __launch_bounds__(RE_BLOCK) void k_run_encode(
    const K* __restrict__ sk, const int32_t* __restrict__ perm, uint32_t n,
    K* __restrict__ uniq, int32_t* __restrict__ inv, uint64_t* __restrict__ lookback,
    uint32_t* __restrict__ tile_ctr, uint32_t epoch, int64_t* __restrict__ ng_out) {
  __shared__ K s_k[RE_TILE + 1];
  __shared__ int32_t s_p[RE_TILE];
  __shared__ uint32_t s_tmp[RE_WAVES];
  __shared__ uint32_t s_tile, s_prefix;
  const int tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * uint32_t(RE_TILE);
  const uint32_t cnt = (n - base) < uint32_t(RE_TILE) ? (n - base) : uint32_t(RE_TILE);
  // Coalesced (striped) loads into LDS; s_k[0] holds the key before the tile.
  for (uint32_t i = tid; i < cnt; i += RE_BLOCK) {
    s_k[i + 1] = sk[base + i];
    s_p[i] = perm[base + i];
  }
  if (tid == 0) s_k[0] = base > 0 ? sk[base - 1] : K(0);
  __syncthreads();
  // Blocked items: thread t owns tile positions t*IPT .. t*IPT+IPT-1.
  uint32_t flags = 0, c = 0;
#pragma unroll
  for (int j = 0; j < RE_IPT; ++j) {
    const uint32_t i = uint32_t(tid * RE_IPT + j);
    if (i < cnt) {
      const bool head = (base + i == 0) || (s_k[i + 1] != s_k[i]);
      flags |= uint32_t(head) << j;
      c += uint32_t(head);
    }
  }
  uint32_t tile_total;
  const uint32_t tex = block_excl_scan<RE_WAVES>(c, s_tmp, &tile_total);
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
  const uint32_t pre = s_prefix;
  uint32_t run = pre + tex;  // heads strictly before this thread's items
#pragma unroll
  for (int j = 0; j < RE_IPT; ++j) {
    const uint32_t i = uint32_t(tid * RE_IPT + j);
    if (i < cnt) {
      if ((flags >> j) & 1u) {
        uniq[run] = s_k[i + 1];
        ++run;
      }
      inv[s_p[i]] = int32_t(run - 1);
    }
  }
  if (base + cnt == n && tid == RE_BLOCK - 1) *ng_out = int64_t(pre + tile_total);
}

template <class K>
int64_t unique_inverse_device(const K* keys, size_t n_, K* uniq, int32_t* inv, int32_t* perm,
                              word_t<K>* t0, word_t<K>* t1, int32_t* tp, hipStream_t s) {
  const uint32_t n = uint32_t(n_);
  if (n == 0) return 0;
  K* sorted = reinterpret_cast<K*>(t0);  // free after the last sort pass
  radix_argsort<K>(keys, n, false, perm, t0, t1, tp, sorted, s);
  const uint32_t tiles = (n + RE_TILE - 1) / RE_TILE;
  uint64_t* lb = lookback_words(tiles, s);
  const uint32_t epoch = lookback_epoch(s);
  uint32_t* meta = scratch_t<uint32_t>(S_META2, 4);
  fill_u32(meta, 0, 4, s);
  int64_t* ng_dev = reinterpret_cast<int64_t*>(meta + 2);
  {
    // reads sorted key + row id, writes the inverse (+ the distinct keys)
    ProfScope prof("unique_run_encode", s, double(n) * (sizeof(K) + 8.0));
    hipLaunchKernelGGL((k_run_encode<K>), dim3(tiles), dim3(RE_BLOCK), 0, s, sorted, perm, n,
                       uniq, inv, lb, meta, epoch, ng_dev);
    NFH_LAUNCH_CHECK();
  }
  return read_scalar(ng_dev, s);
}

template int64_t unique_inverse_device<int32_t>(const int32_t*, size_t, int32_t*, int32_t*,
                                                int32_t*, uint32_t*, uint32_t*, int32_t*,
                                                hipStream_t);
template int64_t unique_inverse_device<int64_t>(const int64_t*, size_t, int64_t*, int32_t*,
                                                int32_t*, uint64_t*, uint64_t*, int32_t*,
                                                hipStream_t);

template <class K>
static int64_t unique_entry(const K* keys, size_t n, K* uniq_out, int32_t* inv_out,
                            int32_t* perm, word_t<K>* tmp0, word_t<K>* tmp1, int32_t* tmp_p) {
  Stage st;
  const K* k = st.in(keys, n, S_IN0);
  K* u = st.out(uniq_out, n, S_OUT0);
  int32_t* iv = st.out(inv_out, n, S_OUT1);
  int32_t* p = st.tmp(perm, n, S_T3);
  word_t<K>* a = st.tmp(tmp0, n, S_T0);
  word_t<K>* b = st.tmp(tmp1, n, S_T1);
  int32_t* c = st.tmp(tmp_p, n, S_T2);
  const int64_t ng = unique_inverse_device<K>(k, n, u, iv, p, a, b, c, st.s);
  st.set_out_bytes(uniq_out, size_t(ng) * sizeof(K));
  st.finish();
  return ng;
}

}  // namespace nfh

using namespace nfh;

// lib.rs:1420-1456
NF_EXPORT int64_t nf_unique_inverse_i32(const int32_t* keys, size_t n, int32_t* uniq_out,
                                        int32_t* inv_out, int32_t* perm, uint32_t* tmp0,
                                        uint32_t* tmp1, int32_t* tmp_p) {
  if (n == 0) return 0;
  if (!keys || !uniq_out || !inv_out || !perm || !tmp0 || !tmp1 || !tmp_p) return -2;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return unique_entry<int32_t>(keys, n, uniq_out, inv_out, perm, tmp0, tmp1, tmp_p);
  NFH_GUARD_END(int64_t)
}

// lib.rs:1464-1500
NF_EXPORT int64_t nf_unique_inverse_i64(const int64_t* keys, size_t n, int64_t* uniq_out,
                                        int32_t* inv_out, int32_t* perm, uint64_t* tmp0,
                                        uint64_t* tmp1, int32_t* tmp_p) {
  if (n == 0) return 0;
  if (!keys || !uniq_out || !inv_out || !perm || !tmp0 || !tmp1 || !tmp_p) return -2;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return unique_entry<int64_t>(keys, n, uniq_out, inv_out, perm, tmp0, tmp1, tmp_p);
  NFH_GUARD_END(int64_t)
}
