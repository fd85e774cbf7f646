// select.hip -- stable selection (IR `filter`), mask algebra and the
// filter -> derive -> sort -> limit pipeline.
//
// Semantics: numfast-native/src/series/select.rs:13-67 (select_count,
// select_scatter order-preserving compaction, mask and/or/not normalised
// to 0/1) behind lib.rs:616-1062.  Return codes as there: counts for the
// count / scatter symbols, -2 on a null pointer; 0 / -1 for the masks.
//
// Compaction kernel: one workgroup per 4096-row tile (256 threads x 4
// quads of 4 rows; quad-vector loads of values and mask bytes).  Row order
// inside a tile is (quad slot j, wave, lane, element c), so a wave ranks its
// kept rows with one ballot per (j, c) and v_mbcnt; the 16 (j, wave) counts
// are scanned in LDS; tiles chain through decoupled look-back.  Every kept
// row is written once, in order, so the output is written with full lines.
#include <mutex>

#include "device.hpp"
#include "learned.hpp"
#include "runtime.hpp"
#include "select.hpp"
#include "sort.hpp"

namespace nfh {

constexpr int CP_BLOCK = 256;
constexpr int CP_WAVES = CP_BLOCK / WAVE;
constexpr int CP_Q = 4;                            // quads per thread
constexpr int CP_TILE = CP_BLOCK * CP_Q * 4;        // rows per tile

template <class T>
struct alignas(4 * sizeof(T)) Quad {
  T v[4];
};

// Stable compaction skeleton.  Op provides a register type Reg and
//   uint32_t keep(uint32_t q, uint32_t n, Reg& r) -> 4-bit keep mask of quad
//       q (rows 4q .. 4q+3), loading what emit needs into r;
//   void emit(const Reg& r, int c, uint32_t q, uint32_t pos) -> write row
//       4q + c at output position pos.
// keep() runs for every quad of the thread before any emit().
template <class Op>
__global__ __launch_bounds__(CP_BLOCK) void k_compact(Op op, uint32_t n,
                                                      uint64_t* __restrict__ lookback,
                                                      uint32_t* __restrict__ tile_ctr,
                                                      uint32_t epoch,
                                                      uint32_t* __restrict__ total) {
  __shared__ uint32_t s_cnt[CP_Q][CP_WAVES];
  __shared__ uint32_t s_tile, s_prefix;
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t qbase = tile * uint32_t(CP_TILE / 4);
  const uint64_t lt = lanemask_lt();

  uint32_t kb[CP_Q];
  typename Op::Reg reg[CP_Q];
#pragma unroll
  for (int j = 0; j < CP_Q; ++j) kb[j] = op.keep(qbase + uint32_t(j * CP_BLOCK + tid), n, reg[j]);

  // in-wave exclusive ranks per (j, c) and per-(j, wave) counts
  uint32_t r0[CP_Q];
#pragma unroll
  for (int j = 0; j < CP_Q; ++j) {
    uint32_t before = 0, wcnt = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint64_t b = __ballot((kb[j] >> c) & 1u);
      before += uint32_t(__popcll(b & lt));
      wcnt += uint32_t(__popcll(b));
    }
    r0[j] = before;
    if (lane == 0) s_cnt[j][wave] = wcnt;
  }
  __syncthreads();
  // (j, wave) order prefix; thread 0 also publishes / looks back
  if (tid == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int j = 0; j < CP_Q; ++j)
#pragma unroll
      for (int w = 0; w < CP_WAVES; ++w) {
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
    s_prefix = ex;
    const uint32_t ntiles = (n + CP_TILE - 1) / CP_TILE;
    if (tile == ntiles - 1) *total = ex + t;
  }
  __syncthreads();
  const uint32_t pre = s_prefix;
#pragma unroll
  for (int j = 0; j < CP_Q; ++j) {
    uint32_t pos = pre + s_cnt[j][wave] + r0[j];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if ((kb[j] >> c) & 1u) {
        op.emit(reg[j], c, qbase + uint32_t(j * CP_BLOCK + tid), pos);
        ++pos;
      }
    }
  }
}

// out[m++] = src[i] where mask[i] != 0.
template <class T, bool VEC>
struct SelectOp {
  const T* src;
  const uint8_t* mask;
  T* out;
  struct Reg {
    T v[4];
  };
  __device__ uint32_t keep(uint32_t q, uint32_t n, Reg& r) const {
    const uint32_t i0 = q * 4u;
    if (i0 >= n) return 0u;
    if (VEC && i0 + 4u <= n) {
      const Quad<T> x = reinterpret_cast<const Quad<T>*>(src)[q];
      const uint32_t m = reinterpret_cast<const uint32_t*>(mask)[q];
#pragma unroll
      for (int c = 0; c < 4; ++c) r.v[c] = x.v[c];
      return uint32_t((m & 0xFFu) != 0) | (uint32_t((m & 0xFF00u) != 0) << 1) |
             (uint32_t((m & 0xFF0000u) != 0) << 2) | (uint32_t((m & 0xFF000000u) != 0) << 3);
    }
    uint32_t k = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t i = i0 + uint32_t(c);
      if (i < n && mask[i] != 0) {
        r.v[c] = src[i];
        k |= 1u << c;
      }
    }
    return k;
  }
  __device__ void emit(const Reg& r, int c, uint32_t, uint32_t pos) const { out[pos] = r.v[c]; }
};

// count of mask != 0 (no output rows)
struct CountOp {
  const uint8_t* mask;
  struct Reg {};
  __device__ uint32_t keep(uint32_t q, uint32_t n, Reg&) const {
    const uint32_t i0 = q * 4u;
    uint32_t k = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t i = i0 + uint32_t(c);
      if (i < n && mask[i] != 0) k |= 1u << c;
    }
    return k;
  }
  __device__ void emit(const Reg&, int, uint32_t, uint32_t) const {}
};

// filter(v > 0) -> derive(w = v * mul, int32 wrap): w[m] and row[m].
template <bool VEC>
struct FilterDeriveOp {
  const int32_t* v;
  int32_t mul;
  int32_t* w;
  int32_t* row;
  struct Reg {
    int32_t v[4];
  };
  __device__ uint32_t keep(uint32_t q, uint32_t n, Reg& r) const {
    const uint32_t i0 = q * 4u;
    if (i0 >= n) return 0u;
    if (VEC && i0 + 4u <= n) {
      const int4 x = reinterpret_cast<const int4*>(v)[q];
      r.v[0] = x.x;
      r.v[1] = x.y;
      r.v[2] = x.z;
      r.v[3] = x.w;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) r.v[c] = (i0 + uint32_t(c) < n) ? v[i0 + c] : 0;
    }
    uint32_t k = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) k |= uint32_t(r.v[c] > 0) << c;
    return k;
  }
  __device__ void emit(const Reg& r, int c, uint32_t q, uint32_t pos) const {
    // NumPy int32 a*b then astype(int32): two's-complement wrap
    w[pos] = int32_t(uint32_t(r.v[c]) * uint32_t(mul));
    row[pos] = int32_t(q * 4u + uint32_t(c));
  }
};

template <class Op>
static uint32_t run_compact(const Op& op, uint32_t n, const char* prof_name, double bytes,
                            hipStream_t s) {
  const uint32_t tiles = (n + CP_TILE - 1) / CP_TILE;
  uint64_t* lb = lookback_words(tiles, s);
  const uint32_t epoch = lookback_epoch(s);
  uint32_t* ctr = scratch_t<uint32_t>(S_META2, 8);
  fill_u32(ctr, 0, 2, s);
  {
    ProfScope prof(prof_name, s, bytes);
    hipLaunchKernelGGL(k_compact<Op>, dim3(tiles), dim3(CP_BLOCK), 0, s, op, n, lb, ctr, epoch,
                       ctr + 1);
    NFH_LAUNCH_CHECK();
  }
  return read_scalar(ctr + 1, s);
}

static bool al(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

template <class T>
uint32_t select_scatter_device(const T* src, const uint8_t* mask, uint32_t n, T* out,
                               hipStream_t s) {
  const double bytes = double(n) * (sizeof(T) + 1.0);
  if (al(src, 4 * sizeof(T)) && al(mask, 4))
    return run_compact(SelectOp<T, true>{src, mask, out}, n, "select_scatter", bytes, s);
  return run_compact(SelectOp<T, false>{src, mask, out}, n, "select_scatter", bytes, s);
}

uint32_t filter_derive_device(const int32_t* v, uint32_t n, int32_t mul, int32_t* w,
                              int32_t* row, hipStream_t s) {
  const double bytes = double(n) * 4.0;  // + 8 B per kept row, added by the caller's model
  if (al(v, 16))
    return run_compact(FilterDeriveOp<true>{v, mul, w, row}, n, "filter_derive", bytes, s);
  return run_compact(FilterDeriveOp<false>{v, mul, w, row}, n, "filter_derive", bytes, s);
}

__global__ void k_mask_op(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                          uint32_t n, int op, uint8_t* __restrict__ out) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const bool x = a[i] != 0;
    out[i] = uint8_t(op == 0 ? (x && b[i] != 0) : op == 1 ? (x || b[i] != 0) : !x);
  }
}

// out[k] = vals[perm[k]], rows_out[k] = rows[perm[k]] for k < L.
__global__ void k_gather_limit(const int32_t* __restrict__ perm, const int32_t* __restrict__ w,
                               const int32_t* __restrict__ rows, uint32_t L,
                               int32_t* __restrict__ w_out, int32_t* __restrict__ rows_out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L) return;
  const int32_t p = perm[k];
  if (w_out) w_out[k] = w[p];
  if (rows_out) rows_out[k] = rows[p];
}

template uint32_t select_scatter_device<int32_t>(const int32_t*, const uint8_t*, uint32_t,
                                                 int32_t*, hipStream_t);

}  // namespace nfh

using namespace nfh;

// lib.rs:630-641 -- kept rows of a u8 mask; -2 on a null pointer.
NF_EXPORT int64_t nf_select_count(const uint8_t* mask, size_t n) {
  if (n == 0) return 0;
  if (!mask) return int64_t(BAD_RANGE);
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const uint8_t* m = st.in(mask, n, S_IN0);
  const uint32_t r = run_compact(CountOp{m}, uint32_t(n), "select_count", double(n), st.s);
  st.finish();
  return int64_t(r);
  NFH_GUARD_END(int64_t)
}

template <class T>
static int64_t select_scatter_entry(const T* src, const uint8_t* mask, size_t n, T* out) {
  if (n == 0) return 0;
  if (!src || !mask || !out) return int64_t(BAD_RANGE);
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const T* s = st.in(src, n, S_IN0);
  const uint8_t* m = st.in(mask, n, S_IN1);
  T* o = st.out(out, n, S_OUT0);
  const uint32_t r = select_scatter_device<T>(s, m, uint32_t(n), o, st.s);
  st.set_out_bytes(out, size_t(r) * sizeof(T));
  st.finish();
  return int64_t(r);
  NFH_GUARD_END(int64_t)
}

// lib.rs:643-700 -- order-preserving compaction; returns m, -2 on null.
NF_EXPORT int64_t nf_select_scatter_i32(const int32_t* src, const uint8_t* mask, size_t n,
                                        int32_t* out) {
  return select_scatter_entry(src, mask, n, out);
}
NF_EXPORT int64_t nf_select_scatter_i64(const int64_t* src, const uint8_t* mask, size_t n,
                                        int64_t* out) {
  return select_scatter_entry(src, mask, n, out);
}
NF_EXPORT int64_t nf_select_scatter_f32(const float* src, const uint8_t* mask, size_t n,
                                        float* out) {
  return select_scatter_entry(src, mask, n, out);
}
NF_EXPORT int64_t nf_select_scatter_f64(const double* src, const uint8_t* mask, size_t n,
                                        double* out) {
  return select_scatter_entry(src, mask, n, out);
}
NF_EXPORT int64_t nf_select_scatter_u8(const uint8_t* src, const uint8_t* mask, size_t n,
                                       uint8_t* out) {
  return select_scatter_entry(src, mask, n, out);
}

static int32_t mask_entry(const uint8_t* a, const uint8_t* b, size_t n, uint8_t* out, int op) {
  if (n == 0) return OK;
  if (!a || (op < 2 && !b) || !out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const uint8_t* da = st.in(a, n, S_IN0);
  const uint8_t* db = op < 2 ? st.in(b, n, S_IN1) : nullptr;
  uint8_t* o = st.out(out, n, S_OUT0);
  hipLaunchKernelGGL(k_mask_op, dim3(grid_for(n, 256, 16 * num_cus())), dim3(256), 0, st.s, da,
                     db, uint32_t(n), op, o);
  NFH_LAUNCH_CHECK();
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

// lib.rs:1014-1062 -- elementwise mask algebra, output normalised to 0/1.
NF_EXPORT int32_t nf_mask_and(const uint8_t* a, const uint8_t* b, size_t n, uint8_t* out) {
  return mask_entry(a, b, n, out, 0);
}
NF_EXPORT int32_t nf_mask_or(const uint8_t* a, const uint8_t* b, size_t n, uint8_t* out) {
  return mask_entry(a, b, n, out, 1);
}
NF_EXPORT int32_t nf_mask_not(const uint8_t* a, size_t n, uint8_t* out) {
  return mask_entry(a, nullptr, n, out, 2);
}

// filter(v > 0) -> derive(w = v * mul, int32 wrap) -> stable sort by w ->
// limit(L): the first min(L, m) rows as (w, source row).  Returns that
// count, -1 on a null pointer, -3 when n > i32::MAX.  With a model (from
// nfh_cdf_train_i32 on the derived column) the sort is the learned-CDF
// bucket sort instead of the LSD radix sort; the output is identical.
// Semantics: src/Drivers/CPU/_lib/cpu.py:3034-3035, :3074 (int32 wrap of
// the derived column), cpu.py:1174-1225 (stable sort), Chain.sort / limit
// (src/Semantic/TableExpr/_lib/chain.py:978-1008).
NF_EXPORT int64_t nfh_filter_derive_sort_limit_i32(const int32_t* v, size_t n, int32_t mul,
                                                   size_t limit, const CdfModel* model,
                                                   int32_t* w_out, int32_t* rows_out) {
  if (n == 0 || limit == 0) return 0;
  if (!v || (!w_out && !rows_out)) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* dv = st.in(v, n, S_IN0);
  int32_t* w = scratch_t<int32_t>(S_WORK0, n);
  int32_t* rows = scratch_t<int32_t>(S_WORK1, n);
  const uint32_t m = filter_derive_device(dv, uint32_t(n), mul, w, rows, st.s);
  const uint32_t L = uint32_t(std::min<size_t>(limit, m));
  if (L > 0) {
    int32_t* perm = scratch_t<int32_t>(S_WORK2, m);
    uint32_t* t0 = scratch_t<uint32_t>(S_T0, m);
    uint32_t* t1 = scratch_t<uint32_t>(S_T1, m);
    int32_t* tp = scratch_t<int32_t>(S_T2, m);
    if (model != nullptr)
      learned_argsort_i32(w, m, model, perm, st.s);
    else
      radix_argsort<int32_t>(w, m, false, perm, t0, t1, tp, nullptr, st.s);
    int32_t* wo = w_out ? st.out(w_out, L, S_OUT0) : nullptr;
    int32_t* ro = rows_out ? st.out(rows_out, L, S_OUT1) : nullptr;
    hipLaunchKernelGGL(k_gather_limit, dim3((L + 255) / 256), dim3(256), 0, st.s, perm, w, rows,
                       L, wo, ro);
    NFH_LAUNCH_CHECK();
  }
  st.finish();
  return int64_t(L);
  NFH_GUARD_END(int64_t)
}
