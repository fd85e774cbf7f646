// groupby.hip -- dense-code group-by engine and the nf_group_* entry points.
//
// Reference semantics (numfast-native):
//   dense f64 scatter        groupby/dense.rs:18-37       nf_group_sum_count
//   SoA multi-column f64     groupby/fused.rs:25-50       nf_group_multi_sum_count
//   int32/int64 checked      groupby/scaled.rs:27-105     nf_group_variant_i32/_i64
//   sorted-dense variants    groupby/scaled.rs:113-254    variant 2
//   variant dispatch         groupby/variants.rs:45-101
//   sum-only + count-only    groupby/mixed.rs:42-90       nf_group_mixed_sum_only/count_only
//   owner shard              groupby/mixed.rs:106-149     nf_group_owner_2i32_1f64
//   fused pack + aggregate   groupby/pack_fused.rs:26-90  nf_pack_sum_count_i32/_f64
//
// GPU design.  One accumulate pass visits every row once: the key is
// range-checked (bad rows feed a wave-reduced atomicMin of the first bad
// row), and the row's values are added into per-group accumulators --
// privatised in LDS per workgroup when the group range fits, else global
// atomics.  Integer sums accumulate exactly in 64 bits together with the
// sum of |v|; a group whose sum of |v| stays inside the target type cannot
// overflow at any row-order prefix, so only groups above that bound need
// the exact row-order check (stable radix sort of those groups' rows, then
// one wave per group scanning its rows in row order).  Float64 sums use the
// order-independent exact accumulation of exactsum.hpp.
#include <algorithm>

#include "device.hpp"
#include "exactsum.hpp"
#include "groupby.hpp"
#include "runtime.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace nfh {

typedef unsigned long long u64;

struct KJob {
  const int32_t* keys;
  const int32_t* k2;
  long long m2;
  uint32_t n;
  u64 g, lo, R;
  int ni32, ni64, nf64;
  bool chk32, chk64;
  const int32_t* v32[MAXC];
  const long long* v64[MAXC];
  const double* vf[MAXC];
  // accumulator array indices (arrays of R words)
  int i_s32[MAXC], i_a32[MAXC], i_s64[MAXC], i_a64h[MAXC], i_a64l[MAXC], i_fc[MAXC];
  int i_ff[MAXC];  // into the 32-bit arrays (array 0 is the count)
  int i_fb[MAXC];  // f64: sum of ceil(|v| * 2^(30-E)), a bound on sum |v|
  bool f64_seq_first;  // sequential f64 sums start from the first value
  int n64arr, n32arr;
  u64* acc64;       // n64arr * R
  uint32_t* acc32;  // n32arr * R
  const u64* emax;  // per f64 column: max finite |v| bits
  u64* err;         // first bad-key row (min), ~0 if none
};

// Key of row i; false for a bad key (pack failure or outside [0, g)).
__device__ __forceinline__ bool resolve_key(const KJob& J, uint32_t i, u64& k) {
  long long kk;
  if (J.k2 != nullptr) {
    // core/numeric.rs:29-38: negative operand or a*m+b > i32::MAX fails;
    // the packed value is then taken as int32 (Rust `as i32`).
    const int32_t a = J.keys[i], b = J.k2[i];
    if (a < 0 || b < 0) return false;
    const long long v = (long long)a * J.m2 + (long long)b;
    if (v > 2147483647ll) return false;
    kk = (long long)(int32_t)(uint32_t)(unsigned long long)v;
  } else {
    kk = J.keys[i];
  }
  const u64 u = (u64)kk;  // widen_usize: negative -> huge -> out of range
  if (u >= J.g) return false;
  k = u;
  return true;
}

__device__ __forceinline__ void wave_first_bad(bool bad, uint32_t row, u64* err) {
  const u64 m = __ballot(bad);
  if (m != 0) {
    const int first = __ffsll((unsigned long long)m) - 1;
    if (lane_id() == first) atomicMin(err, (u64)row);
  }
}

template <bool LDS>
__global__ __launch_bounds__(256) void k_dense_acc(KJob J) {
  extern __shared__ __align__(16) unsigned char smem[];
  u64* l64 = reinterpret_cast<u64*>(smem);
  uint32_t* l32 = reinterpret_cast<uint32_t*>(smem + size_t(J.n64arr) * J.R * 8);
  u64* a64 = LDS ? l64 : J.acc64;
  uint32_t* a32 = LDS ? l32 : J.acc32;
  if (LDS) {
    for (u64 i = threadIdx.x; i < u64(J.n64arr) * J.R; i += blockDim.x) l64[i] = 0;
    for (u64 i = threadIdx.x; i < u64(J.n32arr) * J.R; i += blockDim.x) l32[i] = 0;
    __syncthreads();
  }
  int E[MAXC];
  for (int c = 0; c < J.nf64; ++c) E[c] = xs_exponent_of_bits(J.emax[c]);
  const uint32_t stride = gridDim.x * blockDim.x;
  // Uniform trip count so every lane reaches the ballot in wave_first_bad.
  const uint32_t start = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t first_of_wave = start - lane_id();
  for (uint32_t base = first_of_wave; base < J.n; base += stride) {
    const uint32_t i = base + lane_id();
    u64 k = 0;
    const bool in = i < J.n;
    const bool ok = in && resolve_key(J, i, k);
    wave_first_bad(in && !ok, i, J.err);
    if (!ok) continue;
    if (k < J.lo || k - J.lo >= J.R) continue;
    const u64 r = k - J.lo;
    if (LDS || J.acc32 != nullptr) atomicAdd(&a32[r], 1u);
    for (int c = 0; c < J.ni32; ++c) {
      const long long v = J.v32[c][i];
      atomicAdd(&a64[size_t(J.i_s32[c]) * J.R + r], (u64)v);
      if (J.chk32) atomicAdd(&a64[size_t(J.i_a32[c]) * J.R + r], (u64)(v < 0 ? -v : v));
    }
    for (int c = 0; c < J.ni64; ++c) {
      const long long v = J.v64[c][i];
      atomicAdd(&a64[size_t(J.i_s64[c]) * J.R + r], (u64)v);
      if (J.chk64) {
        const u64 av = v < 0 ? (u64)0 - (u64)v : (u64)v;
        atomicAdd(&a64[size_t(J.i_a64h[c]) * J.R + r], av >> 32);
        atomicAdd(&a64[size_t(J.i_a64l[c]) * J.R + r], av & 0xFFFFFFFFull);
      }
    }
    for (int c = 0; c < J.nf64; ++c) {
      const double v = J.vf[c][i];
      if (xs_finite(v)) {
        int64_t ch[XS_CHUNKS];
        xs_chunks(v, E[c], ch);
#pragma unroll
        for (int q = 0; q < XS_CHUNKS; ++q)
          if (ch[q] != 0) atomicAdd(&a64[size_t(J.i_fc[c] + q) * J.R + r], (u64)ch[q]);
        const u64 b = (u64)ceil(ldexp(fabs(v), 30 - E[c]));
        if (b != 0) atomicAdd(&a64[size_t(J.i_fb[c]) * J.R + r], b);
      } else {
        atomicOr(&a32[size_t(J.i_ff[c]) * J.R + r], xs_special_flag(v));
      }
    }
  }
  if (LDS) {
    __syncthreads();
    for (u64 r = threadIdx.x; r < J.R; r += blockDim.x) {
      const uint32_t c = l32[r];
      if (c == 0) continue;
      atomicAdd(&J.acc32[r], c);
      for (int a = 0; a < J.n64arr; ++a) {
        const u64 v = l64[size_t(a) * J.R + r];
        if (v != 0) atomicAdd(&J.acc64[size_t(a) * J.R + r], v);
      }
      for (int a = 1; a < J.n32arr; ++a) {
        const uint32_t v = l32[size_t(a) * J.R + r];
        if (v != 0) atomicOr(&J.acc32[size_t(a) * J.R + r], v);
      }
    }
  }
}

// Largest finite |v| (as bits) per float64 column.
__global__ __launch_bounds__(256) void k_absmax(const double* __restrict__ v, uint32_t n,
                                                u64* __restrict__ out) {
  u64 m = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const u64 b = (u64)__double_as_longlong(v[i]) & 0x7FFFFFFFFFFFFFFFull;
    if ((b >> 52) != 0x7FF && b > m) m = b;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const u64 o = __shfl_down(m, off);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m != 0) atomicMax(out, m);
}

struct FinOut {
  int32_t* o32[MAXC];
  int64_t* o32w[MAXC];  // int64 sums of int32 values
  int64_t* o64[MAXC];
  double* of[MAXC];
  int64_t* counts;
  uint8_t* amb;      // per slot: needs the ordered check
  uint32_t* anyamb;  // set when any slot needs it
};

__global__ __launch_bounds__(256) void k_dense_fin(KJob J, FinOut O) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < J.R; r += stride) {
    const u64 k = J.lo + r;
    bool amb = false, ambf = false;
    if (O.counts) O.counts[k] = (int64_t)J.acc32[r];
    for (int c = 0; c < J.ni32; ++c) {
      const long long s = (long long)J.acc64[size_t(J.i_s32[c]) * J.R + r];
      if (J.chk32 && J.acc64[size_t(J.i_a32[c]) * J.R + r] > 2147483647ull) amb = true;
      if (O.o32[c]) O.o32[c][k] = (int32_t)s;
      if (O.o32w[c]) O.o32w[c][k] = s;
    }
    for (int c = 0; c < J.ni64; ++c) {
      const u64 s = J.acc64[size_t(J.i_s64[c]) * J.R + r];
      if (J.chk64) {
        const unsigned __int128 b =
            ((unsigned __int128)J.acc64[size_t(J.i_a64h[c]) * J.R + r] << 32) +
            J.acc64[size_t(J.i_a64l[c]) * J.R + r];
        if (b > (unsigned __int128)0x7FFFFFFFFFFFFFFFull) amb = true;
      }
      if (O.o64[c]) O.o64[c][k] = (int64_t)s;
    }
    for (int c = 0; c < J.nf64; ++c) {
      int64_t S[XS_CHUNKS];
#pragma unroll
      for (int q = 0; q < XS_CHUNKS; ++q) S[q] = (int64_t)J.acc64[size_t(J.i_fc[c] + q) * J.R + r];
      const uint32_t fl = J.acc32[size_t(J.i_ff[c]) * J.R + r];
      const int E = xs_exponent_of_bits(J.emax[c]);
      const double sum = xs_finish(S, E, fl);
      if (O.of[c]) O.of[c][k] = sum;
      // The row-order sum of the reference differs from the exact sum by at
      // most (cnt-1) * 2^-53 * sum|v|.  When that bound could exceed half the
      // 1e-12 tolerance the group's sum is redone in row order (bit-exact).
      if (fl == 0 && O.amb) {
        const double B = ldexp(double(J.acc64[size_t(J.i_fb[c]) * J.R + r]), E - 30) *
                         (1.0 + 0x1p-40);
        const double cnt = double(J.acc32[r]);
        if (cnt * 0x1p-53 * B > 0.5e-12 * fmax(1.0, fabs(sum))) ambf = true;
      }
    }
    if (O.amb) O.amb[r] = uint8_t((amb ? 1 : 0) | (ambf ? 2 : 0));
    if (amb || ambf) atomicOr(O.anyamb, 1u);
  }
}

// ---------------------------------------------------------------- ordered check
// Rows of flagged groups, compacted in row order (key slot kept per row).
struct FlagRowsOp {
  KJob J;
  const uint8_t* amb;
  int32_t* rows;
  int32_t* slots;
  uint32_t* total_out;
  __device__ bool flagged(uint32_t i, u64& r) const {
    u64 k;
    if (!resolve_key(J, i, k)) return false;
    if (k < J.lo || k - J.lo >= J.R) return false;
    r = k - J.lo;
    return amb[r] != 0;
  }
  __device__ uint32_t value(uint32_t i) const {
    u64 r;
    return flagged(i, r) ? 1u : 0u;
  }
  __device__ void emit(uint32_t i, uint32_t ex) const {
    u64 r;
    if (flagged(i, r)) {
      rows[ex] = int32_t(i);
      slots[ex] = int32_t(r);
    }
  }
  __device__ void total(uint32_t t) const { *total_out = t; }
};

// Start of each flagged slot's rows in key order.
struct SlotStartOp {
  const uint8_t* amb;
  const uint32_t* cnt;
  uint32_t* start;
  __device__ uint32_t value(uint32_t r) const { return amb[r] ? cnt[r] : 0u; }
  __device__ void emit(uint32_t r, uint32_t ex) const { start[r] = ex; }
  __device__ void total(uint32_t) const {}
};

__device__ __forceinline__ void shfl_up_i128(long long& lo, long long& hi, int off) {
  lo = __shfl_up(lo, off);
  hi = __shfl_up(hi, off);
}

// One wave per flagged slot: scan the slot's rows in row order and record
// the first row whose running sum leaves the target type.
__global__ __launch_bounds__(256) void k_ordered_scan(KJob J, const uint8_t* __restrict__ amb,
                                                      const uint32_t* __restrict__ start,
                                                      const int32_t* __restrict__ rows,
                                                      const int32_t* __restrict__ perm,
                                                      u64* __restrict__ ovf) {
  const int lane = lane_id();
  const u64 waves = u64(gridDim.x) * (blockDim.x / WAVE);
  const u64 w0 = u64(blockIdx.x) * (blockDim.x / WAVE) + wave_id();
  for (u64 r = w0; r < J.R; r += waves) {
    if (!(amb[r] & 1)) continue;
    const uint32_t st = start[r];
    const uint32_t cnt = J.acc32[r];
    for (int c = 0; c < J.ni32 + J.ni64; ++c) {
      const bool is32 = c < J.ni32;
      if (is32 && !J.chk32) continue;
      if (!is32 && !J.chk64) continue;
      // running sum as signed 128-bit (lo, hi)
      __int128 carry = 0;
      for (uint32_t t0 = 0; t0 < cnt; t0 += WAVE) {
        const uint32_t t = t0 + lane;
        const bool valid = t < cnt;
        uint32_t row = 0;
        __int128 v = 0;
        if (valid) {
          const uint32_t pos = st + t;
          row = uint32_t(rows[perm ? perm[pos] : pos]);
          v = is32 ? (__int128)J.v32[c][row] : (__int128)J.v64[c - J.ni32][row];
        }
        // wave inclusive scan of v
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
          long long lo = (long long)(unsigned long long)v;
          long long hi = (long long)(v >> 64);
          shfl_up_i128(lo, hi, off);
          const __int128 o = ((__int128)hi << 64) | (unsigned long long)lo;
          if (lane >= off) v += o;
        }
        const __int128 run = carry + v;
        const __int128 mx = is32 ? (__int128)2147483647 : (__int128)0x7FFFFFFFFFFFFFFFll;
        const __int128 mn = is32 ? -(__int128)2147483648ll : -(__int128)0x7FFFFFFFFFFFFFFFll - 1;
        const bool bad = valid && (run > mx || run < mn);
        const u64 bm = __ballot(bad);
        if (bm != 0) {
          const int first = __ffsll((unsigned long long)bm) - 1;
          if (lane == first) atomicMin(ovf, (u64)row);
          break;
        }
        long long lo = (long long)(unsigned long long)run;
        long long hi = (long long)(run >> 64);
        lo = __shfl(lo, WAVE - 1);
        hi = __shfl(hi, WAVE - 1);
        carry = ((__int128)hi << 64) | (unsigned long long)lo;
      }
    }
  }
}

// One wave per slot flagged for a row-order float64 sum: the wave loads 64
// values at a time and every lane adds them in row order, so the result is
// the reference's sequential sum bit for bit (dense.rs:32-36 starts from
// 0.0; the sorted-run lanes, sorted.rs:36-52, start from the first value).
__global__ __launch_bounds__(256) void k_ordered_f64(KJob J, const uint8_t* __restrict__ amb,
                                                     const uint32_t* __restrict__ start,
                                                     const int32_t* __restrict__ rows,
                                                     const int32_t* __restrict__ perm,
                                                     double* o0, double* o1, double* o2,
                                                     double* o3) {
  const int lane = lane_id();
  const u64 waves = u64(gridDim.x) * (blockDim.x / WAVE);
  const u64 w0 = u64(blockIdx.x) * (blockDim.x / WAVE) + wave_id();
  double* outs[MAXC] = {o0, o1, o2, o3};
  for (u64 r = w0; r < J.R; r += waves) {
    if (!(amb[r] & 2)) continue;
    const uint32_t st = start[r];
    const uint32_t cnt = J.acc32[r];
    for (int c = 0; c < J.nf64; ++c) {
      double acc = 0.0;
      for (uint32_t t0 = 0; t0 < cnt; t0 += WAVE) {
        const uint32_t t = t0 + lane;
        double v = 0.0;
        if (t < cnt) {
          const uint32_t pos = st + t;
          v = J.vf[c][rows[perm ? perm[pos] : pos]];
        }
        const uint32_t m = (cnt - t0) < uint32_t(WAVE) ? (cnt - t0) : uint32_t(WAVE);
        for (uint32_t l = 0; l < m; ++l) {
          const double x = __shfl(v, int(l));
          acc = (t0 == 0 && l == 0 && J.f64_seq_first) ? x : acc + x;
        }
      }
      if (lane == 0 && outs[c]) outs[c][J.lo + r] = acc;
    }
  }
}

__global__ void k_first_regress(const int32_t* __restrict__ keys, uint32_t n, u64* out) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t start = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t base = start - lane_id(); base < n; base += stride) {
    const uint32_t i = base + lane_id();
    const bool bad = i > 0 && i < n && keys[i] < keys[i - 1];
    wave_first_bad(bad, i, out);
  }
}

size_t first_regress_row(const int32_t* keys, size_t n, hipStream_t s) {
  if (n < 2) return n;
  u64* d = scratch_t<u64>(S_WORK2, 1);
  fill_u64(d, ~0ull, 1, s);
  hipLaunchKernelGGL(k_first_regress, dim3(grid_for(n, 256, 8 * num_cus())), dim3(256), 0, s,
                     keys, uint32_t(n), d);
  NFH_LAUNCH_CHECK();
  const u64 r = read_scalar(d, s);
  return r == ~0ull ? n : size_t(r);
}

static size_t lds_limit() { return 64 * 1024; }

AggResult dense_aggregate(const DenseSpec& sp, hipStream_t s) {
  AggResult res{0, sp.n};
  if (sp.n == 0) {
    // Outputs still zeroed over [lo, hi) (fill_zero in the reference).
  }
  KJob J{};
  J.keys = sp.keys;
  J.k2 = sp.k2;
  J.m2 = sp.m2;
  J.n = uint32_t(sp.n);
  J.g = sp.g;
  J.lo = sp.lo;
  J.R = sp.hi - sp.lo;
  J.ni32 = sp.ni32;
  J.ni64 = sp.ni64;
  J.nf64 = sp.nf64;
  J.chk32 = sp.check_i32;
  J.chk64 = sp.check_i64;
  int a64 = 0, a32 = 1;  // 32-bit array 0 = count
  for (int c = 0; c < sp.ni32; ++c) {
    J.v32[c] = sp.v32[c];
    J.i_s32[c] = a64++;
    J.i_a32[c] = sp.check_i32 ? a64++ : -1;
  }
  for (int c = 0; c < sp.ni64; ++c) {
    J.v64[c] = reinterpret_cast<const long long*>(sp.v64[c]);
    J.i_s64[c] = a64++;
    if (sp.check_i64) {
      J.i_a64h[c] = a64++;
      J.i_a64l[c] = a64++;
    }
  }
  for (int c = 0; c < sp.nf64; ++c) {
    J.vf[c] = sp.vf[c];
    J.i_fc[c] = a64;
    a64 += XS_CHUNKS;
    J.i_fb[c] = a64++;
    J.i_ff[c] = a32++;
  }
  J.f64_seq_first = sp.f64_seq_first;
  J.n64arr = a64;
  J.n32arr = a32;
  const u64 R = J.R;
  if (R == 0 && sp.n == 0) return res;
  // scratch: acc64 | acc32 | emax[MAXC] | err | anyamb
  const size_t b64 = size_t(a64) * R * 8, b32 = size_t(a32) * R * 4;
  const size_t off32 = b64, offm = (b64 + b32 + 15) & ~size_t(15);
  unsigned char* base = static_cast<unsigned char*>(scratch(S_WORK0, offm + 128));
  J.acc64 = reinterpret_cast<u64*>(base);
  J.acc32 = reinterpret_cast<uint32_t*>(base + off32);
  u64* emax = reinterpret_cast<u64*>(base + offm);
  J.emax = emax;
  J.err = emax + MAXC;
  uint32_t* anyamb = reinterpret_cast<uint32_t*>(emax + MAXC + 1);
  // zero accumulators + metadata, err = ~0
  fill_u32(base, 0, (offm + 128) / 4, s);
  fill_u64(J.err, ~0ull, 1, s);
  for (int c = 0; c < sp.nf64; ++c) {
    if (sp.n == 0) break;
    hipLaunchKernelGGL(k_absmax, dim3(grid_for(sp.n, 256, 4 * num_cus())), dim3(256), 0, s,
                       sp.vf[c], uint32_t(sp.n), emax + c);
    NFH_LAUNCH_CHECK();
  }
  if (sp.n > 0) {
    const size_t lds = b64 + b32;
    if (R > 0 && lds <= lds_limit()) {
      const unsigned blocks = std::max(1u, std::min<unsigned>(4 * num_cus(), unsigned((sp.n + 4095) / 4096)));
      hipLaunchKernelGGL(k_dense_acc<true>, dim3(blocks), dim3(256), lds, s, J);
    } else {
      hipLaunchKernelGGL(k_dense_acc<false>, dim3(grid_for(sp.n, 256, 16 * num_cus())),
                         dim3(256), 0, s, J);
    }
    NFH_LAUNCH_CHECK();
  }
  FinOut O{};
  for (int c = 0; c < sp.ni32; ++c) {
    O.o32[c] = sp.o32[c].s32;
    O.o32w[c] = sp.o32[c].s64;
  }
  for (int c = 0; c < sp.ni64; ++c) O.o64[c] = sp.o64[c].s64;
  for (int c = 0; c < sp.nf64; ++c) O.of[c] = sp.of[c].f64;
  O.counts = sp.counts;
  const bool need_check =
      (sp.check_i32 && sp.ni32 > 0) || (sp.check_i64 && sp.ni64 > 0) || sp.nf64 > 0;
  uint8_t* amb = need_check ? scratch_t<uint8_t>(S_WORK1, R) : nullptr;
  O.amb = amb;
  O.anyamb = anyamb;
  if (R > 0) {
    hipLaunchKernelGGL(k_dense_fin, dim3(grid_for(R, 256, 8 * num_cus())), dim3(256), 0, s, J,
                       O);
    NFH_LAUNCH_CHECK();
  }
  // read back {err, anyamb}
  u64* h = static_cast<u64*>(pinned(16));
  NFH_CHECK(hipMemcpyAsync(h, J.err, 16, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  const u64 err_row = h[0];
  const bool any_amb = (uint32_t(h[1]) != 0);
  u64 ovf_row = ~0ull;
  if (any_amb) {
    // rows of flagged slots in row order, then stable-sorted by slot
    const size_t n = sp.n;
    int32_t* rows = scratch_t<int32_t>(S_T0, n);
    int32_t* slots = scratch_t<int32_t>(S_T1, n);
    uint32_t* m_dev = scratch_t<uint32_t>(S_MISC, 4);
    fill_u32(m_dev, 0, 1, s);
    scan_tiles(FlagRowsOp{J, amb, rows, slots, m_dev}, n, s);
    const uint32_t m = read_scalar(m_dev, s);
    const int32_t* perm = nullptr;
    if (!sp.presorted && m > 0) {
      int32_t* p = scratch_t<int32_t>(S_T2, m);
      uint32_t* t0 = scratch_t<uint32_t>(S_T3, m);
      uint32_t* t1 = scratch_t<uint32_t>(S_T4, m);
      int32_t* tp = scratch_t<int32_t>(S_T5, m);
      radix_argsort<int32_t>(slots, m, false, p, t0, t1, tp, nullptr, s);
      perm = p;
    }
    uint32_t* start = scratch_t<uint32_t>(S_WORK3, R);
    scan_tiles(SlotStartOp{amb, J.acc32, start}, R, s);
    u64* ovf = scratch_t<u64>(S_WORK2, 1);
    fill_u64(ovf, ~0ull, 1, s);
    if (J.ni32 + J.ni64 > 0) {
      hipLaunchKernelGGL(k_ordered_scan, dim3(grid_for(R, 4, 4 * num_cus())), dim3(256), 0, s,
                         J, amb, start, rows, perm, ovf);
      NFH_LAUNCH_CHECK();
    }
    if (J.nf64 > 0) {
      hipLaunchKernelGGL(k_ordered_f64, dim3(grid_for(R, 4, 4 * num_cus())), dim3(256), 0, s, J,
                         amb, start, rows, perm, O.of[0], O.of[1], O.of[2], O.of[3]);
      NFH_LAUNCH_CHECK();
    }
    ovf_row = read_scalar(ovf, s);
  }
  if (err_row == ~0ull && ovf_row == ~0ull) return res;
  if (err_row <= ovf_row) return AggResult{BAD_RANGE, err_row};
  return AggResult{OVERFLOW_RC, ovf_row};
}

// ---------------------------------------------------------------- helpers
// Run one aggregation over more columns than MAXC by batching; counts go
// with the first batch.
static AggResult aggregate_batched(DenseSpec base, int ni32, const int32_t* i32v,
                                   int32_t* i32out, int nf64, const double* f64v,
                                   double* f64out, size_t n, size_t g) {
  hipStream_t s = stream();
  AggResult acc{0, n};
  int c32 = 0, cf = 0;
  bool first = true;
  do {
    DenseSpec sp = base;
    sp.ni32 = std::min(MAXC, ni32 - c32);
    sp.nf64 = std::min(MAXC, nf64 - cf);
    for (int c = 0; c < sp.ni32; ++c) {
      sp.v32[c] = i32v + size_t(c32 + c) * n;
      sp.o32[c].s32 = i32out + size_t(c32 + c) * g;
    }
    for (int c = 0; c < sp.nf64; ++c) {
      sp.vf[c] = f64v + size_t(cf + c) * n;
      sp.of[c].f64 = f64out + size_t(cf + c) * g;
    }
    if (!first) sp.counts = nullptr;
    acc = agg_first(acc, dense_aggregate(sp, s));
    c32 += sp.ni32;
    cf += sp.nf64;
    first = false;
  } while (c32 < ni32 || cf < nf64);
  return acc;
}

}  // namespace nfh

using namespace nfh;

// lib.rs:85-103 (dense_scatter, dense.rs:18-37).  An out-of-range key, a
// panic (process abort) in the reference, returns -2 here.
NF_EXPORT int32_t nf_group_sum_count(const int32_t* keys, const double* values, size_t n,
                                     double* sums, int64_t* counts, size_t g) {
  if (!keys || !values || !sums || !counts) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.nf64 = 1;
  sp.vf[0] = st.in(values, n, S_IN1);
  sp.of[0].f64 = st.out(sums, g, S_OUT0);
  sp.counts = st.out(counts, g, S_OUT1);
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

// lib.rs:117-135 (fused_scatter_soa, fused.rs:25-50)
NF_EXPORT int32_t nf_group_multi_sum_count(const int32_t* keys, const double* values, size_t n,
                                           size_t ncols, double* sums, int64_t* counts,
                                           size_t g) {
  if (!keys || !values || !sums || !counts) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.counts = st.out(counts, g, S_OUT1);
  const double* v = st.in(values, ncols * n, S_IN1);
  double* o = st.out(sums, ncols * g, S_OUT0);
  AggResult r = aggregate_batched(sp, 0, nullptr, nullptr, int(ncols), v, o, n, g);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

// lib.rs:150-178 (mixed_sum_only_i32_f64, mixed.rs:42-74)
NF_EXPORT int32_t nf_group_mixed_sum_only(const int32_t* keys, const int32_t* i32_vals,
                                          const double* f64_vals, size_t n, size_t n_i32,
                                          size_t n_f64, int32_t* i32_sums, double* f64_sums,
                                          size_t g) {
  if (!keys) return NULL_OR_ABORT;
  if ((n_i32 > 0 && (!i32_vals || !i32_sums)) || (n_f64 > 0 && (!f64_vals || !f64_sums)))
    return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.check_i32 = true;
  const int32_t* iv = n_i32 ? st.in(i32_vals, n_i32 * n, S_IN1) : nullptr;
  const double* fv = n_f64 ? st.in(f64_vals, n_f64 * n, S_IN2) : nullptr;
  int32_t* io = n_i32 ? st.out(i32_sums, n_i32 * g, S_OUT0) : nullptr;
  double* fo = n_f64 ? st.out(f64_sums, n_f64 * g, S_OUT1) : nullptr;
  AggResult r = aggregate_batched(sp, int(n_i32), iv, io, int(n_f64), fv, fo, n, g);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

// lib.rs:184-197 (count_only, mixed.rs:79-90)
NF_EXPORT int32_t nf_group_count_only(const int32_t* keys, size_t n, int64_t* counts,
                                      size_t g) {
  if (!keys || !counts) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.counts = st.out(counts, g, S_OUT0);
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

// lib.rs:213-250 (owner_2i32_1f64, mixed.rs:106-149): only lanes [lo, hi)
// are written; every row's key is still range-checked.
NF_EXPORT int32_t nf_group_owner_2i32_1f64(const int32_t* keys, const int32_t* a,
                                           const int32_t* b, const double* c, size_t n,
                                           size_t lo, size_t hi, int32_t* s1, int32_t* s2,
                                           double* s3, int64_t* counts, size_t g) {
  if (!keys || !a || !b || !c || !s1 || !s2 || !s3 || !counts || lo > hi || hi > g)
    return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.n = n;
  sp.g = g;
  sp.lo = lo;
  sp.hi = hi;
  sp.check_i32 = true;
  sp.ni32 = 2;
  sp.nf64 = 1;
  sp.v32[0] = st.in(a, n, S_IN1);
  sp.v32[1] = st.in(b, n, S_IN2);
  sp.vf[0] = st.in(c, n, S_IN3);
  // in/out: lanes outside [lo, hi) belong to other owners and are kept
  sp.o32[0].s32 = st.out(s1, g, S_OUT0, true);
  sp.o32[1].s32 = st.out(s2, g, S_OUT1, true);
  sp.of[0].f64 = st.out(s3, g, S_OUT2, true);
  sp.counts = st.out(counts, g, S_OUT3, true);
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

// Variant entry points: lib.rs:539-614, dispatch variants.rs:45-101.
// V_SORTED (2) reports -5 at the first regressing row unless an earlier
// row already failed (scaled.rs:113-254 check order: regress, overflow).
template <class V>
static int32_t variant_entry(const int32_t* keys, const V* vals, size_t n, uint32_t variant,
                             V* sums, int64_t* counts, size_t g) {
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(keys, n, S_IN0);
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.counts = st.out(counts, g, S_OUT1);
  const V* v = st.in(vals, n, S_IN1);
  V* o = st.out(sums, g, S_OUT0);
  if constexpr (std::is_same<V, int32_t>::value) {
    sp.ni32 = 1;
    sp.v32[0] = v;
    sp.o32[0].s32 = o;
    sp.check_i32 = true;
  } else if constexpr (std::is_same<V, int64_t>::value) {
    sp.ni64 = 1;
    sp.v64[0] = v;
    sp.o64[0].s64 = o;
    sp.check_i64 = true;
  } else {
    sp.nf64 = 1;
    sp.vf[0] = v;
    sp.of[0].f64 = o;
  }
  size_t nn = n;
  if (variant == 2) {
    nn = first_regress_row(sp.keys, n, st.s);
    sp.presorted = true;
    sp.f64_seq_first = true;
  }
  sp.n = nn;
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  if (r.rc != 0) return r.rc;
  if (nn < n) return UNSORTED;
  return OK;
}

NF_EXPORT int32_t nf_group_variant_f64(const int32_t* keys, const double* values, size_t n,
                                       uint32_t variant, double* sums, int64_t* counts,
                                       size_t g) {
  if (!keys || !values || !sums || !counts) return NULL_OR_ABORT;
  if (variant > 2) return BAD_VARIANT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return variant_entry<double>(keys, values, n, variant, sums, counts, g);
  NFH_GUARD_END(int32_t)
}

NF_EXPORT int32_t nf_group_variant_i64(const int32_t* keys, const int64_t* ticks, size_t n,
                                       uint32_t variant, int64_t* sums, int64_t* counts,
                                       size_t g) {
  if (!keys || !ticks || !sums || !counts) return NULL_OR_ABORT;
  if (variant > 2) return BAD_VARIANT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return variant_entry<int64_t>(keys, ticks, n, variant, sums, counts, g);
  NFH_GUARD_END(int32_t)
}

// The int32 family has no V_CHECKED lane: variant 1 is -3 (variants.rs:88-101).
NF_EXPORT int32_t nf_group_variant_i32(const int32_t* keys, const int32_t* ticks, size_t n,
                                       uint32_t variant, int32_t* sums, int64_t* counts,
                                       size_t g) {
  if (!keys || !ticks || !sums || !counts) return NULL_OR_ABORT;
  if (variant != 0 && variant != 2) return BAD_VARIANT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return variant_entry<int32_t>(keys, ticks, n, variant, sums, counts, g);
  NFH_GUARD_END(int32_t)
}

// lib.rs:276-331 (pack_fused_i32/_f64, pack_fused.rs:26-90)
NF_EXPORT int32_t nf_pack_sum_count_i32(const int32_t* k1, const int32_t* k2, int32_t m2,
                                        const int32_t* vals, size_t n, int32_t* sums,
                                        int64_t* counts, size_t g) {
  if (!k1 || !k2 || !vals || !sums || !counts) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(k1, n, S_IN0);
  sp.k2 = st.in(k2, n, S_IN2);
  sp.m2 = m2;
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.ni32 = 1;
  sp.check_i32 = true;
  sp.v32[0] = st.in(vals, n, S_IN1);
  sp.o32[0].s32 = st.out(sums, g, S_OUT0);
  sp.counts = st.out(counts, g, S_OUT1);
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}

NF_EXPORT int32_t nf_pack_sum_count_f64(const int32_t* k1, const int32_t* k2, int32_t m2,
                                        const double* vals, size_t n, double* sums,
                                        int64_t* counts, size_t g) {
  if (!k1 || !k2 || !vals || !sums || !counts) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  DenseSpec sp;
  sp.keys = st.in(k1, n, S_IN0);
  sp.k2 = st.in(k2, n, S_IN2);
  sp.m2 = m2;
  sp.n = n;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  sp.nf64 = 1;
  sp.vf[0] = st.in(vals, n, S_IN1);
  sp.of[0].f64 = st.out(sums, g, S_OUT0);
  sp.counts = st.out(counts, g, S_OUT1);
  AggResult r = dense_aggregate(sp, st.s);
  st.finish();
  return r.rc;
  NFH_GUARD_END(int32_t)
}
