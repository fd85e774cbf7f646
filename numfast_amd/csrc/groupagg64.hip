// groupagg64.hip -- NULL-aware group-by over raw int32 keys with a float64
// or int64 value column: sum / count / min / max / mean, keys ascending.
//
// This is the executor's min/max branch for 64-bit value columns
// (src/Drivers/CPU/_lib/cpu.py:3507-3563): rows whose key or value is NULL
// are dropped first (_valid_mask, cpu.py:1374-1382), groups come from the
// surviving keys in ascending order (groupindex.py:3-8), min / max follow
// np.minimum.at / np.maximum.at over a +-inf (float) or iinfo (int) fill
// (_group_minmax, cpu.py:2570-2601), sums follow np.bincount (float64, row
// order; int64 exact under the executor's 2^53 gate, cpu.py:424-446) and
// mean = sums.astype(float64) / counts (carry.py:55-62).
//
// Pipeline:
//   1. stable compaction of the valid rows (keys + 8-byte values);
//   2. codes: key - kmin when the key span is small against the rows, else
//      the radix unique's sorted-position codes;
//   3. counts and sums by the dense-code engine (groupby.hip): float64
//      sums exactly rounded, and redone in row order for the groups where
//      the reference's own rounding could matter (within 1e-12 relative of
//      the row-order sum, bit-equal where redone); int64 sums wrap;
//   4. min / max by one pass of order-preserving 64-bit atomics per row;
//   5. one compaction of the present groups writes the outputs.
//
// np.minimum / np.maximum return their second argument on ties and their
// NaN argument when there is one, so over a row-order fold the result is
// the first NaN of the group (its bits kept), and a zero extreme takes the
// sign of the group's last zero (-0.0 == +0.0).  Step 4 tracks both: the
// first NaN row (atomic min) and the last zero row with its sign bit
// (atomic max of row << 1 | sign); zeros enter the min / max as +0.0.
#include <algorithm>

#include "device.hpp"
#include "groupagg.hpp"
#include "groupby.hpp"
#include "runtime.hpp"
#include "scan.hpp"
#include "unique.hpp"

namespace nfh {

typedef unsigned long long u64;

constexpr u64 ENC_ZERO = 0x8000000000000000ull;  // +0.0 in the float order
constexpr uint32_t NO_ROW = 0xFFFFFFFFu;

// Compact the rows whose key and value are valid, in row order.
struct ValidRows64Op {
  const int32_t* keys;
  const u64* vals;
  const uint8_t* v;
  const uint8_t* kv;
  int32_t* ck;
  u64* cv;
  uint32_t* m;
  __device__ bool ok(uint32_t i) const {
    return (v == nullptr || v[i] != 0) && (kv == nullptr || kv[i] != 0);
  }
  __device__ uint32_t value(uint32_t i) const { return ok(i) ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t ex) const {
    if (ok(i)) {
      ck[ex] = keys[i];
      cv[ex] = vals[i];
    }
  }
  __device__ void total(uint32_t t) const { *m = t; }
};

__global__ void k_codes(const int32_t* __restrict__ keys, uint32_t m, int32_t kmin,
                        int32_t* __restrict__ codes) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    codes[i] = int32_t(uint32_t(keys[i]) - uint32_t(kmin));
}

// Per-group min / max state, structure of arrays over g slots.
struct MinMax64 {
  u64* mn;          // order-preserving code, ~0 = none yet
  u64* mx;          // order-preserving code, 0 = none yet
  uint32_t* nan;    // first NaN row, NO_ROW = none
  u64* zero;        // max of (row << 1 | sign) over the zero rows
};

__global__ void k_mm_init(MinMax64 S, uint32_t g) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g; i += stride) {
    S.mn[i] = ~0ull;
    S.mx[i] = 0ull;
    S.nan[i] = NO_ROW;
    S.zero[i] = 0ull;
  }
}

// FLOAT: values are float64 bit patterns; else int64.
template <bool FLOAT>
__global__ __launch_bounds__(256) void k_minmax64(const int32_t* __restrict__ codes,
                                                  const u64* __restrict__ vals, uint32_t m,
                                                  MinMax64 S) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const uint32_t c = uint32_t(codes[i]);
    const u64 b = vals[i];
    u64 e;
    if constexpr (FLOAT) {
      const u64 mag = b & 0x7FFFFFFFFFFFFFFFull;
      if (mag > 0x7FF0000000000000ull) {  // NaN
        if (i < S.nan[c]) atomicMin(&S.nan[c], i);
        continue;
      }
      if (mag == 0) {
        const u64 z = (u64(i) << 1) | (b >> 63);
        if (z > S.zero[c]) atomicMax(&S.zero[c], z);
        e = ENC_ZERO;
      } else {
        e = (b >> 63) ? ~b : (b | ENC_ZERO);
      }
    } else {
      e = b ^ ENC_ZERO;
    }
    // both only move one way: a stale read costs a redundant atomic only
    if (e < S.mn[c]) atomicMin(&S.mn[c], e);
    if (e > S.mx[c]) atomicMax(&S.mx[c], e);
  }
}

struct Agg64Out {
  int64_t* ukeys;
  int64_t* counts;
  u64* sums;   // float64 bits or int64
  u64* mins;
  u64* maxs;
  double* means;
};

template <bool FLOAT>
__device__ __forceinline__ u64 decode_extreme(u64 e, uint32_t nan_row, u64 zero,
                                              const u64* __restrict__ vals) {
  if constexpr (FLOAT) {
    if (nan_row != NO_ROW) return vals[nan_row];
    if (e == ENC_ZERO) return (zero & 1ull) << 63;
    return (e & ENC_ZERO) ? (e & ~ENC_ZERO) : ~e;
  } else {
    return e ^ ENC_ZERO;
  }
}

// Present groups (count != 0) in slot order -> outputs.
template <bool FLOAT>
struct Emit64Op {
  const int64_t* cnt;
  const u64* sum;
  MinMax64 S;
  const u64* vals;
  const int32_t* uniq;  // sparse codes: key of each code (else null)
  int32_t kmin;
  Agg64Out o;
  int64_t* ng;
  __device__ uint32_t value(uint32_t s) const { return cnt[s] != 0 ? 1u : 0u; }
  __device__ void emit(uint32_t s, uint32_t ex) const {
    const int64_t c = cnt[s];
    if (c == 0) return;
    const u64 sm = sum[s];
    if (o.ukeys) o.ukeys[ex] = uniq ? int64_t(uniq[s]) : int64_t(kmin) + int64_t(s);
    if (o.counts) o.counts[ex] = c;
    if (o.sums) o.sums[ex] = sm;
    if (o.mins) o.mins[ex] = decode_extreme<FLOAT>(S.mn[s], S.nan[s], S.zero[s], vals);
    if (o.maxs) o.maxs[ex] = decode_extreme<FLOAT>(S.mx[s], S.nan[s], S.zero[s], vals);
    if (o.means) {
      const double num = FLOAT ? __longlong_as_double((long long)sm) : double((long long)sm);
      o.means[ex] = num / double(c);
    }
  }
  __device__ void total(uint32_t t) const { *ng = int64_t(t); }
};

// Codes over a dense span up to this many slots: max(2^20, 2m), at most 2^27
// (48 B of accumulators per slot); wider spans go through the radix unique.
static bool dense_codes_ok(uint64_t span, size_t m) {
  const uint64_t lim = std::min<uint64_t>(std::max<uint64_t>(uint64_t(1) << 20, 2 * uint64_t(m)),
                                          uint64_t(1) << 27);
  return span <= lim;
}

// Groups of the valid rows into o (lanes for every group: min(n, span)).
// Returns ng.  Synchronous.
template <bool FLOAT>
static int64_t group_agg64_device(const int32_t* keys, const u64* vals, const uint8_t* v,
                                  const uint8_t* kv, size_t n_, const Agg64Out& o,
                                  hipStream_t s) {
  const uint32_t n = uint32_t(n_);
  const int32_t* ck = keys;
  const u64* cv = vals;
  uint32_t m = n;
  if (v != nullptr || kv != nullptr) {
    int32_t* k2 = scratch_t<int32_t>(S_L0, n);
    u64* v2 = scratch_t<u64>(S_L1, n);
    uint32_t* m_d = scratch_t<uint32_t>(S_MISC, 4) + 2;
    scan_tiles(ValidRows64Op{keys, vals, v, kv, k2, v2, m_d}, n, s);
    m = read_scalar(m_d, s);
    ck = k2;
    cv = v2;
  }
  if (m == 0) return 0;
  int32_t kmin, kmax;
  if (!key_range(ck, nullptr, nullptr, m, &kmin, &kmax, s)) return 0;
  const uint64_t span = uint64_t(int64_t(kmax) - int64_t(kmin) + 1);
  int32_t* codes = scratch_t<int32_t>(S_L3, m);
  const int32_t* uniq = nullptr;
  uint64_t g;
  if (dense_codes_ok(span, m)) {
    hipLaunchKernelGGL(k_codes, dim3(grid_for(m, 256, 16 * num_cus())), dim3(256), 0, s, ck, m,
                       kmin, codes);
    NFH_LAUNCH_CHECK();
    g = span;
  } else {
    int32_t* u = scratch_t<int32_t>(S_L2, m);
    int32_t* perm = scratch_t<int32_t>(S_T2, m);
    uint32_t* t0 = scratch_t<uint32_t>(S_T3, m);
    uint32_t* t1 = scratch_t<uint32_t>(S_T4, m);
    int32_t* tp = scratch_t<int32_t>(S_T5, m);
    g = uint64_t(unique_inverse_device<int32_t>(ck, m, u, codes, perm, t0, t1, tp, s));
    uniq = u;
  }
  // per-slot state: counts | sums | min | max | zero (u64 each) | nan (u32)
  unsigned char* st = static_cast<unsigned char*>(scratch(S_LSEG0, size_t(g) * 44 + 1024));
  int64_t* cnt = reinterpret_cast<int64_t*>(st);
  u64* sum = reinterpret_cast<u64*>(st + size_t(g) * 8);
  MinMax64 S;
  S.mn = reinterpret_cast<u64*>(st + size_t(g) * 16);
  S.mx = reinterpret_cast<u64*>(st + size_t(g) * 24);
  S.zero = reinterpret_cast<u64*>(st + size_t(g) * 32);
  S.nan = reinterpret_cast<uint32_t*>(st + size_t(g) * 40);
  DenseSpec sp;
  sp.keys = codes;
  sp.n = m;
  sp.g = g;
  sp.lo = 0;
  sp.hi = g;
  if (FLOAT) {
    sp.nf64 = 1;
    sp.vf[0] = reinterpret_cast<const double*>(cv);
    sp.of[0].f64 = reinterpret_cast<double*>(sum);
  } else {
    sp.ni64 = 1;
    sp.v64[0] = reinterpret_cast<const int64_t*>(cv);
    sp.o64[0].s64 = reinterpret_cast<int64_t*>(sum);
  }
  sp.counts = cnt;
  const AggResult r = dense_aggregate(sp, s);
  if (r.rc != 0) throw DeviceError{hipErrorUnknown, "group codes out of range"};
  hipLaunchKernelGGL(k_mm_init, dim3(grid_for(g, 256, 16 * num_cus())), dim3(256), 0, s, S,
                     uint32_t(g));
  NFH_LAUNCH_CHECK();
  {
    ProfScope prof("group_minmax64", s, double(m) * 12.0);
    hipLaunchKernelGGL(k_minmax64<FLOAT>, dim3(grid_for(m, 256, 16 * num_cus())), dim3(256), 0,
                       s, codes, cv, m, S);
    NFH_LAUNCH_CHECK();
  }
  int64_t* ng_d = scratch_t<int64_t>(S_WORK2, 1);
  scan_tiles(Emit64Op<FLOAT>{cnt, sum, S, cv, uniq, kmin, o, ng_d}, size_t(g), s);
  return read_scalar(ng_d, s);
}

template <bool FLOAT>
static int64_t group_agg64_entry(const int32_t* keys, const void* vals, const uint8_t* valid,
                                 const uint8_t* key_valid, size_t n, int64_t* ukeys,
                                 int64_t* counts, void* sums, void* mins, void* maxs,
                                 double* means, size_t cap) {
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  const u64* v = st.in(static_cast<const u64*>(vals), n, S_IN1);
  const uint8_t* vm = valid ? st.in(valid, n, S_IN2) : nullptr;
  const uint8_t* km = key_valid ? st.in(key_valid, n, S_IN3) : nullptr;
  Agg64Out d{};
  if (cap >= n) {
    d.ukeys = st.out(ukeys, n, S_OUT0);
    d.counts = st.out(counts, n, S_OUT1);
    d.sums = st.out(static_cast<u64*>(sums), n, S_OUT2);
    d.mins = st.out(static_cast<u64*>(mins), n, S_OUT3);
    d.maxs = st.out(static_cast<u64*>(maxs), n, S_OUT4);
    d.means = st.out(means, n, S_OUT5);
    const int64_t ng = group_agg64_device<FLOAT>(k, v, vm, km, n, d, st.s);
    for (const void* p : {static_cast<const void*>(ukeys), static_cast<const void*>(counts),
                          static_cast<const void*>(sums), static_cast<const void*>(mins),
                          static_cast<const void*>(maxs), static_cast<const void*>(means)})
      st.set_out_bytes(p, size_t(ng) * 8);
    st.finish();
    return ng;
  }
  // results in device scratch first, moved once ng <= cap is known
  d.ukeys = ukeys ? scratch_t<int64_t>(S_OUT0, n) : nullptr;
  d.counts = counts ? scratch_t<int64_t>(S_OUT1, n) : nullptr;
  d.sums = sums ? scratch_t<u64>(S_OUT2, n) : nullptr;
  d.mins = mins ? scratch_t<u64>(S_OUT3, n) : nullptr;
  d.maxs = maxs ? scratch_t<u64>(S_OUT4, n) : nullptr;
  d.means = means ? scratch_t<double>(S_OUT5, n) : nullptr;
  const int64_t ng = group_agg64_device<FLOAT>(k, v, vm, km, n, d, st.s);
  if (size_t(ng) > cap) {
    st.finish();
    return TOO_SMALL;
  }
  auto copy = [&](void* dst, const void* src) {
    if (dst && ng > 0)
      NFH_CHECK(hipMemcpyAsync(dst, src, size_t(ng) * 8, hipMemcpyDefault, st.s));
  };
  copy(ukeys, d.ukeys);
  copy(counts, d.counts);
  copy(sums, d.sums);
  copy(mins, d.mins);
  copy(maxs, d.maxs);
  copy(means, d.means);
  st.finish();
  return ng;
}

}  // namespace nfh

using namespace nfh;

// Group-by over raw int32 keys with a float64 value column and NULL masks
// (valid / key_valid may be null): groups of the rows with both set, ukeys
// ascending; counts int64; sums, mins, maxs, means float64.  Outputs hold
// `cap` lanes (null outputs are skipped).  Returns ng, -1 null input, -3
// n > i32::MAX, -6 device error, -7 cap < ng.
NF_EXPORT int64_t nfh_group_agg_f64(const int32_t* keys, const double* vals,
                                    const uint8_t* valid, const uint8_t* key_valid, size_t n,
                                    int64_t* ukeys, int64_t* counts, double* sums, double* mins,
                                    double* maxs, double* means, size_t cap) {
  if (n == 0) return 0;
  if (!keys || !vals) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return group_agg64_entry<true>(keys, vals, valid, key_valid, n, ukeys, counts, sums, mins,
                                 maxs, means, cap);
  NFH_GUARD_END(int64_t)
}

// The same with an int64 value column: sums int64 (wrapping), mins / maxs
// int64, means float64.
NF_EXPORT int64_t nfh_group_agg_i64(const int32_t* keys, const int64_t* vals,
                                    const uint8_t* valid, const uint8_t* key_valid, size_t n,
                                    int64_t* ukeys, int64_t* counts, int64_t* sums, int64_t* mins,
                                    int64_t* maxs, double* means, size_t cap) {
  if (n == 0) return 0;
  if (!keys || !vals) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return group_agg64_entry<false>(keys, vals, valid, key_valid, n, ukeys, counts, sums, mins,
                                  maxs, means, cap);
  NFH_GUARD_END(int64_t)
}
