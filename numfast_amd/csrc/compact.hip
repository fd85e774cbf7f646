// compact.hip -- the remaining eager symbols of the reference loader:
// carry compaction, sorted-run aggregation, composite pack and the
// prefix+int string encode (src/Drivers/CPU/_lib/native_cpu.py:125-140
// binds all eight; one missing symbol disables the whole library).
//
//   nf_carry_build_i64/_f64   result/carry.rs:23-67, core/compact.rs:16-37
//   nf_sorted_run_i64/_f64    groupby/sorted.rs:24-58, lib.rs:399-447
//   nf_pack_i32_direct        transform/pack.rs:15-32, lib.rs:252-272
//   nf_pattern_encode         transform/pattern.rs:50-131, lib.rs:340-373
#include "device.hpp"
#include "exactsum.hpp"
#include "groupby.hpp"
#include "runtime.hpp"
#include "scan.hpp"

namespace nfh {

typedef unsigned long long u64;

// ---------------------------------------------------------------- carry
template <class S>
struct CarryOp {
  const int64_t* counts_m;
  const S* sums_m;
  int64_t* ukeys;
  int64_t* counts;
  S* sums;
  int64_t* ng;
  __device__ uint32_t value(uint32_t j) const { return counts_m[j] != 0 ? 1u : 0u; }
  __device__ void emit(uint32_t j, uint32_t ex) const {
    const int64_t c = counts_m[j];
    if (c != 0) {
      ukeys[ex] = int64_t(j);
      counts[ex] = c;
      sums[ex] = sums_m[j];
    }
  }
  __device__ void total(uint32_t t) const { *ng = int64_t(t); }
};

template <class S>
static int64_t carry_entry(const int64_t* counts_m, const S* sums_m, size_t m, int64_t* ukeys,
                           int64_t* counts, S* sums) {
  Stage st;
  const int64_t* cm = st.in(counts_m, m, S_IN0);
  const S* sm = st.in(sums_m, m, S_IN1);
  int64_t* uk = st.out(ukeys, m, S_OUT0);
  int64_t* ct = st.out(counts, m, S_OUT1);
  S* su = st.out(sums, m, S_OUT2);
  int64_t* ng_d = scratch_t<int64_t>(S_MISC, 1);
  scan_tiles(CarryOp<S>{cm, sm, uk, ct, su, ng_d}, m, st.s);
  const int64_t ng = read_scalar(ng_d, st.s);
  st.set_out_bytes(ukeys, size_t(ng) * 8);
  st.set_out_bytes(counts, size_t(ng) * 8);
  st.set_out_bytes(sums, size_t(ng) * sizeof(S));
  st.finish();
  return ng;
}

// ---------------------------------------------------------------- sorted runs
struct RunHeadOp {
  const int32_t* keys;
  int32_t* rid;
  int64_t* ukeys;
  int64_t* ng;
  __device__ uint32_t value(uint32_t i) const {
    return (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
  }
  __device__ void emit(uint32_t i, uint32_t ex) const {
    const bool head = i == 0 || keys[i] != keys[i - 1];
    if (head) {
      rid[i] = int32_t(ex);
      ukeys[ex] = int64_t(keys[i]);
    } else {
      rid[i] = int32_t(ex - 1);
    }
  }
  __device__ void total(uint32_t t) const { *ng = int64_t(t); }
};

// groupby/sorted.rs:24-58: ng >= 0, -1 on a regressing key.  int64 sums
// wrap like the reference's release-mode `acc += v`; float64 sums use the
// order-independent exact accumulation.
template <class V>
static int64_t sorted_run_entry(const int32_t* keys, const V* vals, size_t n, int64_t* ukeys,
                                V* sums, int64_t* counts) {
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  const V* v = st.in(vals, n, S_IN1);
  int64_t* uk = st.out(ukeys, n, S_OUT0);
  V* su = st.out(sums, n, S_OUT1);
  int64_t* ct = st.out(counts, n, S_OUT2);
  if (first_regress_row(k, n, st.s) < n) {
    st.set_out_bytes(ukeys, 0);
    st.set_out_bytes(sums, 0);
    st.set_out_bytes(counts, 0);
    st.finish();
    return -1;
  }
  // S_IN5: not used by this entry's staging nor by dense_aggregate, whose
  // row-order path reuses the S_T* slots.
  int32_t* rid = scratch_t<int32_t>(S_IN5, n);
  int64_t* ng_d = scratch_t<int64_t>(S_MISC, 1);
  scan_tiles(RunHeadOp{k, rid, uk, ng_d}, n, st.s);
  const int64_t ng = read_scalar(ng_d, st.s);
  DenseSpec sp;
  sp.keys = rid;
  sp.n = n;
  sp.g = size_t(ng);
  sp.lo = 0;
  sp.hi = size_t(ng);
  sp.counts = ct;
  sp.presorted = true;
  sp.f64_seq_first = true;
  if constexpr (std::is_same<V, int64_t>::value) {
    sp.ni64 = 1;
    sp.v64[0] = v;
    sp.o64[0].s64 = su;
  } else {
    sp.nf64 = 1;
    sp.vf[0] = v;
    sp.of[0].f64 = su;
  }
  (void)dense_aggregate(sp, st.s);  // run ids are in range by construction
  st.set_out_bytes(ukeys, size_t(ng) * 8);
  st.set_out_bytes(sums, size_t(ng) * sizeof(V));
  st.set_out_bytes(counts, size_t(ng) * 8);
  st.finish();
  return ng;
}

// ---------------------------------------------------------------- pack
__global__ void k_pack(const int32_t* __restrict__ k1, const int32_t* __restrict__ k2,
                       long long m2, uint32_t n, int32_t* __restrict__ out, u64* err) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t start = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t base = start - lane_id(); base < n; base += stride) {
    const uint32_t i = base + lane_id();
    bool bad = false;
    if (i < n) {
      const int32_t a = k1[i], b = k2[i];
      const long long v = (long long)a * m2 + (long long)b;
      bad = a < 0 || b < 0 || v > 2147483647ll;
      if (!bad) out[i] = (int32_t)(uint32_t)(unsigned long long)v;
    }
    const u64 m = __ballot(bad);
    if (m != 0 && lane_id() == __ffsll((unsigned long long)m) - 1) atomicMin(err, (u64)i);
  }
}

// ---------------------------------------------------------------- pattern
struct PatRow {
  int32_t code;
  int32_t width;  // digits, -1 when the row is not valid
  int32_t status;  // 0 ok, 2 overflow, 3 malformed end
};

// Parse row i (transform/pattern.rs:59-131).  Row status: malformed when
// its end offset regresses or leaves the data; overflow when the magnitude
// exceeds the int32 limit for its sign.
__global__ void k_pattern_rows(const uint8_t* __restrict__ data, uint32_t total,
                               const int32_t* __restrict__ offs, uint32_t n,
                               const uint8_t* __restrict__ prefix, uint32_t plen,
                               PatRow* __restrict__ rows, u64* __restrict__ first_err) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    PatRow r{0, -1, 0};
    const int32_t prev = offs[i];
    const int32_t end = offs[i + 1];
    // offs[0] was checked by the host; a regressing or out-of-range end is
    // malformed (the previous end is checked by its own row).
    if (end < prev || end < 0 || uint32_t(end) > total) {
      r.status = 3;
    } else if (prev >= 0 && uint32_t(prev) <= total) {
      const uint32_t len = uint32_t(end - prev);
      const uint8_t* row = data + prev;
      bool match = len >= plen;
      for (uint32_t j = 0; match && j < plen; ++j) match = row[j] == prefix[j];
      if (match) {
        uint32_t p = plen;
        const bool neg = p < len && row[p] == '-';
        if (neg) ++p;
        if (p < len) {
          long long mag = 0;
          int w = 0;
          bool ok = true;
          for (; p < len; ++p) {
            const uint8_t ch = row[p];
            if (ch < '0' || ch > '9') {
              ok = false;
              break;
            }
            const long long d = ch - '0';
            mag = (mag > (0x7FFFFFFFFFFFFFFFll - d) / 10) ? 0x7FFFFFFFFFFFFFFFll : mag * 10 + d;
            ++w;
          }
          if (ok) {
            const long long lim = neg ? (1ll << 31) : 2147483647ll;
            if (mag > lim) {
              r.status = 2;
            } else {
              r.code = neg ? (int32_t)(-mag) : (int32_t)mag;
              r.width = w;
            }
          }
        }
      }
    }
    rows[i] = r;
    if (r.status != 0) atomicMin(first_err, (u64)i);
  }
}

// Write rows before the first error (the failing row itself is cleared).
__global__ void k_pattern_write(const PatRow* __restrict__ rows, uint32_t n, uint32_t upto,
                                int32_t* __restrict__ codes, uint8_t* __restrict__ valid,
                                int* __restrict__ width) {
  const uint32_t stride = gridDim.x * blockDim.x;
  int wmax = -1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i < upto) {
      const PatRow r = rows[i];
      codes[i] = r.width >= 0 ? r.code : 0;
      valid[i] = r.width >= 0 ? 1 : 0;
      if (r.width > wmax) wmax = r.width;
    } else if (i == upto) {
      codes[i] = 0;
      valid[i] = 0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_down(wmax, off);
    wmax = o > wmax ? o : wmax;
  }
  if (lane_id() == 0 && wmax >= 0) atomicMax(width, wmax);
}

}  // namespace nfh

using namespace nfh;

// lib.rs:462-489
NF_EXPORT int64_t nf_carry_build_i64(const int64_t* counts_m, const int64_t* sums_m, size_t m,
                                     int64_t* ukeys, int64_t* counts, int64_t* sums) {
  if (m == 0) return 0;
  if (!counts_m || !sums_m || !ukeys || !counts || !sums) return -2;
  if (m > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return carry_entry<int64_t>(counts_m, sums_m, m, ukeys, counts, sums);
  NFH_GUARD_END(int64_t)
}

// lib.rs:492-518
NF_EXPORT int64_t nf_carry_build_f64(const int64_t* counts_m, const double* sums_m, size_t m,
                                     int64_t* ukeys, int64_t* counts, double* sums) {
  if (m == 0) return 0;
  if (!counts_m || !sums_m || !ukeys || !counts || !sums) return -2;
  if (m > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return carry_entry<double>(counts_m, sums_m, m, ukeys, counts, sums);
  NFH_GUARD_END(int64_t)
}

// lib.rs:399-425
NF_EXPORT int64_t nf_sorted_run_i64(const int32_t* keys, const int64_t* vals, size_t n,
                                    int64_t* ukeys, int64_t* sums, int64_t* counts) {
  if (n == 0) return 0;
  if (!keys || !vals || !ukeys || !sums || !counts) return -2;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return sorted_run_entry<int64_t>(keys, vals, n, ukeys, sums, counts);
  NFH_GUARD_END(int64_t)
}

// lib.rs:428-447
NF_EXPORT int64_t nf_sorted_run_f64(const int32_t* keys, const double* vals, size_t n,
                                    int64_t* ukeys, double* sums, int64_t* counts) {
  if (n == 0) return 0;
  if (!keys || !vals || !ukeys || !sums || !counts) return -2;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  return sorted_run_entry<double>(keys, vals, n, ukeys, sums, counts);
  NFH_GUARD_END(int64_t)
}

// lib.rs:252-272 -- out[i] = k1[i]*m2 + k2[i]; -2 on a negative operand or
// an int32 overflow.  Rows are written independently (the reference stops at
// the first failing row; the output is only meaningful when rc == 0).
NF_EXPORT int32_t nf_pack_i32_direct(const int32_t* k1, const int32_t* k2, int32_t m2, size_t n,
                                     int32_t* out) {
  if (!k1 || !k2 || !out) return NULL_OR_ABORT;
  if (n == 0) return OK;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* a = st.in(k1, n, S_IN0);
  const int32_t* b = st.in(k2, n, S_IN1);
  int32_t* o = st.out(out, n, S_OUT0, true);
  u64* err = scratch_t<u64>(S_MISC, 1);
  fill_u64(err, ~0ull, 1, st.s);
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256, 16 * num_cus())), dim3(256), 0, st.s, a, b,
                     (long long)m2, uint32_t(n), o, err);
  NFH_LAUNCH_CHECK();
  const u64 e = read_scalar(err, st.s);
  st.finish();
  return e == ~0ull ? OK : BAD_RANGE;
  NFH_GUARD_END(int32_t)
}

// lib.rs:340-373
NF_EXPORT int32_t nf_pattern_encode(const uint8_t* data, size_t total, const int32_t* offs,
                                    size_t n, const uint8_t* prefix, size_t prefix_len,
                                    int32_t* codes, uint8_t* valid, int32_t* width_out,
                                    int32_t* err_row_out) {
  if (!data || !offs || !prefix || !codes || !valid || !width_out || !err_row_out)
    return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX) || total > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  // Scalars and offs[0] are read on the host: the scalar outputs and the
  // first offset are a handful of bytes.
  int32_t off0 = 0;
  const bool offs_dev = is_device_ptr(offs);
  if (offs_dev) {
    NFH_CHECK(hipMemcpy(&off0, offs, 4, hipMemcpyDeviceToHost));
  } else {
    off0 = offs[0];
  }
  const bool wdev = is_device_ptr(width_out), edev = is_device_ptr(err_row_out);
  auto put = [&](int32_t* p, bool dev, int32_t v) {
    if (dev)
      NFH_CHECK(hipMemcpy(p, &v, 4, hipMemcpyHostToDevice));
    else
      *p = v;
  };
  put(width_out, wdev, -1);
  put(err_row_out, edev, -1);
  if (off0 < 0 || size_t(off0) > total) return MALFORMED;
  if (n == 0) return OK;
  Stage st;
  const uint8_t* d = st.in(data, total, S_IN0);
  const int32_t* of = st.in(offs, n + 1, S_IN1);
  const uint8_t* pf = st.in(prefix, prefix_len, S_IN2);
  int32_t* co = st.out(codes, n, S_OUT0, true);
  uint8_t* va = st.out(valid, n, S_OUT1, true);
  PatRow* rows = scratch_t<PatRow>(S_T0, n);
  u64* meta = scratch_t<u64>(S_MISC, 2);
  fill_u64(meta, ~0ull, 1, st.s);
  fill_u64(meta + 1, ~0ull, 1, st.s);  // width (int, -1)
  hipLaunchKernelGGL(k_pattern_rows, dim3(grid_for(n, 256, 16 * num_cus())), dim3(256), 0, st.s,
                     d, uint32_t(total), of, uint32_t(n), pf, uint32_t(prefix_len), rows, meta);
  NFH_LAUNCH_CHECK();
  const u64 fe = read_scalar(meta, st.s);
  const uint32_t upto = fe == ~0ull ? uint32_t(n) : uint32_t(fe);
  int* wdev_p = reinterpret_cast<int*>(meta + 1);
  hipLaunchKernelGGL(k_pattern_write, dim3(grid_for(n, 256, 16 * num_cus())), dim3(256), 0,
                     st.s, rows, uint32_t(n), upto, co, va, wdev_p);
  NFH_LAUNCH_CHECK();
  int32_t status = 0;
  if (fe != ~0ull) {
    PatRow pr;
    NFH_CHECK(hipMemcpyAsync(&pr, rows + fe, sizeof(PatRow), hipMemcpyDeviceToHost, st.s));
    NFH_CHECK(hipStreamSynchronize(st.s));
    status = pr.status;
  }
  const int w = read_scalar(wdev_p, st.s);
  st.finish();
  put(width_out, wdev, w);
  if (status == 2) {
    put(err_row_out, edev, int32_t(fe));
    return BAD_RANGE;
  }
  if (status == 3) return MALFORMED;
  return OK;
  NFH_GUARD_END(int32_t)
}
