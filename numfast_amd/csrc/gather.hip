// gather.hip -- positional take (the IR gather, cpu.py:1110-1138 _gather_ref
// and :3170-3176): out[i] = values[idx[i]] in index order, with the
// validity bytes taken the same way; any index outside [0, n) refuses the
// whole call, as the reference raises before producing output.
//
// One pass: every thread takes GA_IPT indices, loads them all, then issues
// the (random) value loads together, so each wave keeps 4 x 64 reads in
// flight.  An index outside [0, n) is never dereferenced: the thread
// records the first such position (atomic min) and writes nothing for it;
// the host then returns -2.
#include "device.hpp"
#include "runtime.hpp"

namespace nfh {

constexpr int GA_IPT = 4;

template <class T, class I>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ values, uint64_t n,
                                                const I* __restrict__ idx, uint32_t k,
                                                T* __restrict__ out,
                                                const uint8_t* __restrict__ vin,
                                                uint8_t* __restrict__ vout,
                                                unsigned long long* __restrict__ bad) {
  const uint32_t stride = gridDim.x * blockDim.x * GA_IPT;
  for (uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x); b < k; b += stride) {
    I ix[GA_IPT];
#pragma unroll
    for (int j = 0; j < GA_IPT; ++j) {
      const uint32_t i = b + uint32_t(j) * gridDim.x * blockDim.x;
      ix[j] = i < k ? idx[i] : I(0);
    }
    bool ok[GA_IPT];
    T v[GA_IPT];
    uint8_t m[GA_IPT];
#pragma unroll
    for (int j = 0; j < GA_IPT; ++j) {
      const uint32_t i = b + uint32_t(j) * gridDim.x * blockDim.x;
      ok[j] = i < k && ix[j] >= I(0) && uint64_t(ix[j]) < n;
      if (i < k && !ok[j]) atomicMin(bad, (unsigned long long)i);
      const uint64_t s = ok[j] ? uint64_t(ix[j]) : 0u;
      v[j] = values[s];
      m[j] = vin ? vin[s] : uint8_t(0);
    }
#pragma unroll
    for (int j = 0; j < GA_IPT; ++j) {
      const uint32_t i = b + uint32_t(j) * gridDim.x * blockDim.x;
      if (!ok[j]) continue;
      out[i] = v[j];
      if (vout) vout[i] = m[j];
    }
  }
}

template <class T>
static void launch_gather(const void* values, size_t n, const void* idx, bool idx64, size_t k,
                          void* out, const uint8_t* vin, uint8_t* vout, unsigned long long* bad,
                          hipStream_t s) {
  const unsigned grid = grid_for((k + GA_IPT - 1) / GA_IPT, 256, 16 * num_cus());
  ProfScope prof("gather", s, double(k) * (2.0 * sizeof(T) + (idx64 ? 8.0 : 4.0)));
  if (idx64)
    hipLaunchKernelGGL((k_gather<T, int64_t>), dim3(grid), dim3(256), 0, s,
                       static_cast<const T*>(values), uint64_t(n),
                       static_cast<const int64_t*>(idx), uint32_t(k), static_cast<T*>(out), vin,
                       vout, bad);
  else
    hipLaunchKernelGGL((k_gather<T, int32_t>), dim3(grid), dim3(256), 0, s,
                       static_cast<const T*>(values), uint64_t(n),
                       static_cast<const int32_t*>(idx), uint32_t(k), static_cast<T*>(out), vin,
                       vout, bad);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh

using namespace nfh;

// Positional take of k elements of elem_size bytes (1, 4 or 8) from
// values[0, n) at int32 (idx_bytes 4) or int64 (idx_bytes 8) positions.
// valid_in / valid_out (both or neither) carry the validity bytes.  Returns
// 0, -1 null pointer, -2 an index outside [0, n) (outputs unspecified), -3
// bad element or index size or k > i32::MAX, -6 device error.
NF_EXPORT int32_t nfh_gather(const void* values, size_t n, size_t elem_size, const void* idx,
                             size_t idx_bytes, size_t k, void* out, const uint8_t* valid_in,
                             uint8_t* valid_out) {
  if (k == 0) return OK;
  if (!values || !idx || !out || (valid_in == nullptr) != (valid_out == nullptr))
    return NULL_OR_ABORT;
  if ((elem_size != 1 && elem_size != 4 && elem_size != 8) || (idx_bytes != 4 && idx_bytes != 8) ||
      k > size_t(INT32_MAX))
    return MALFORMED;
  if (n == 0) return BAD_RANGE;
  NFH_GUARD_BEGIN
  Stage st;
  const void* v = st.in(static_cast<const uint8_t*>(values), n * elem_size, S_IN0);
  const void* ix = st.in(static_cast<const uint8_t*>(idx), k * idx_bytes, S_IN1);
  const uint8_t* vin = valid_in ? st.in(valid_in, n, S_IN2) : nullptr;
  void* o = st.out(static_cast<uint8_t*>(out), k * elem_size, S_OUT0);
  uint8_t* vout = valid_out ? st.out(valid_out, k, S_OUT1) : nullptr;
  unsigned long long* bad = scratch_t<unsigned long long>(S_MISC, 1);
  fill_u64(bad, ~0ull, 1, st.s);
  const bool i64 = idx_bytes == 8;
  if (elem_size == 1)
    launch_gather<uint8_t>(v, n, ix, i64, k, o, vin, vout, bad, st.s);
  else if (elem_size == 4)
    launch_gather<uint32_t>(v, n, ix, i64, k, o, vin, vout, bad, st.s);
  else
    launch_gather<unsigned long long>(v, n, ix, i64, k, o, vin, vout, bad, st.s);
  const unsigned long long first_bad = read_scalar(bad, st.s);
  st.finish();
  return first_bad == ~0ull ? OK : BAD_RANGE;
  NFH_GUARD_END(int32_t)
}
