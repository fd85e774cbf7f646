// noalloc.hip -- a stable radix argsort that allocates nothing.
//
// nfh_sort_perm_i32_noalloc is nf_sort_perm_i32 for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc.  The caller passes device
// buffers only (the keys, the permutation and the three n-element work
// buffers of the reference ABI, numfast-native/src/lib.rs:1326-1357); the
// sort's own scratch -- per-chain digit counts and bins, tile counters and
// the decoupled look-back words -- lives in `scratch`, a fixed device
// array of this file sized for NA_MAX_ROWS keys.
//
// Differences from nf_sort_perm_i32, all of them kept small on purpose:
//   * host buffers are refused (-3): staging them would allocate;
//   * n > NA_MAX_ROWS is refused (-3);
//   * every digit pass runs: the OR / AND readback that lets the plain
//     entry point skip passes whose digit all keys share needs a pinned
//     host buffer and a host synchronisation.  An identity pass leaves the
//     result unchanged, so the permutation is the same;
//   * the work goes to the stream set on this thread by nfh_set_stream,
//     else the null stream (creating the library's own stream would
//     allocate), and the call returns after that stream has finished it.
// Each GPU holds its own copy of `scratch`; the first call on a device
// takes its address with hipGetSymbolAddress (with that device current)
// and keeps it.  Every call on one GPU shares that copy: the look-back
// words and counters are cleared at the start of each call, and calls that
// overlap on one GPU (two host threads, two streams) race -- callers
// serialise them.
#include <mutex>

#include "device.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {
constexpr size_t NA_MAX_ROWS = size_t(1) << 24;
constexpr size_t NA_LB_WORDS =
    (NA_MAX_ROWS + SORT_TILE_I32 - 1) / SORT_TILE_I32 * 256;         // 2 MiB
constexpr size_t NA_META_U64 = (sort_meta_words<int32_t>() + 1) / 2;  // 64 KiB
constexpr size_t NA_WORDS = NA_LB_WORDS + NA_META_U64;
}  // namespace nfh

// The scratch of the allocation-free sort (about 2.1 MiB per GPU).  In host
// code the name is a stand-in, not a device address: only the pointer that
// hipGetSymbolAddress returns is passed to kernels.
static __device__ __align__(256) uint64_t scratch[nfh::NA_WORDS];

namespace nfh {
namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
uint64_t* g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of this device's `scratch`, taken once per device.
uint64_t* na_scratch() {
  int dev = 0;
  NFH_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) throw DeviceError{hipErrorInvalidDevice, "device id"};
  std::lock_guard<std::mutex> g(g_na_mu);
  if (g_na_ptr[dev] == nullptr) {
    void* p = nullptr;
    NFH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    g_na_ptr[dev] = static_cast<uint64_t*>(p);
  }
  return g_na_ptr[dev];
}
}  // namespace

}  // namespace nfh

using namespace nfh;

// No counterpart in the reference: nf_sort_perm_i32 (lib.rs:1326-1357)
// for callers that cannot allocate.  Device buffers only.  Returns 0, -1
// null pointer, -3 host buffer or n > 2^24, -6 device error.
NF_EXPORT int32_t nfh_sort_perm_i32_noalloc(const int32_t* keys, size_t n, uint8_t descending,
                                            int32_t* perm_out, uint32_t* tmp0, uint32_t* tmp1,
                                            int32_t* tmp_p) {
  if (n == 0) return OK;
  if (!keys || !perm_out || !tmp0 || !tmp1 || !tmp_p) return NULL_OR_ABORT;
  if (n > NA_MAX_ROWS) return MALFORMED;
  NFH_GUARD_BEGIN
  ensure_device();
  if (!is_device_ptr(keys) || !is_device_ptr(perm_out) || !is_device_ptr(tmp0) ||
      !is_device_ptr(tmp1) || !is_device_ptr(tmp_p))
    return MALFORMED;
  ProfSuspend no_events;
  Ctx& c = ctx();
  hipStream_t s = c.stream_set_by_user ? c.stream : nullptr;
  uint64_t* base = na_scratch();
  SortWs ws;
  ws.lb = base;
  ws.meta = reinterpret_cast<uint32_t*>(base + NA_LB_WORDS);
  ws.epoch0 = 1;  // the words below are cleared to epoch 0 first
  ws.all_passes = true;
  fill_u64(ws.lb, 0, sort_lookback_words<int32_t>(n), s);
  radix_argsort_ws<int32_t>(keys, n, descending != 0, perm_out, tmp0, tmp1, tmp_p, nullptr, ws,
                            s);
  NFH_CHECK(hipStreamSynchronize(s));
  return OK;
  NFH_GUARD_END(int32_t)
}
