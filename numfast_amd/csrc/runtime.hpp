// runtime.hpp -- host runtime shared by every entry point of libnumfast_hip.so.
//
// The C ABI (numfast-native/src/lib.rs) carries plain pointers and sizes.
// NumFast's ctypes caller passes NumPy host buffers; a device-resident
// caller (torch tensors, other HIP code) passes hipMalloc'd pointers.  Each
// entry point classifies every pointer once: device pointers are used in
// place, host pointers are staged through a per-thread device pool and the
// outputs copied back before the call returns.  Every call is synchronous
// like the reference (the result is in the caller's buffers on return).
//
// Threading: the reference calls its symbols concurrently from a
// ThreadPoolExecutor (src/Drivers/CPU/_lib/cpu.py:558-581).  All mutable
// state (stream, pool, look-back epoch) is thread_local, so concurrent
// calls never share scratch.  GPU initialisation is lazy (first call).
//
// Errors never abort: any HIP failure becomes NF_DEVICE_ERROR (-6), an
// additive code outside the reference table (core/errors.rs:13-38), which
// the reference's wrappers treat like any other non-zero rc (fallback).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#define NF_EXPORT extern "C" __attribute__((visibility("default")))

namespace nfh {

// Return codes: numfast-native/src/core/errors.rs:13-38 (+ NF_DEVICE_ERROR).
constexpr int32_t OK = 0;
constexpr int32_t NULL_OR_ABORT = -1;
constexpr int32_t BAD_RANGE = -2;
constexpr int32_t MALFORMED = -3;
constexpr int32_t BAD_VARIANT = -3;
constexpr int32_t OVERFLOW_RC = -4;
constexpr int32_t UNSORTED = -5;
constexpr int32_t DEVICE_ERROR = -6;
constexpr int32_t TOO_SMALL = -7;  // additive: caller capacity too small

struct DeviceError {
  hipError_t err;
  const char* what;
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError{e, what};
}
#define NFH_CHECK(call) ::nfh::check((call), #call)
#define NFH_LAUNCH_CHECK() ::nfh::check(hipGetLastError(), "kernel launch")

// Per-thread device context: stream, scratch pool, look-back epoch.
// Slot ids.  Slots 0..15 hold staged user buffers, 16.. hold internal
// scratch, so staging and scratch never alias within one call.
enum Slot : int {
  S_IN0 = 0, S_IN1, S_IN2, S_IN3, S_IN4, S_IN5,
  S_OUT0 = 6, S_OUT1, S_OUT2, S_OUT3, S_OUT4, S_OUT5, S_OUT6,
  S_T0 = 16, S_T1, S_T2, S_T3, S_T4, S_T5,
  S_LOOKBACK = 24, S_META, S_META2, S_WORK0, S_WORK1, S_WORK2, S_WORK3, S_MISC,
  // learned-CDF sort: code-ordered (word, row) pairs, model / knots, buckets,
  // and the overflow segment sort
  S_L0 = 32, S_L1, S_L2, S_L3, S_LMODEL, S_LBUCKET, S_LSEG0, S_LSEG1, S_LSEG2, S_LSEG3,
  S_LSEG4, S_LTRAIN,
  S_COUNT
};

struct Ctx {
  hipStream_t stream = nullptr;
  bool stream_set_by_user = false;
  void* slot[S_COUNT] = {};
  size_t cap[S_COUNT] = {};
  void* pinned = nullptr;  // small pinned buffer for scalar readbacks
  size_t pinned_cap = 0;
  uint32_t epoch = 0;  // 30-bit look-back epoch (see sort.hip)
  ~Ctx();
};

Ctx& ctx();
void ensure_device();  // throws DeviceError when no GPU is usable
hipStream_t stream();
uint32_t next_epoch();  // fresh look-back epoch, never 0

// Scratch slot `id` grown to at least `bytes` (contents undefined).
void* scratch(int id, size_t bytes);
template <class T>
T* scratch_t(int id, size_t count) {
  return static_cast<T*>(scratch(id, count * sizeof(T) + 256));
}
// Small pinned host buffer (for scalar results read back per call).
void* pinned(size_t bytes);

bool is_device_ptr(const void* p);


// Staging of one call's buffers.  in() returns a device pointer holding
// the caller's input; out() returns a device pointer whose contents are
// copied back to the caller's host buffer by finish().  Device pointers
// pass through untouched.
class Stage {
 public:
  Stage();
  template <class T>
  const T* in(const T* p, size_t count, int slot) {
    return static_cast<const T*>(in_raw(p, count * sizeof(T), slot));
  }
  template <class T>
  T* out(T* p, size_t count, int slot, bool copy_in = false) {
    return static_cast<T*>(out_raw(p, count * sizeof(T), slot, copy_in));
  }
  // Device scratch: the caller's buffer when it is device memory, else pool.
  template <class T>
  T* tmp(T* p, size_t count, int slot) {
    if (p != nullptr && is_device_ptr(p)) return p;
    return scratch_t<T>(slot, count);
  }
  // Copy `bytes` of staged output back (only the first `bytes` of a buffer
  // may be meaningful, e.g. the first ng lanes).
  void set_out_bytes(const void* host_p, size_t bytes);
  void finish();  // D2H of staged outputs + stream sync
  hipStream_t s;

 private:
  const void* in_raw(const void* p, size_t bytes, int slot);
  void* out_raw(void* p, size_t bytes, int slot, bool copy_in);
  struct Pending {
    void* host;
    void* dev;
    size_t bytes;
  };
  std::vector<Pending> pend_;
};

// Device-wide reduction of a scalar into pinned memory then read: helper
// used by entry points that return counts (ng).
template <class T>
T read_scalar(const T* dev_p, hipStream_t s) {
  T* h = static_cast<T*>(pinned(sizeof(T)));
  NFH_CHECK(hipMemcpyAsync(h, dev_p, sizeof(T), hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  return *h;
}

inline unsigned grid_for(size_t n, unsigned block, unsigned max_blocks = 1u << 20) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return static_cast<unsigned>(g);
}

// Fill kernel launcher (graph-safe replacement for hipMemsetAsync).
void fill_u32(void* p, uint32_t v, size_t count, hipStream_t s);
void fill_u64(void* p, uint64_t v, size_t count, hipStream_t s);

int num_cus();

// True when LDS atomics return old values in ascending lane order for
// lanes of one instruction that hit one address (checked once per process
// on the current device; the stable rank_lds ranking depends on it).
bool lds_rank_supported();

// Optional per-kernel timing with HIP events on the launch stream
// (nfh_profile_enable).  A scope brackets one kernel launch; durations are
// collected after the stream is synchronised (Stage::finish / prof_collect).
class ProfScope {
 public:
  // `bytes`: algorithmic bytes the launch must move (for roofline reports).
  ProfScope(const char* name, hipStream_t s, double bytes = 0.0);
  ~ProfScope();

 private:
  const char* name_;
  hipStream_t s_;
  double bytes_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};
void prof_collect();
// Adds `bytes` to the algorithmic bytes of this thread's latest pending
// ProfScope named `name` (a launch whose byte count was known only after it).
void prof_add_bytes(const char* name, double bytes);
// While one lives on this thread, ProfScope creates no events (entry points
// that must not allocate run under it).
class ProfSuspend {
 public:
  ProfSuspend();
  ~ProfSuspend();

 private:
  bool prev_;
};

}  // namespace nfh

// Wrap an entry point body: every exception becomes a return code.
#define NFH_GUARD_BEGIN try {
#define NFH_GUARD_END(rc_type)                                   \
  }                                                              \
  catch (const ::nfh::DeviceError&) {                            \
    return static_cast<rc_type>(::nfh::DEVICE_ERROR);            \
  }                                                              \
  catch (...) {                                                  \
    return static_cast<rc_type>(::nfh::DEVICE_ERROR);            \
  }
