// runtime.hip -- per-thread device context, scratch pool and staging.
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "device.hpp"
#include "runtime.hpp"

namespace nfh {

// Context memory is intentionally not released at thread exit: thread-local
// destructors of the main thread can run after the HIP runtime has been torn
// down, and a hipFree at that point is unsafe.  Worker threads of a pool live
// for the process, so the pool is bounded by (threads x high-water mark).
Ctx::~Ctx() {}

Ctx& ctx() {
  static thread_local Ctx c;
  return c;
}

namespace {
std::once_flag g_init_once;
hipError_t g_init_err = hipSuccess;
int g_cus = 0;
}  // namespace

void ensure_device() {
  std::call_once(g_init_once, [] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
      g_init_err = e;
      return;
    }
    if (n <= 0) {
      g_init_err = hipErrorNoDevice;
      return;
    }
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) {
      g_init_err = e;
      return;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
      g_init_err = e;
      return;
    }
    g_cus = prop.multiProcessorCount;
  });
  if (g_init_err != hipSuccess) throw DeviceError{g_init_err, "device init"};
}

int num_cus() {
  ensure_device();
  return g_cus > 0 ? g_cus : 256;
}

// Check of the LDS property rank_lds relies on: the old values that
// ds_add_rtn returns to lanes of one instruction hitting one address rise
// with the lane index.  Compares against the ballot-match rank over digit
// patterns with 2..256 distinct values per wave; counts mismatches.
__global__ __launch_bounds__(1024) void k_lds_rank_check(uint32_t trials,
                                                         unsigned long long* bad) {
  __shared__ uint32_t cnt[16][256];
  const int lane = lane_id(), wave = wave_id();
  for (int i = threadIdx.x; i < 16 * 256; i += blockDim.x) (&cnt[0][0])[i] = 0;
  __syncthreads();
  unsigned long long v = 0;
  for (uint32_t t = 0; t < trials; ++t) {
    uint32_t h = (t * 2654435761u) ^ (blockIdx.x * 40503u + uint32_t(wave) * 7919u + lane);
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    const uint32_t mask = (t % 4 == 0) ? 255u : (t % 4 == 1) ? 1u : (t % 4 == 2) ? 3u : 31u;
    const uint32_t d = (h >> 7) & mask;
    const uint32_t old = atomicAdd(&cnt[wave][d], 1u);
    const uint64_t peers = match_digit8(d, true);
    const uint32_t r = uint32_t(__popcll(peers & lanemask_lt()));
    const uint32_t lead_old = __shfl(old, __ffsll((unsigned long long)peers) - 1);
    v += (old != lead_old + r) ? 1u : 0u;
  }
  if (v) atomicAdd(bad, v);
}

// Mismatch counter of the check (a static device word: no allocation).
static __device__ unsigned long long g_rank_bad;

bool lds_rank_supported() {
  static std::once_flag once;
  static bool ok = false;
  std::call_once(once, [] {
    ensure_device();
    unsigned long long* d = nullptr;
    if (hipGetSymbolAddress(reinterpret_cast<void**>(&d), HIP_SYMBOL(g_rank_bad)) != hipSuccess)
      return;
    unsigned long long h = 1;
    // the null stream: no stream is created here
    hipStream_t s = nullptr;
    fill_u64(d, 0, 1, s);
    hipLaunchKernelGGL(k_lds_rank_check, dim3(64), dim3(1024), 0, s, 256u, d);
    if (hipGetLastError() == hipSuccess &&
        hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess)
      ok = (h == 0);
  });
  return ok;
}

hipStream_t stream() {
  Ctx& c = ctx();
  if (c.stream == nullptr && !c.stream_set_by_user) {
    // Blocking stream at default priority: orders after work on the legacy
    // null stream (where torch's default-stream kernels run).
    NFH_CHECK(hipStreamCreate(&c.stream));
  }
  return c.stream;
}

uint32_t next_epoch() {
  Ctx& c = ctx();
  c.epoch = (c.epoch + 1) & 0x3FFFFFFFu;
  if (c.epoch == 0) c.epoch = 1;
  return c.epoch;
}

void* scratch(int id, size_t bytes) {
  Ctx& c = ctx();
  if (bytes == 0) bytes = 256;
  if (c.cap[id] < bytes) {
    if (c.slot[id] != nullptr) {
      // The stream may still use the old buffer.
      NFH_CHECK(hipStreamSynchronize(stream()));
      NFH_CHECK(hipFree(c.slot[id]));
      c.slot[id] = nullptr;
      c.cap[id] = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom for slowly growing sizes
    want = (want + 4095) & ~size_t(4095);
    void* p = nullptr;
    NFH_CHECK(hipMalloc(&p, want));
    c.slot[id] = p;
    c.cap[id] = want;
  }
  return c.slot[id];
}

void* pinned(size_t bytes) {
  Ctx& c = ctx();
  if (c.pinned_cap < bytes) {
    if (c.pinned != nullptr) {
      NFH_CHECK(hipStreamSynchronize(stream()));
      NFH_CHECK(hipHostFree(c.pinned));
      c.pinned = nullptr;
      c.pinned_cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    NFH_CHECK(hipHostMalloc(&c.pinned, want, hipHostMallocDefault));
    c.pinned_cap = want;
  }
  return c.pinned;
}

bool is_device_ptr(const void* p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: clear the status
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
         a.type == hipMemoryTypeUnified;
}

Stage::Stage() {
  ensure_device();
  s = stream();
  // one-time device checks run here, before any caller could be capturing
  (void)lds_rank_supported();
}

const void* Stage::in_raw(const void* p, size_t bytes, int slot) {
  if (p == nullptr || bytes == 0) return p;
  if (is_device_ptr(p)) return p;
  void* d = scratch(slot, bytes);
  NFH_CHECK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, s));
  return d;
}

void* Stage::out_raw(void* p, size_t bytes, int slot, bool copy_in) {
  if (p == nullptr || bytes == 0) return p;
  if (is_device_ptr(p)) return p;
  void* d = scratch(slot, bytes);
  if (copy_in) NFH_CHECK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, s));
  pend_.push_back({p, d, bytes});
  return d;
}

void Stage::set_out_bytes(const void* host_p, size_t bytes) {
  for (auto& q : pend_)
    if (q.host == host_p && bytes < q.bytes) q.bytes = bytes;
}

void Stage::finish() {
  for (auto& q : pend_)
    if (q.bytes > 0)
      NFH_CHECK(hipMemcpyAsync(q.host, q.dev, q.bytes, hipMemcpyDeviceToHost, s));
  pend_.clear();
  NFH_CHECK(hipStreamSynchronize(s));
  prof_collect();
}

// ---------------------------------------------------------------- profiling
namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
struct ProfAcc {
  double ms = 0;
  uint64_t launches = 0;
  double bytes = 0;
};
std::map<std::string, ProfAcc> g_prof;
struct Pending {
  const char* name;
  hipEvent_t a, b;
  double bytes;
};
thread_local std::vector<Pending> t_pending;
}  // namespace

namespace {
thread_local bool t_prof_suspended = false;
}  // namespace

ProfSuspend::ProfSuspend() : prev_(t_prof_suspended) { t_prof_suspended = true; }
ProfSuspend::~ProfSuspend() { t_prof_suspended = prev_; }

ProfScope::ProfScope(const char* name, hipStream_t s, double bytes)
    : name_(name), s_(s), bytes_(bytes) {
  if (!g_prof_on || t_prof_suspended) return;
  NFH_CHECK(hipEventCreate(&a_));
  NFH_CHECK(hipEventCreate(&b_));
  NFH_CHECK(hipEventRecord(a_, s_));
}

ProfScope::~ProfScope() {
  if (a_ == nullptr) return;
  if (hipEventRecord(b_, s_) == hipSuccess) {
    t_pending.push_back({name_, a_, b_, bytes_});
  } else {
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
  }
}

void prof_add_bytes(const char* name, double bytes) {
  for (auto it = t_pending.rbegin(); it != t_pending.rend(); ++it)
    if (std::strcmp(it->name, name) == 0) {
      it->bytes += bytes;
      return;
    }
}

void prof_collect() {
  if (t_pending.empty()) return;
  std::vector<Pending> p;
  p.swap(t_pending);
  for (auto& e : p) {
    float ms = 0;
    if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      std::lock_guard<std::mutex> g(g_prof_mu);
      ProfAcc& acc = g_prof[e.name];
      acc.ms += ms;
      acc.launches += 1;
      acc.bytes += e.bytes;
    }
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t count) {
  size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, size_t count) {
  size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

void fill_u32(void* p, uint32_t v, size_t count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s,
                     static_cast<uint32_t*>(p), v, count);
  NFH_LAUNCH_CHECK();
}
void fill_u64(void* p, uint64_t v, size_t count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s,
                     static_cast<uint64_t*>(p), v, count);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh

// ------------------------------------------------------------ additive ABI
// Run this thread's calls on a caller-owned stream (nullptr restores the
// library's own per-thread stream).  No reference counterpart: a
// device-resident caller uses it to order library work after its own.
NF_EXPORT int32_t nfh_set_stream(void* stream) {
  nfh::Ctx& c = nfh::ctx();
  if (stream == nullptr) {
    c.stream_set_by_user = false;
    c.stream = nullptr;
  } else {
    c.stream = static_cast<hipStream_t>(stream);
    c.stream_set_by_user = true;
  }
  return 0;
}

// 1 when a GPU is usable by this library, 0 otherwise (never throws).
NF_EXPORT int32_t nfh_device_ok(void) {
  try {
    nfh::ensure_device();
    return 1;
  } catch (...) {
    return 0;
  }
}

// ABI version of this library's additive surface.
NF_EXPORT int32_t nfh_abi_version(void) { return 1; }

// Per-kernel event timing on the launch stream: on != 0 enables it.
NF_EXPORT int32_t nfh_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> g(nfh::g_prof_mu);
  nfh::g_prof_on = on != 0;
  return 0;
}

NF_EXPORT int32_t nfh_profile_reset(void) {
  std::lock_guard<std::mutex> g(nfh::g_prof_mu);
  nfh::g_prof.clear();
  return 0;
}

// Total milliseconds and launch count recorded for kernel `name`; returns 0
// when the name is known, -1 otherwise.
NF_EXPORT int32_t nfh_profile_get(const char* name, double* ms, uint64_t* launches) {
  try {
    nfh::prof_collect();
  } catch (...) {
  }
  std::lock_guard<std::mutex> g(nfh::g_prof_mu);
  auto it = nfh::g_prof.find(name ? std::string(name) : std::string());
  if (it == nfh::g_prof.end()) return -1;
  if (ms) *ms = it->second.ms;
  if (launches) *launches = it->second.launches;
  return 0;
}

// Algorithmic bytes recorded for kernel `name` (sum over its launches).
NF_EXPORT int32_t nfh_profile_bytes(const char* name, double* bytes) {
  std::lock_guard<std::mutex> g(nfh::g_prof_mu);
  auto it = nfh::g_prof.find(name ? std::string(name) : std::string());
  if (it == nfh::g_prof.end()) return -1;
  if (bytes) *bytes = it->second.bytes;
  return 0;
}

// Names of every kernel recorded so far, '\n'-separated, into buf (cap
// bytes, NUL-terminated); returns the length needed.
NF_EXPORT int64_t nfh_profile_names(char* buf, size_t cap) {
  std::lock_guard<std::mutex> g(nfh::g_prof_mu);
  std::string all;
  for (auto& kv : nfh::g_prof) {
    all += kv.first;
    all += '\n';
  }
  if (buf && cap) {
    const size_t k = all.size() < cap - 1 ? all.size() : cap - 1;
    std::memcpy(buf, all.data(), k);
    buf[k] = 0;
  }
  return int64_t(all.size());
}
