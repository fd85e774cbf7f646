// gbpart.hpp -- partitioned dense group-by (no per-row global atomics).
#pragma once
#include "runtime.hpp"

namespace nfh {

// Keys per LDS bucket: 2^11 by default; 2^12 halves the partition count
// (the strategy selector's choice between the two).
constexpr int GP_BUCKET_BITS = 11;
constexpr int GP_BUCKET_BITS_WIDE = 12;

// One 32-byte group state slot, indexed by key - kmin.
struct __align__(32) GpSlot {
  unsigned long long sum;
  uint32_t cnt;
  int32_t mn;
  int32_t mx;
  uint32_t pad[3];
};
static_assert(sizeof(GpSlot) == 32, "slot is one 32-byte sector");

// Spans this path handles with 2^bucket_bits-key buckets (at most 2^13
// buckets: the histogram pass's LDS; two 7-bit partition passes).
bool gp_supported(uint64_t span, int bucket_bits = GP_BUCKET_BITS);

struct GroupAggOut;

// Groups {key, count, sum, min, max, mean} of the valid rows (vm may be
// null), keys ascending, written to `out`; returns the group count.
// `table` / `hot_keys` / `hot_n` describe the sampled hot keys (table may be
// null).  Synchronous.
int64_t gp_aggregate(const int32_t* keys, const int32_t* vals, const uint8_t* vm, size_t n,
                     int32_t kmin, uint64_t span, const int2* table, const int32_t* hot_keys,
                     const uint32_t* hot_n, const GroupAggOut& out, hipStream_t s,
                     int bucket_bits = GP_BUCKET_BITS);

}  // namespace nfh
