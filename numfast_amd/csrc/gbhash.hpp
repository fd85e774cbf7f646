// gbhash.hpp -- hash-scatter group-by (strategy 4 of nfh_group_agg_i32).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "groupagg.hpp"

namespace nfh {

// Groups of the rows whose key and value are valid (masks may be null),
// keys ascending, into `o` (ng lanes each).  `table` / `hot_keys` /
// `hot_n`: the sampled hot-key table (table may be null).  Returns ng, or
// -1 when the global hash table overflowed (the caller then uses another
// strategy).  Synchronous.
int64_t hash_group_agg(const int32_t* keys, const int32_t* vals, const uint8_t* v,
                       const uint8_t* kv, size_t n, uint64_t span, const int2* table,
                       const int32_t* hot_keys, const uint32_t* hot_n, const GroupAggOut& o,
                       hipStream_t s);

}  // namespace nfh
