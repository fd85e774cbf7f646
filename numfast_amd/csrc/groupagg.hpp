// groupagg.hpp -- device-pointer NULL-aware group-by and unique.
#pragma once
#include "runtime.hpp"

namespace nfh {

struct GroupAggOut {
  int64_t* ukeys = nullptr;
  int64_t* counts = nullptr;
  int64_t* sums = nullptr;
  int32_t* mins = nullptr;
  int32_t* maxs = nullptr;
  double* means = nullptr;
};

// Groups of the rows whose key and value are valid (masks may be null).
// Outputs need ng lanes (ng <= n).  Returns ng.  Synchronous.
// `known_range` ({kmin, kmax}) skips the key-range pass when the caller
// already ran it for the same rows and masks.
int64_t group_agg_device(const int32_t* keys, const int32_t* vals, const uint8_t* valid,
                         const uint8_t* key_valid, size_t n, const GroupAggOut& out,
                         hipStream_t s, const int32_t* known_range = nullptr);

// np.unique over valid rows; inv = -1 on NULL rows.  Returns ng.
int64_t unique_valid_device(const int32_t* keys, const uint8_t* valid, size_t n, int32_t* uniq,
                            int32_t* inv, hipStream_t s);

// min / max of the valid keys; false when no row is valid.
bool key_range(const int32_t* keys, const uint8_t* v, const uint8_t* kv, size_t n,
               int32_t* kmin, int32_t* kmax, hipStream_t s);

}  // namespace nfh
