// groupby.hpp -- the dense-code group-by engine behind the nf_group_*,
// nf_pack_sum_count_* and nf_sorted_run_* entry points.
#pragma once
#include "runtime.hpp"

namespace nfh {

constexpr int MAXC = 4;  // value columns per accumulate launch

// Output sink for one value column.
struct ColOut {
  int32_t* s32 = nullptr;  // int32 sums (i32 family)
  int64_t* s64 = nullptr;  // int64 sums (i64 family)
  double* f64 = nullptr;   // float64 sums
};

// One dense aggregation over int32 codes in [0, g) (or packed k1*m2+k2).
// Only keys in [lo, hi) are accumulated; outputs are written for [lo, hi).
struct DenseSpec {
  const int32_t* keys = nullptr;
  const int32_t* k2 = nullptr;  // packed keys when not null
  int64_t m2 = 0;
  size_t n = 0;
  size_t g = 0;
  size_t lo = 0, hi = 0;
  int ni32 = 0, ni64 = 0, nf64 = 0;
  const int32_t* v32[MAXC] = {};
  const int64_t* v64[MAXC] = {};
  const double* vf[MAXC] = {};
  ColOut o32[MAXC], o64[MAXC], of[MAXC];
  int64_t* counts = nullptr;  // int64 counts out (nullptr: not wanted)
  bool check_i32 = false;     // -4 on a row-order prefix leaving int32
  bool check_i64 = false;     // -4 on a row-order prefix leaving int64
  bool presorted = false;     // keys nondecreasing over [0, n) (V_SORTED)
  bool f64_seq_first = false; // row-order f64 sums start from the first value
};

// rc (0, -2 bad key, -4 prefix overflow) of the first failing row, in row
// order like the reference loops; row = n when rc == 0.
struct AggResult {
  int32_t rc;
  uint64_t row;
};
AggResult dense_aggregate(const DenseSpec& spec, hipStream_t s);

// Combine results of column batches over the same rows.
inline AggResult agg_first(AggResult a, AggResult b) {
  if (a.rc == 0) return b;
  if (b.rc == 0) return a;
  if (a.row != b.row) return a.row < b.row ? a : b;
  return a.rc == -2 ? a : b;
}

// First row i with keys[i] < keys[i-1], or n when nondecreasing.
size_t first_regress_row(const int32_t* keys, size_t n, hipStream_t s);

}  // namespace nfh
