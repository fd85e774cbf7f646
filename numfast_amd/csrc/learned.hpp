// learned.hpp -- learned-CDF bucket sort: a 1 -> 32 -> 32 -> 1 fp32 MLP
// fitted to the key CDF drives a monotone bucket code; keys are bucketed by
// that code in one or two stable passes and each bucket is sorted exactly in
// LDS, so the permutation equals the plain stable radix argsort's.
#pragma once
#include "runtime.hpp"

namespace nfh {

constexpr int MLP_H = 32;
// parameter layout (floats): w1[32] b1[32] w2[32][32] (row = output unit)
// b2[32] w3[32] b3
constexpr int MLP_W1 = 0, MLP_B1 = 32, MLP_W2 = 64, MLP_B2 = 64 + 1024, MLP_W3 = MLP_B2 + 32,
              MLP_B3 = MLP_W3 + 32, MLP_P = MLP_B3 + 1;

// Host / C-ABI layout of a trained model (struct nfh_cdf_model).
struct CdfModel {
  float p[MLP_P];
  uint32_t wmin;  // smallest sampled key word (order-preserving uint32)
  uint32_t span;  // largest sampled word - wmin (0: every sample equal)
  uint32_t hints;  // NFH_CDF_HINT_* bits: speed only, never the result
};
// The training sample was nearly in ascending order (at least 3/4 of its
// neighbours ascending): the code passes rank runs of rows sharing a code
// in one step.
constexpr uint32_t NFH_CDF_HINT_SORTED = 1;
static_assert(sizeof(CdfModel) == MLP_P * 4 + 12, "model layout");

// Stable ascending argsort of device int32 keys with a trained model
// (host or device pointer).  Same result as the radix argsort.
// Synchronous.  `stats` (optional, host, 4 words): buckets, largest bucket,
// buckets over the LDS capacity, of which not all-equal (sorted by the
// overflow fix-up).
void learned_argsort_i32(const int32_t* keys, uint32_t n, const CdfModel* model, int32_t* perm,
                         hipStream_t s, uint64_t* stats = nullptr);

// Train on a strided sample of the device keys (plain radix-sorted on the
// device; its digit histograms add CDF points), by normalised gradient
// descent with backtracking.  Writes *model (host).
// Returns the final training loss (mean squared CDF error).
// flags bit 0: start from the linear CDF instead of the greedy fit.
float cdf_train_i32(const int32_t* keys, uint32_t n, uint32_t steps, uint32_t flags,
                    CdfModel* model, hipStream_t s);

}  // namespace nfh
