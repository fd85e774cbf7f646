// sort.hpp -- device-pointer radix argsort shared by sort, unique and the
// sort-based pipelines.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include <type_traits>

namespace nfh {

template <class K>
using word_t = typename std::conditional<sizeof(K) == 4, uint32_t, uint64_t>::type;

// Stable argsort of device keys into device `perm` (int32 positions).
// W = uint32_t for int32 keys, uint64_t for int64 keys; t0/t1/tp are device
// scratch of n elements each.  When `sorted_out` is not null the keys are
// also written in sorted order (it may alias t0).  Asynchronous on `s`
// except for one small readback that decides which digit passes to skip.
template <class K>
void radix_argsort(const K* keys, size_t n, bool desc, int32_t* perm, word_t<K>* t0,
                   word_t<K>* t1, int32_t* tp, K* sorted_out, hipStream_t s);

// Device workspace of one radix argsort: `meta` holds sort_meta_words<K>()
// 32-bit words, `lb` sort_lookback_words<K>(n) look-back words.
// epoch0 == 0: each pass takes a fresh epoch from the thread's context;
// else pass a uses epoch0 + a, and lb must hold no word of those epochs.
// all_passes: every digit pass runs, with no readback (no host sync).
struct SortWs {
  uint32_t* meta = nullptr;
  uint64_t* lb = nullptr;
  uint32_t epoch0 = 0;
  bool all_passes = false;
};
constexpr int SORT_NCHAIN = 8;
constexpr uint32_t SORT_TILE_I32 = 16384;
template <class K>
constexpr size_t sort_meta_words() {
  // or / and words | per pass: per-chain hist and bins | tile counters
  return 2 * sizeof(K) / 4 + size_t(2) * sizeof(K) * size_t(SORT_NCHAIN) * 256 +
         sizeof(K) * size_t(SORT_NCHAIN);
}
template <class K>
size_t sort_lookback_words(size_t n);
template <class K>
void radix_argsort_ws(const K* keys, size_t n, bool desc, int32_t* perm, word_t<K>* t0,
                      word_t<K>* t1, int32_t* tp, K* sorted_out, const SortWs& ws,
                      hipStream_t s);

// The digit histograms the radix argsort's first kernel computes (4 x 256
// bins of the ascending key word, bin p*256+d = keys whose byte p is d).
void radix_histograms_i32(const int32_t* keys, uint32_t n, uint32_t* hist, hipStream_t s);

// Look-back scratch of `words` 64-bit words (zeroed whenever it grows) and
// a fresh epoch for one decoupled look-back pass.
uint64_t* lookback_words(size_t words, hipStream_t s);
uint32_t lookback_epoch(hipStream_t s);

}  // namespace nfh
