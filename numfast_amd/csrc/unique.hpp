// unique.hpp -- device-pointer sorted unique + inverse.
#pragma once
#include "sort.hpp"

namespace nfh {

// np.unique(keys, return_inverse=True) on device buffers; returns ng.
// perm/t0/t1/tp are device scratch of n elements each.  Synchronous.
template <class K>
int64_t unique_inverse_device(const K* keys, size_t n, K* uniq, int32_t* inv, int32_t* perm,
                              word_t<K>* t0, word_t<K>* t1, int32_t* tp, hipStream_t s);

}  // namespace nfh
