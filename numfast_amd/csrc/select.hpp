// select.hpp -- device-pointer stable compaction and the filter / derive
// stage of the sort pipeline.
#pragma once
#include "runtime.hpp"

namespace nfh {

// out[m++] = src[i] for mask[i] != 0, order kept.  Returns m.  Synchronous.
template <class T>
uint32_t select_scatter_device(const T* src, const uint8_t* mask, uint32_t n, T* out,
                               hipStream_t s);

// filter(v > 0) -> derive(w = v * mul, int32 wrap): w[m], source row[m].
// Returns m.  Synchronous.
uint32_t filter_derive_device(const int32_t* v, uint32_t n, int32_t mul, int32_t* w,
                              int32_t* row, hipStream_t s);

}  // namespace nfh
