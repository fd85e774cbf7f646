// hot.hpp -- the sampled hot-key table shared by the group-by and unique
// kernels.
//
// Up to HOT_MAX hot keys live in a 3-choice cuckoo table of HOT_TBL slots
// (load <= 0.5).  A key may sit in one of three slots, taken from three
// 10-bit fields of one mixed hash, so a lookup is three independent LDS
// reads and no probe loop.  Empty slots hold a decoy key none of whose three
// slots is that slot, so a lookup never matches an empty slot and needs no
// separate "occupied" test.  Global form: int2 {key, hot index} per slot,
// index -1 for empty slots.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nfh {

constexpr int HOT_MAX = 512;   // hot keys privatised per workgroup
constexpr int HOT_TBL = 1024;  // cuckoo slots
constexpr int HOT_TBL_BITS = 10;
constexpr int HOT_SAMPLE = 1 << 16;  // sampled keys the hot set is chosen from

// Key hash: xor-shifts around two 24-bit multiplies (full-rate
// v_mul_u32_u24 on the GPU; a 32-bit multiply issues at a quarter of the
// rate, and every row of the hot-key kernels hashes its key).  The first
// shift folds the key's top byte into the multiplied low 24 bits, the last
// brings the product's high bits into the low slot fields.
__host__ __device__ __forceinline__ uint32_t hot_mul24(uint32_t a, uint32_t b) {
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}
__host__ __device__ __forceinline__ uint32_t hot_mix(int32_t k) {
  uint32_t h = uint32_t(k);
  h ^= h >> 16;
  h = hot_mul24(h, 0x9E3779u);
  h ^= h >> 15;
  h = hot_mul24(h, 0x85EBCAu);
  h ^= h >> 16;
  return h;
}

__host__ __device__ __forceinline__ uint32_t hot_slot(uint32_t h, int c) {
  return (h >> (HOT_TBL_BITS * c)) & uint32_t(HOT_TBL - 1);
}

// Slot of k in the LDS key array, or -1.
__device__ __forceinline__ int hot_slot_of(const int32_t* t_key, int32_t k) {
  const uint32_t h = hot_mix(k);
  const uint32_t s0 = hot_slot(h, 0), s1 = hot_slot(h, 1), s2 = hot_slot(h, 2);
  const int32_t k0 = t_key[s0], k1 = t_key[s1], k2 = t_key[s2];
  return k0 == k ? int(s0) : k1 == k ? int(s1) : k2 == k ? int(s2) : -1;
}

// Hot index of k (0 .. HOT_MAX-1), or -1.
__device__ __forceinline__ int hot_index(const int32_t* t_key, const int32_t* t_idx, int32_t k) {
  const int s = hot_slot_of(t_key, k);
  return s >= 0 ? t_idx[s] : -1;
}

// Hot indices of N keys at once (-1: not hot).  All 3N key reads are
// issued before any compare and the N index reads after them: written row
// by row, the compiler moves the second and third reads behind the first
// compare, and each row then waits for its own LDS reads in turn.
template <int N>
__device__ __forceinline__ void hot_index_batch(const int32_t* t_key, const int32_t* t_idx,
                                                const int32_t (&k)[N], int (&hi)[N]) {
  uint32_t sl[N][3];
  int32_t tk[N][3];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t h = hot_mix(k[i]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sl[i][c] = hot_slot(h, c);
      tk[i][c] = t_key[sl[i][c]];
    }
  }
  // every read above is complete here (each value is an operand)
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) asm volatile("" : "+v"(tk[i][c]));
  int s[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    s[i] = tk[i][0] == k[i] ? int(sl[i][0])
         : tk[i][1] == k[i] ? int(sl[i][1])
         : tk[i][2] == k[i] ? int(sl[i][2]) : -1;
    hi[i] = t_idx[s[i] >= 0 ? s[i] : 0];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) hi[i] = s[i] >= 0 ? hi[i] : -1;
}

// Copy of the global table into LDS.  A null table (no hot keys) loads
// key 0 / index -1 everywhere, which a lookup of key 0 would match: callers
// make no lookups without a table.  t_idx may be null (membership only).
__device__ __forceinline__ void hot_load(const int2* __restrict__ table, int32_t* t_key,
                                         int32_t* t_idx) {
  for (int b = threadIdx.x; b < HOT_TBL; b += blockDim.x) {
    const int2 e = table ? table[b] : make_int2(0, -1);
    t_key[b] = e.x;
    if (t_idx) t_idx[b] = e.y;
  }
}

// Build the table in LDS from `n` hot keys (n <= HOT_MAX, key i has index
// i).  Parallel cuckoo insertion by atomic exchange; a key still homeless
// after the walk limit is left out (it is then cold everywhere, which every
// kernel handles).  `cell` is HOT_TBL words of LDS scratch; on return
// t_key / t_idx hold the table.  Called by a whole workgroup.
__device__ inline void hot_build(const int32_t* __restrict__ keys, uint32_t n,
                                 unsigned long long* cell, int32_t* t_key, int32_t* t_idx) {
  constexpr unsigned long long EMPTY = ~0ull;
  for (int b = threadIdx.x; b < HOT_TBL; b += blockDim.x) cell[b] = EMPTY;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    unsigned long long item = (unsigned long long)uint32_t(keys[i]) |
                              ((unsigned long long)i << 32);
    uint32_t slot = hot_slot(hot_mix(int32_t(uint32_t(item))), 0);
    for (int it = 0; it < 256; ++it) {
      const unsigned long long old = atomicExch(&cell[slot], item);
      if (old == EMPTY) {
        item = EMPTY;
        break;
      }
      // the evicted key moves to its next choice after the slot it held
      item = old;
      const uint32_t h = hot_mix(int32_t(uint32_t(item)));
      const int c = hot_slot(h, 0) == slot ? 1 : hot_slot(h, 1) == slot ? 2 : 0;
      slot = hot_slot(h, c);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < HOT_TBL; b += blockDim.x) {
    const unsigned long long e = cell[b];
    if (e != EMPTY) {
      t_key[b] = int32_t(uint32_t(e));
      t_idx[b] = int32_t(uint32_t(e >> 32));
    } else {
      // decoy: a key whose three slots all differ from b
      uint32_t d = 0x9E3779B9u * uint32_t(b + 1);
      for (;; ++d) {
        const uint32_t h = hot_mix(int32_t(d));
        if (hot_slot(h, 0) != uint32_t(b) && hot_slot(h, 1) != uint32_t(b) &&
            hot_slot(h, 2) != uint32_t(b))
          break;
      }
      t_key[b] = int32_t(d);
      t_idx[b] = -1;
    }
  }
  __syncthreads();
}

}  // namespace nfh
