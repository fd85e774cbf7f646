// exactsum.hpp -- order-independent float64 sums for the dense group-by.
//
// The reference sums float64 in row order (numfast-native/src/groupby/
// dense.rs:5-7); atomics on a GPU visit rows in no fixed order.  Instead
// every value is turned into an integer multiple of 2^L, with L = E - 116
// for E the largest binary exponent of the column, and added as four
// 30-bit chunks into int64 accumulators.  Integer adds commute, so the
// result does not depend on the visiting order, and the final conversion
// rounds the exact total once (round-to-nearest-even).  Values below 2^L
// (more than 64 binades under the column maximum) are truncated: an error
// below 2^-116 relative to the column maximum per value.
// NaN / +-inf are tracked as flags and give NaN, +inf or -inf like a
// sequential IEEE sum would.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nfh {

constexpr int XS_CHUNKS = 4;
constexpr int XS_CHUNK_BITS = 30;
constexpr uint32_t XS_FLAG_NAN = 1, XS_FLAG_PINF = 2, XS_FLAG_NINF = 4;

// Binary exponent (as used by the chunk scale) of a finite |v| bit pattern.
__host__ __device__ inline int xs_exponent_of_bits(uint64_t absbits) {
  const int b = int(absbits >> 52);
  return b == 0 ? -1022 : b - 1023;
}

// Split one finite value into signed chunks of |v| * 2^-L, L = E - 116.
__device__ inline void xs_chunks(double v, int E, int64_t c[XS_CHUNKS]) {
  const uint64_t bits = __double_as_longlong(v);
  const bool neg = bits >> 63;
  const int b = int((bits >> 52) & 0x7FF);
  uint64_t m = bits & 0xFFFFFFFFFFFFFull;
  int e;
  if (b == 0) {
    e = -1022;
  } else {
    m |= 1ull << 52;
    e = b - 1023;
  }
  const int shift = e - E + 64;  // |q| = m * 2^shift, shift <= 64
  unsigned __int128 q;
  if (shift >= 0)
    q = (unsigned __int128)m << shift;
  else if (shift > -64)
    q = (unsigned __int128)(m >> (-shift));
  else
    q = 0;
  const uint64_t mask = (1ull << XS_CHUNK_BITS) - 1;
#pragma unroll
  for (int i = 0; i < XS_CHUNKS; ++i) {
    const int64_t ch = int64_t(uint64_t(q >> (XS_CHUNK_BITS * i)) & mask);
    c[i] = neg ? -ch : ch;
  }
}

__device__ inline uint32_t xs_special_flag(double v) {
  if (v != v) return XS_FLAG_NAN;
  return v > 0 ? XS_FLAG_PINF : XS_FLAG_NINF;
}

__device__ inline bool xs_finite(double v) {
  return ((__double_as_longlong(v) >> 52) & 0x7FF) != 0x7FF;
}

// sum_i S[i] * 2^(30 i) * 2^L, correctly rounded (ties to even).
__device__ inline double xs_finish(const int64_t S[XS_CHUNKS], int E, uint32_t flags) {
  if (flags & XS_FLAG_NAN) return __longlong_as_double(0x7FF8000000000000ll);
  if ((flags & XS_FLAG_PINF) && (flags & XS_FLAG_NINF))
    return __longlong_as_double(0x7FF8000000000000ll);
  if (flags & XS_FLAG_PINF) return __longlong_as_double(0x7FF0000000000000ll);
  if (flags & XS_FLAG_NINF) return __longlong_as_double(0xFFF0000000000000ll);
  // 192-bit two's complement accumulate: T = sum S[i] << (30 i).
  uint64_t t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
  for (int i = 0; i < XS_CHUNKS; ++i) {
    const int sh = XS_CHUNK_BITS * i;  // 0, 30, 60, 90
    const int64_t s = S[i];
    const uint64_t ext = s < 0 ? ~0ull : 0ull;  // sign extension
    // limbs of (s << sh) as 192-bit two's complement
    uint64_t a0, a1, a2;
    const uint64_t u = uint64_t(s);
    if (sh == 0) {
      a0 = u;
      a1 = ext;
      a2 = ext;
    } else if (sh < 64) {
      a0 = u << sh;
      a1 = (u >> (64 - sh)) | (ext << sh);
      a2 = ext;
    } else {
      const int r = sh - 64;  // 26
      a0 = 0;
      a1 = u << r;
      a2 = (u >> (64 - r)) | (ext << r);
    }
    uint64_t c = 0;
    uint64_t x = t0 + a0;
    c = x < t0;
    t0 = x;
    uint64_t y = t1 + a1;
    uint64_t c1 = y < t1;
    y += c;
    c1 |= (y < c);
    t1 = y;
    t2 = t2 + a2 + c1;
  }
  const bool neg = (t2 >> 63) != 0;
  if (neg) {  // magnitude = -T
    t0 = ~t0;
    t1 = ~t1;
    t2 = ~t2;
    t0 += 1;
    if (t0 == 0) {
      t1 += 1;
      if (t1 == 0) t2 += 1;
    }
  }
  if ((t0 | t1 | t2) == 0) return 0.0;
  int h;  // index of the highest set bit
  if (t2)
    h = 128 + 63 - __clzll(t2);
  else if (t1)
    h = 64 + 63 - __clzll(t1);
  else
    h = 63 - __clzll(t0);
  const int L = E - 116;
  // Drop `drop` low bits, keep 53 (or fewer in the subnormal range).
  int drop = h - 52;
  if (h - 52 + L < -1074) drop = -1074 - L;
  uint64_t mant;
  bool round = false, sticky = false;
  if (drop <= 0) {
    mant = t0 << (-drop);
  } else {
    // mant = T >> drop, round = bit drop-1, sticky = bits below
    auto bit = [&](int k) -> uint64_t {
      if (k < 64) return (t0 >> k) & 1;
      if (k < 128) return (t1 >> (k - 64)) & 1;
      return (t2 >> (k - 128)) & 1;
    };
    auto shr = [&](int k) -> uint64_t {  // low 64 bits of T >> k
      if (k == 0) return t0;
      if (k < 64) return (t0 >> k) | (t1 << (64 - k));
      if (k == 64) return t1;
      if (k < 128) return (t1 >> (k - 64)) | (t2 << (128 - k));
      if (k == 128) return t2;
      return t2 >> (k - 128);
    };
    mant = shr(drop);
    round = bit(drop - 1) != 0;
    const int sb = drop - 1;  // bits [0, sb) form the sticky
    if (sb >= 128)
      sticky = t0 != 0 || t1 != 0 || (sb > 128 && (t2 << (192 - sb)) != 0);
    else if (sb >= 64)
      sticky = t0 != 0 || (sb > 64 && (t1 << (128 - sb)) != 0);
    else
      sticky = sb > 0 && (t0 << (64 - sb)) != 0;
  }
  if (round && (sticky || (mant & 1))) mant += 1;
  double r = ldexp(double(mant), drop + L);
  return neg ? -r : r;
}

}  // namespace nfh
