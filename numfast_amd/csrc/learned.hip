// learned.hip -- learned-CDF bucket sort (exact) and its GPU training.
//
// Model: a 1 -> 32 -> 32 -> 1 fp32 MLP with ReLU hidden layers maps the
// normalised key word x = (w - wmin) / span to an estimate of the key CDF.
//
// Training (cdf_train_i32): a strided sample of 2^16 keys is sorted by the
// plain radix argsort, whose histogram kernel also runs over the full
// column; the training set is the sample's empirical CDF points plus the
// exact CDF at every top-byte boundary from that histogram.  The network
// starts as a 32-piece piecewise-linear interpolant of the sample CDF
// (layer 1: ReLU ramps at greedily placed knots of the empirical CDF,
// layer 2: identity, layer 3: slope changes) and is refined on the device
// by full-batch gradient descent: k_mlp_grad (forward + backward, one
// sample per thread, per-workgroup partial gradients) and k_mlp_step
// (fixed-order reduction, normalised step with backtracking: a step that
// raises the training loss is undone and the step length halved, so the
// result is never worse than the initial fit).  Nothing about correctness
// depends on the fit.
//
// Sort (learned_argsort_i32):
//   1. the model is tabulated at 1025 knots into integer codes scaled to
//      [0, 2^bits * 16], made non-decreasing by a running max;
//   2. the bucket code of a key word interpolates the knots in integer
//      arithmetic, so code(w) is non-decreasing in w: every key of bucket b
//      is <= every key of bucket b+1;
//   3. one or two stable onesweep passes (the plain sort's pass kernel with
//      the code's bytes as digits) order (word, row) pairs by code;
//   4. bucket starts come from a binary search per code; one workgroup per
//      bucket sorts it in LDS (stable LSD over the bucket-relative key,
//      9-bit digits, as many passes as the bucket's key range needs) and
//      writes the rows;
//   5. buckets beyond the LDS capacity are either all-equal (already in
//      stable order: rows copied) or re-sorted by the plain radix argsort
//      (the overflow fix-up).
// The result is the unique stable ascending order, identical to the plain
// LSD argsort's, whatever the model predicts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "learned.hpp"
#include "onesweep.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace nfh {

constexpr int KNOTS = 1024;
constexpr int LS_SAMPLE = 1 << 16;
constexpr int LS_BLOCK = 512;
constexpr int LS_WAVES = LS_BLOCK / WAVE;
constexpr int LS_IPT = 10;
constexpr int LS_CAP = LS_BLOCK * LS_IPT;  // 5120 keys per bucket in LDS
constexpr int LS_DBITS = 9;                // in-bucket digit width
constexpr int LS_DIG = 1 << LS_DBITS;
static_assert(LS_BLOCK >= LS_DIG, "a thread per in-bucket digit in the scan");

// ------------------------------------------------------------ code map
// 32-bit integer arithmetic: with u = clamp(w - wmin, 0, span),
// p = umulhi(u << shl, mul) = floor(u * 2^21 / (span + 1)) up to a relative
// 2^-21 (shl = clz(span) normalises, mul = 2^53 / ((span + 1) << shl) in
// [2^21, 2^22]), a knot index . 11-bit fraction below 2^21; knots are in
// units of code / 16 (<= 2^20), so every product fits 32 bits.
constexpr int KN_FRAC = 11;  // fraction bits of the knot position
constexpr int KN_UNIT = 4;   // knot values carry 4 fractional code bits
struct LearnedMap {
  const uint32_t* knots;  // KNOTS + 1 non-decreasing, units of code / 16
  uint32_t wmin, span;
  uint32_t shl, mul;
  uint32_t maxcode;
};

// Non-decreasing in w: every step is a non-decreasing map (clamp, shift,
// multiply-high by mul >= 0, then linear interpolation between
// non-decreasing knots, whose value at fraction 2^KN_FRAC - 1 of knot k
// never exceeds knot k + 1).
__device__ __forceinline__ uint32_t lcode(uint32_t w, const uint32_t* kn, const LearnedMap& m) {
  uint32_t u = w > m.wmin ? w - m.wmin : 0u;
  u = u < m.span ? u : m.span;
  const uint32_t p = __umulhi(u << m.shl, m.mul);
  const uint32_t k = p >> KN_FRAC;
  const uint32_t f = p & ((1u << KN_FRAC) - 1u);
  const uint32_t t0 = kn[k], t1 = kn[k + 1];
  const uint32_t c = ((t0 << KN_FRAC) + (t1 - t0) * f) >> (KN_FRAC + KN_UNIT);
  return c < m.maxcode ? c : m.maxcode;
}

// Digit generator of the onesweep passes: the pass's digit is byte
// shift / 8 of the key's bucket code (the knots live in the pass's LDS).
// P: rank and count the run of lanes sharing the first lane's code in one
// step (nearly sorted keys share one code over long runs of rows; on
// unordered keys the ballots only cost time, so the model's hint picks).
template <bool P>
struct LearnedDigitT {
  LearnedMap m;
  int shift, next_shift;
  static constexpr size_t EXTRA_LDS = ((KNOTS + 1) * 4 + 15) / 16 * 16;
  static constexpr bool PEEL = P;
  // the first pass carries the code beside the row (codes < 2^16)
  static constexpr bool PACK = true;
  __device__ void init(unsigned char* s) const {
    uint32_t* k = reinterpret_cast<uint32_t*>(s);
    for (int i = threadIdx.x; i <= KNOTS; i += blockDim.x) k[i] = m.knots[i];
  }
  template <class W>
  __device__ uint32_t code(W w, const unsigned char* s) const {
    return lcode(uint32_t(w), reinterpret_cast<const uint32_t*>(s), m);
  }
};

// ------------------------------------------------------------ MLP
__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

__device__ float mlp_eval(const float* P, float x) {
  float h1[MLP_H];
#pragma unroll
  for (int k = 0; k < MLP_H; ++k) h1[k] = relu(fmaf(P[MLP_W1 + k], x, P[MLP_B1 + k]));
  float out = P[MLP_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float a = P[MLP_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) a = fmaf(P[MLP_W2 + i * MLP_H + j], h1[j], a);
    out = fmaf(P[MLP_W3 + i], relu(a), out);
  }
  return out;
}

constexpr int TG_BLOCK = 64;  // one wave per workgroup
constexpr int TG_GRID = 128;  // workgroups of the gradient kernel (at most)

// Forward + backward of the mean squared error over samples (x, y).  Each
// workgroup loops over chunks of TG_BLOCK samples (one per thread).  A
// thread keeps its sample's activations as a row of LDS arrays (h1, h2,
// dh2, dh1); per-chunk gradient sums run over the rows in a fixed order and
// accumulate in registers.  The workgroup then writes its gradient
// (MLP_P floats) and loss share to part[blockIdx.x * (MLP_P + 1) ...].
// pred (optional) gets the outputs.  Deterministic for a given grid.
__global__ __launch_bounds__(TG_BLOCK) void k_mlp_grad(const float* __restrict__ Pg,
                                                       const float* __restrict__ xs,
                                                       const float* __restrict__ ys, uint32_t S,
                                                       float inv_s, float* __restrict__ part,
                                                       float* __restrict__ pred) {
  __shared__ float P[MLP_P];
  __shared__ float s_h1[TG_BLOCK][MLP_H + 1];
  __shared__ float s_h2[TG_BLOCK][MLP_H + 1];
  __shared__ float s_d2[TG_BLOCK][MLP_H + 1];
  __shared__ float s_d1[TG_BLOCK][MLP_H + 1];
  __shared__ float s_x[TG_BLOCK], s_dout[TG_BLOCK], s_r2[TG_BLOCK];
  constexpr int NW2 = MLP_H * MLP_H / TG_BLOCK;  // W2 entries per thread
  const int t = threadIdx.x;
  for (int q = t; q < MLP_P; q += TG_BLOCK) P[q] = Pg[q];
  __syncthreads();
  float acc_w2[NW2];
#pragma unroll
  for (int e = 0; e < NW2; ++e) acc_w2[e] = 0.f;
  float acc_a0 = 0.f, acc_a1 = 0.f, acc_b0 = 0.f, acc_b1 = 0.f;
  for (uint32_t c0 = blockIdx.x * TG_BLOCK; c0 < S; c0 += gridDim.x * TG_BLOCK) {
    const uint32_t sidx = c0 + t;
    const bool valid = sidx < S;
    const float x = valid ? xs[sidx] : 0.f;
    const float y = valid ? ys[sidx] : 0.f;
    for (int k = 0; k < MLP_H; ++k) s_h1[t][k] = relu(fmaf(P[MLP_W1 + k], x, P[MLP_B1 + k]));
    float out = P[MLP_B3];
#pragma unroll 1
    for (int i = 0; i < MLP_H; ++i) {
      float a = P[MLP_B2 + i];
#pragma unroll 8
      for (int jj = 0; jj < MLP_H; ++jj) a = fmaf(P[MLP_W2 + i * MLP_H + jj], s_h1[t][jj], a);
      const float h2 = relu(a);
      s_h2[t][i] = h2;
      out = fmaf(P[MLP_W3 + i], h2, out);
    }
    if (valid && pred) pred[sidx] = out;
    const float r = out - y;
    const float dout = valid ? 2.f * r * inv_s : 0.f;
    for (int i = 0; i < MLP_H; ++i) s_d2[t][i] = s_h2[t][i] > 0.f ? dout * P[MLP_W3 + i] : 0.f;
#pragma unroll 1
    for (int jj = 0; jj < MLP_H; ++jj) {
      float g = 0.f;
#pragma unroll 8
      for (int i = 0; i < MLP_H; ++i) g = fmaf(P[MLP_W2 + i * MLP_H + jj], s_d2[t][i], g);
      s_d1[t][jj] = s_h1[t][jj] > 0.f ? g : 0.f;
    }
    s_x[t] = x;
    s_dout[t] = dout;
    s_r2[t] = valid ? r * r * inv_s : 0.f;
    __syncthreads();
    // sums over the chunk's samples; W2 entry t + e * TG_BLOCK is
    // (row t / 32 + 2e, column t % 32)
    {
      const int jj = t % MLP_H, i0 = t / MLP_H;
#pragma unroll 1
      for (int q = 0; q < TG_BLOCK; ++q) {
        const float h = s_h1[q][jj];
#pragma unroll
        for (int e = 0; e < NW2; ++e) acc_w2[e] = fmaf(s_d2[q][i0 + 2 * e], h, acc_w2[e]);
      }
    }
    if (t < MLP_H) {
      float gb2 = 0.f, gw1 = 0.f, gb1 = 0.f;
      for (int q = 0; q < TG_BLOCK; ++q) {
        gb2 += s_d2[q][t];
        gw1 = fmaf(s_d1[q][t], s_x[q], gw1);
        gb1 += s_d1[q][t];
      }
      acc_a0 += gb2;
      acc_b0 += gw1;
      acc_b1 += gb1;
    } else {
      const int i = t - MLP_H;  // w3[i]; lane MLP_H also sums b3 and the loss
      float gw3 = 0.f, gb3 = 0.f, l = 0.f;
      for (int q = 0; q < TG_BLOCK; ++q) {
        gw3 = fmaf(s_dout[q], s_h2[q][i], gw3);
        gb3 += s_dout[q];
        l += s_r2[q];
      }
      acc_a0 += gw3;
      acc_b0 += gb3;
      acc_b1 += l;
    }
    __syncthreads();
  }
  float* out_p = part + size_t(blockIdx.x) * (MLP_P + 1);
#pragma unroll
  for (int e = 0; e < NW2; ++e) out_p[MLP_W2 + t + e * TG_BLOCK] = acc_w2[e];
  if (t < MLP_H) {
    out_p[MLP_B2 + t] = acc_a0;
    out_p[MLP_W1 + t] = acc_b0;
    out_p[MLP_B1 + t] = acc_b1;
  } else {
    out_p[MLP_W3 + t - MLP_H] = acc_a0;
    if (t == MLP_H) {
      out_p[MLP_B3] = acc_b0;
      out_p[MLP_P] = acc_b1;
    }
  }
}

// Sum of the per-workgroup partials in workgroup order (gradient and
// loss).  With step > 0 also one training step: normalised gradient
// descent, P -= step_len * g / |g|, with backtracking.  best_p / best_loss
// hold the lowest-loss parameters seen (the loss in `part` is that of P
// before this step); when P's loss is below it the step length grows by
// 1.2, otherwise P returns to best_p, the step length halves and no update
// is made.  grad_out / loss_out (optional) receive the summed gradient and
// loss.
__global__ __launch_bounds__(256) void k_mlp_step(float* __restrict__ P,
                                                  const float* __restrict__ part, uint32_t nb,
                                                  float step, float* __restrict__ grad_out,
                                                  float* __restrict__ loss_out,
                                                  float* __restrict__ best_p,
                                                  float* __restrict__ best_loss,
                                                  float* __restrict__ step_scale) {
  __shared__ float s_g[MLP_P];
  __shared__ float s_red[4];
  __shared__ float s_loss;
  __shared__ int s_state;  // 0 plain, 1 better (save), 2 worse (revert)
  if (threadIdx.x == 0) {
    float l = 0.f;
    for (uint32_t b = 0; b < nb; ++b) l += part[size_t(b) * (MLP_P + 1) + MLP_P];
    s_loss = l;
    s_state = 0;
    if (best_loss != nullptr) s_state = (l < *best_loss) ? 1 : 2;  // NaN counts as worse
    if (loss_out) *loss_out = l;
  }
  float sq = 0.f;
  for (int q = threadIdx.x; q < MLP_P; q += blockDim.x) {
    float g = 0.f;
    for (uint32_t b = 0; b < nb; ++b) g += part[size_t(b) * (MLP_P + 1) + q];
    s_g[q] = g;
    sq = fmaf(g, g, sq);
    if (grad_out) grad_out[q] = g;
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
  if (lane_id() == 0) s_red[wave_id()] = sq;
  __syncthreads();
  const int state = s_state;
  if (state == 1) {
    for (int q = threadIdx.x; q < MLP_P; q += blockDim.x) best_p[q] = P[q];
    if (threadIdx.x == 0) {
      *best_loss = s_loss;
      if (step_scale) *step_scale *= 1.2f;
    }
  } else if (state == 2) {
    for (int q = threadIdx.x; q < MLP_P; q += blockDim.x) P[q] = best_p[q];
    if (threadIdx.x == 0 && step_scale) *step_scale *= 0.5f;
    return;
  }
  if (!(step > 0.f)) return;
  __syncthreads();
  const float norm = sqrtf(s_red[0] + s_red[1] + s_red[2] + s_red[3]);
  if (!(norm > 0.f)) return;
  const float len = step * (step_scale ? *step_scale : 1.f) / norm;
  for (int q = threadIdx.x; q < MLP_P; q += blockDim.x) P[q] -= len * s_g[q];
}

// Tabulate the model into integer knot codes, non-decreasing.
__global__ __launch_bounds__(1024) void k_cdf_knots(const float* __restrict__ P, float scale,
                                                    uint32_t* __restrict__ knots) {
  __shared__ float Ps[MLP_P];
  __shared__ uint32_t s_k[KNOTS + 1];
  __shared__ uint32_t s_tmp[16];
  for (int q = threadIdx.x; q < MLP_P; q += blockDim.x) Ps[q] = P[q];
  __syncthreads();
  for (int k = threadIdx.x; k <= KNOTS; k += blockDim.x) {
    float f = mlp_eval(Ps, float(k) / float(KNOTS));
    f = f > 0.f ? (f < 1.f ? f : 1.f) : 0.f;  // NaN -> 0
    s_k[k] = uint32_t(f * scale + 0.5f);
  }
  __syncthreads();
  // running max: per-thread chunk max, block scan of maxima, apply
  constexpr int PER = (KNOTS + 1 + 1023) / 1024;
  const int t = threadIdx.x;
  uint32_t mx = 0;
  for (int k = t * PER; k < (t + 1) * PER && k <= KNOTS; ++k) mx = s_k[k] > mx ? s_k[k] : mx;
  // inclusive max-scan over threads (wave scan + wave carries)
  uint32_t v = mx;
  const int lane = lane_id();
  for (int off = 1; off < WAVE; off <<= 1) {
    const uint32_t o = __shfl_up(v, off);
    if (lane >= off) v = o > v ? o : v;
  }
  if (lane == WAVE - 1) s_tmp[wave_id()] = v;
  __syncthreads();
  uint32_t carry = 0;
  for (int w = 0; w < wave_id(); ++w) carry = s_tmp[w] > carry ? s_tmp[w] : carry;
  uint32_t run = __shfl_up(v, 1);
  run = lane == 0 ? 0u : run;
  run = run > carry ? run : carry;  // max of everything before this thread
  for (int k = t * PER; k < (t + 1) * PER && k <= KNOTS; ++k) {
    run = s_k[k] > run ? s_k[k] : run;
    knots[k] = run;
  }
}

// ------------------------------------------------------------ sort kernels
// start[b] = first position whose code >= b (b in [0, B]); start[B] = n.
__global__ __launch_bounds__(256) void k_bucket_bounds(const uint32_t* __restrict__ W, uint32_t n,
                                                       LearnedMap m, uint32_t B,
                                                       uint32_t* __restrict__ start) {
  __shared__ uint32_t s_kn[KNOTS + 1];
  for (int i = threadIdx.x; i <= KNOTS; i += blockDim.x) s_kn[i] = m.knots[i];
  __syncthreads();
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > B) return;
  uint32_t lo = 0, hi = n;
  if (b == B) {
    lo = n;
  } else if (b > 0) {
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (lcode(W[mid], s_kn, m) < b)
        lo = mid + 1;
      else
        hi = mid;
    }
  }
  start[b] = lo;
}

// One workgroup per bucket: stable LSD sort in LDS of the bucket-relative
// key w - min (9-bit digits, only as many passes as max - min needs); the
// items carry their 16-bit position in the bucket, and the rows (held in
// registers since the start) are gathered by it at the end (6 bytes of LDS
// per key: three workgroups per CU).  Buckets over LS_CAP are listed for the large-bucket path.  Items
// are wave-contiguous (wave w holds positions [w * 64 * LS_IPT,
// (w + 1) * 64 * LS_IPT)) and ranked with lane-ordered LDS atomics.
__global__ __launch_bounds__(LS_BLOCK, 6) void k_bucket_sort(const uint32_t* __restrict__ W,
                                                             const uint32_t* __restrict__ I,
                                                             const uint32_t* __restrict__ start,
                                                             int32_t* __restrict__ perm,
                                                             uint32_t* __restrict__ large,
                                                             uint32_t* __restrict__ nlarge) {
  __shared__ uint32_t s_w[LS_CAP];
  __shared__ uint16_t s_l[LS_CAP];
  __shared__ uint32_t s_cnt[LS_WAVES][LS_DIG];
  __shared__ uint32_t s_dst[LS_DIG];
  __shared__ uint32_t s_tmp[2 * LS_WAVES];
  const uint32_t b = blockIdx.x;
  const uint32_t lo = start[b], hi = start[b + 1];
  const uint32_t size = hi - lo;
  if (size == 0) return;
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  if (size > uint32_t(LS_CAP)) {
    if (tid == 0) large[atomicAdd(nlarge, 1u)] = b;
    return;
  }
  const uint32_t wb = uint32_t(wave) * (WAVE * LS_IPT);
  // words and row ids are both loaded now (every lane from an in-range
  // position, so all loads are in flight at once); the rows are gathered
  // through LDS at the end without a second trip to memory
  // lp: the items' positions in the bucket (below 2^13), two per register
  uint32_t kw[LS_IPT], lp[LS_IPT / 2], ix[LS_IPT];
  static_assert(LS_IPT % 2 == 0, "positions are packed in pairs");
  auto lp_get = [&](int j) -> uint32_t { return (lp[j >> 1] >> (16 * (j & 1))) & 0xFFFFu; };
  auto lp_set = [&](int j, uint32_t v) {
    lp[j >> 1] = (lp[j >> 1] & ~(0xFFFFu << (16 * (j & 1)))) | (v << (16 * (j & 1)));
  };
#pragma unroll
  for (int j = 0; j < LS_IPT; ++j) {
    const uint32_t pos = wb + uint32_t(j * WAVE + lane);
    const uint32_t pc = lo + (pos < size ? pos : size - 1u);
    kw[j] = W[pc];
    ix[j] = I[pc];
  }
  for (int i = tid; i < LS_WAVES * LS_DIG; i += LS_BLOCK) (&s_cnt[0][0])[i] = 0;
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
  for (int j = 0; j < LS_IPT; ++j) {
    const uint32_t pos = wb + uint32_t(j * WAVE + lane);
    lp_set(j, pos);
    if (pos < size) {
      mn = kw[j] < mn ? kw[j] : mn;
      mx = kw[j] > mx ? kw[j] : mx;
    } else {
      kw[j] = 0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t a = __shfl_xor(mn, off), c = __shfl_xor(mx, off);
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
  }
  if (lane == 0) {
    s_tmp[wave] = mn;
    s_tmp[LS_WAVES + wave] = mx;
  }
  __syncthreads();
  mn = 0xFFFFFFFFu;
  mx = 0;
#pragma unroll
  for (int w = 0; w < LS_WAVES; ++w) {
    mn = s_tmp[w] < mn ? s_tmp[w] : mn;
    mx = s_tmp[LS_WAVES + w] > mx ? s_tmp[LS_WAVES + w] : mx;
  }
  const uint32_t range = mx - mn;
  const int bits = range == 0 ? 0 : 32 - __clz(range);
  __syncthreads();
  for (int shift = 0; shift < bits; shift += LS_DBITS) {  // uniform over the workgroup
    // (the counters are zero here: cleared at the start and after each pass)
    uint32_t rk[LS_IPT];
#pragma unroll
    for (int j = 0; j < LS_IPT; ++j) {
      const uint32_t pos = wb + uint32_t(j * WAVE + lane);
      const uint32_t d = ((kw[j] - mn) >> shift) & uint32_t(LS_DIG - 1);
      const uint32_t r = pos < size ? atomicAdd(&s_cnt[wave][d], 1u) : 0u;
      rk[j] = (r << LS_DBITS) | d;
    }
    __syncthreads();
    {
      // thread == digit (threads past the last digit scan zeros): offsets
      // of each wave's run, digit starts
      uint32_t c = 0;
      if (tid < LS_DIG) {
#pragma unroll
        for (int w = 0; w < LS_WAVES; ++w) {
          const uint32_t x = s_cnt[w][tid];
          s_cnt[w][tid] = c;
          c += x;
        }
      }
      uint32_t tot;
      const uint32_t ex = block_excl_scan<LS_WAVES>(c, s_tmp, &tot);
      if (tid < LS_DIG) s_dst[tid] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LS_IPT; ++j) {
      const uint32_t pos = wb + uint32_t(j * WAVE + lane);
      if (pos < size) {
        const uint32_t d = rk[j] & uint32_t(LS_DIG - 1);
        const uint32_t p = s_dst[d] + s_cnt[wave][d] + (rk[j] >> LS_DBITS);
        s_w[p] = kw[j];
        s_l[p] = uint16_t(lp_get(j));
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LS_IPT; ++j) {
      const uint32_t pos = wb + uint32_t(j * WAVE + lane);
      if (pos < size) {
        kw[j] = s_w[pos];
        lp_set(j, s_l[pos]);
      }
    }
    for (int i = tid; i < LS_WAVES * LS_DIG; i += LS_BLOCK) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
  }
  // the bucket's row ids into LDS by position (s_w is free after the last
  // pass), then gathered in sorted order
#pragma unroll
  for (int j = 0; j < LS_IPT; ++j) {
    const uint32_t pos = wb + uint32_t(j * WAVE + lane);
    if (pos < size) s_w[pos] = ix[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LS_IPT; ++j) {
    const uint32_t pos = wb + uint32_t(j * WAVE + lane);
    if (pos < size) perm[lo + pos] = int32_t(s_w[lp_get(j)]);
  }
}

// Over-capacity buckets, first stage: chunk offsets over the large-bucket
// list (one workgroup), then LS_CHUNK rows per workgroup copied in their
// current (stable) order, with a flag raised for each bucket holding two
// different keys.  Buckets of equal keys are then done; k_large_sort sorts
// the flagged ones.
constexpr int LS_CHUNK = 4096;
constexpr int LS_PREP_BLOCK = 1024;
constexpr int LS_COPY_GRID = 8192;

__global__ __launch_bounds__(LS_PREP_BLOCK) void k_large_prep(
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ large,
    const uint32_t* __restrict__ nlarge, uint32_t* __restrict__ cst,
    uint32_t* __restrict__ unequal) {
  __shared__ uint32_t s_tmp[LS_PREP_BLOCK / WAVE];
  const uint32_t nl = *nlarge;
  const uint32_t per = (nl + LS_PREP_BLOCK - 1) / LS_PREP_BLOCK;
  const uint32_t q0 = threadIdx.x * per;
  const uint32_t q1 = q0 + per < nl ? q0 + per : nl;
  uint32_t c = 0;
  for (uint32_t q = q0; q < q1; ++q) {
    const uint32_t b = large[q];
    c += (start[b + 1] - start[b] + LS_CHUNK - 1) / LS_CHUNK;
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<LS_PREP_BLOCK / WAVE>(c, s_tmp, &tot);
  for (uint32_t q = q0; q < q1; ++q) {
    const uint32_t b = large[q];
    cst[q] = run;
    run += (start[b + 1] - start[b] + LS_CHUNK - 1) / LS_CHUNK;
    unequal[q] = 0;
  }
  if (threadIdx.x == 0) cst[nl] = tot;
}

__global__ __launch_bounds__(256) void k_large_copy(const uint32_t* __restrict__ W,
                                                    const uint32_t* __restrict__ I,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ large,
                                                    const uint32_t* __restrict__ nlarge,
                                                    const uint32_t* __restrict__ cst,
                                                    int32_t* __restrict__ perm,
                                                    uint32_t* __restrict__ unequal) {
  const uint32_t nl = *nlarge;
  if (nl == 0) return;
  const uint32_t total = cst[nl];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    uint32_t lo = 0, hi = nl;  // last bucket with cst <= c
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (cst[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t q = lo, b = large[q];
    const uint32_t s0 = start[b], s1 = start[b + 1];
    const uint32_t first = s0 + (c - cst[q]) * uint32_t(LS_CHUNK);
    const uint32_t last = first + uint32_t(LS_CHUNK) < s1 ? first + uint32_t(LS_CHUNK) : s1;
    const uint32_t w0 = W[s0];
    bool diff = false;
    for (uint32_t i = first + threadIdx.x; i < last; i += blockDim.x) {
      perm[i] = int32_t(I[i]);
      diff |= W[i] != w0;
    }
    if (__ballot(diff) != 0 && lane_id() == 0) unequal[q] = 1u;
  }
}

// Over-capacity buckets holding different keys, one workgroup each:
//   * at most LS_LARGE_MAX keys: a stable LSD sort of w - min with 8-bit
//     digits through global memory (the bucket's own range of the free
//     ping-pong buffers).  Each pass counts the digits, then ranks chunks of
//     LS_CAP keys in order (per-wave lane-ordered LDS counters, as in the
//     in-LDS sort) on top of running digit offsets; the last pass writes
//     the permutation;
//   * larger and unequal: listed in `huge` for the host's plain radix sort.
constexpr uint32_t LS_LARGE_MAX = 1u << 17;
constexpr int LS_LARGE_GRID = 1024;

__global__ __launch_bounds__(LS_BLOCK) void k_large_sort(
    uint32_t* __restrict__ W, uint32_t* __restrict__ I, uint32_t* __restrict__ TW,
    uint32_t* __restrict__ TI, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ large, const uint32_t* __restrict__ nlarge,
    const uint32_t* __restrict__ unequal, int32_t* __restrict__ perm, uint32_t* __restrict__ huge,
    uint32_t* __restrict__ nhuge) {
  __shared__ uint32_t s_cnt[LS_WAVES][256];
  __shared__ uint32_t s_hist[256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_dst[256];
  __shared__ uint32_t s_tmp[2 * LS_WAVES];
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  const uint32_t nl = *nlarge;
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    if (!unequal[q]) continue;  // equal keys: copied by k_large_copy
    const uint32_t b = large[q];
    const uint32_t lo = start[b], m = start[b + 1] - lo;
    if (m > LS_LARGE_MAX) {
      if (threadIdx.x == 0) huge[atomicAdd(nhuge, 1u)] = b;
      continue;
    }
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (uint32_t i = uint32_t(tid); i < m; i += LS_BLOCK) {
      const uint32_t w = W[lo + i];
      mn = w < mn ? w : mn;
      mx = w > mx ? w : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t a = __shfl_xor(mn, off), c = __shfl_xor(mx, off);
      mn = a < mn ? a : mn;
      mx = c > mx ? c : mx;
    }
    if (lane == 0) {
      s_tmp[wave] = mn;
      s_tmp[LS_WAVES + wave] = mx;
    }
    __syncthreads();
    mn = 0xFFFFFFFFu;
    mx = 0;
#pragma unroll
    for (int w = 0; w < LS_WAVES; ++w) {
      mn = s_tmp[w] < mn ? s_tmp[w] : mn;
      mx = s_tmp[LS_WAVES + w] > mx ? s_tmp[LS_WAVES + w] : mx;
    }
    __syncthreads();
    const int bits = 32 - __clz(mx - mn);  // mx > mn: the bucket is flagged
    const int passes = (bits + 7) / 8;
    uint32_t* sw = W + lo;
    uint32_t* si = I + lo;
    uint32_t* dw = TW + lo;
    uint32_t* di = TI + lo;
    for (int p = 0; p < passes; ++p) {
      const int shift = 8 * p;
      const bool last = p == passes - 1;
      if (tid < 256) s_hist[tid] = 0;
      __syncthreads();
      for (uint32_t i = uint32_t(tid); i < m; i += LS_BLOCK)
        atomicAdd(&s_hist[((sw[i] - mn) >> shift) & 255u], 1u);
      __syncthreads();
      {
        uint32_t tot;
        const uint32_t ex = block_excl_scan<LS_WAVES>(tid < 256 ? s_hist[tid] : 0u, s_tmp, &tot);
        if (tid < 256) s_run[tid] = ex;
      }
      __syncthreads();
      for (uint32_t c0 = 0; c0 < m; c0 += uint32_t(LS_CAP)) {
        for (int i = tid; i < LS_WAVES * 256; i += LS_BLOCK) (&s_cnt[0][0])[i] = 0;
        __syncthreads();
        const uint32_t wb = c0 + uint32_t(wave) * (WAVE * LS_IPT);
        uint32_t kw[LS_IPT], kr[LS_IPT], rk[LS_IPT];
#pragma unroll
        for (int j = 0; j < LS_IPT; ++j) {
          const uint32_t pos = wb + uint32_t(j * WAVE + lane);
          const bool in = pos < m;
          kw[j] = in ? sw[pos] : 0u;
          kr[j] = in ? si[pos] : 0u;
        }
#pragma unroll
        for (int j = 0; j < LS_IPT; ++j) {
          const uint32_t pos = wb + uint32_t(j * WAVE + lane);
          const uint32_t d = ((kw[j] - mn) >> shift) & 255u;
          const uint32_t r = pos < m ? atomicAdd(&s_cnt[wave][d], 1u) : 0u;
          rk[j] = (r << 8) | d;
        }
        __syncthreads();
        if (tid < 256) {
          uint32_t c = 0;
#pragma unroll
          for (int w = 0; w < LS_WAVES; ++w) {
            const uint32_t x = s_cnt[w][tid];
            s_cnt[w][tid] = c;
            c += x;
          }
          s_dst[tid] = s_run[tid];
          s_run[tid] += c;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < LS_IPT; ++j) {
          const uint32_t pos = wb + uint32_t(j * WAVE + lane);
          if (pos < m) {
            const uint32_t d = rk[j] & 255u;
            const uint32_t g = s_dst[d] + s_cnt[wave][d] + (rk[j] >> 8);
            if (last) {
              perm[lo + g] = int32_t(kr[j]);
            } else {
              dw[g] = kw[j];
              di[g] = kr[j];
            }
          }
        }
        __syncthreads();
      }
      // the next pass reads what this one wrote: the barrier that ended the
      // last chunk orders those global writes before its loads
      uint32_t* t = sw;
      sw = dw;
      dw = t;
      t = si;
      si = di;
      di = t;
    }
  }
}

__global__ void k_seg_keys(const uint32_t* __restrict__ W, uint32_t lo, uint32_t m,
                           int32_t* __restrict__ out) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    out[i] = int32_t(W[lo + i] ^ 0x80000000u);
}

// perm[lo + i] = I[lo + sp[i] - off]: sp is this segment's slice of the
// concatenation's permutation (indices into the concatenation; the slice
// of a segment holds exactly its own indices, offset by off).
__global__ void k_seg_gather(const int32_t* __restrict__ sp, const uint32_t* __restrict__ I,
                             uint32_t lo, uint32_t m, uint32_t off, int32_t* __restrict__ perm) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    perm[lo + i] = int32_t(I[lo + (uint32_t(sp[i]) - off)]);
}

// ------------------------------------------------------------ host: sort
static uint32_t code_bits_for(uint32_t n) {
  int lg = 0;
  while ((uint64_t(1) << lg) < n) ++lg;
  return uint32_t(std::min(16, std::max(1, lg - 12)));
}

// Code histogram and the stable code passes: the low byte (reads the raw
// keys; counts the high byte per chain of the second pass), then the high
// byte.  (word, row) pairs end in (w1, i1) with two passes, else (w0, i0).
template <bool P>
static void code_passes(const int32_t* keys, uint32_t n, const LearnedMap& map, bool two,
                        uint32_t* hist, uint32_t* bins, uint32_t* ctr, uint64_t* lb, uint32_t* w0,
                        uint32_t* i0, uint32_t* w1, uint32_t* i1, hipStream_t s) {
  constexpr size_t CH = size_t(NCHAIN) * 256;
  chain_hist_first<int32_t>(keys, n, LearnedDigitT<P>{map, 0, 0}, false, hist,
                            static_cast<uint32_t*>(nullptr), s, "learned_code_hist");
  hipLaunchKernelGGL(k_chain_bins<256>, dim3(1), dim3(256), 0, s, hist, bins);
  NFH_LAUNCH_CHECK();
  onesweep_pass<int32_t, true, false>(keys, nullptr, nullptr, w0, reinterpret_cast<int32_t*>(i0),
                                      nullptr, bins, lb, ctr, n, LearnedDigitT<P>{map, 0, 8},
                                      two ? hist + CH : nullptr, false, s, "learned_pass");
  if (!two) return;
  hipLaunchKernelGGL(k_chain_bins<256>, dim3(1), dim3(256), 0, s, hist + CH, bins + CH);
  NFH_LAUNCH_CHECK();
  onesweep_pass<int32_t, false, false>(nullptr, w0, reinterpret_cast<const int32_t*>(i0), w1,
                                       reinterpret_cast<int32_t*>(i1), nullptr, bins + CH, lb,
                                       ctr + NCHAIN, n, LearnedDigitT<P>{map, 8, 0}, nullptr,
                                       false, s, "learned_pass");
}

void learned_argsort_i32(const int32_t* keys, uint32_t n, const CdfModel* model, int32_t* perm,
                         hipStream_t s, uint64_t* stats) {
  CdfModel hm;
  if (is_device_ptr(model)) {
    NFH_CHECK(hipMemcpyAsync(&hm, model, sizeof(hm), hipMemcpyDeviceToHost, s));
    NFH_CHECK(hipStreamSynchronize(s));
  } else {
    hm = *model;
  }
  if (stats)
    for (int i = 0; i < 4; ++i) stats[i] = 0;
  if (n < (1u << 14) || hm.span == 0 || !lds_rank_supported()) {
    // too small to bucket, a degenerate model, or a device whose LDS
    // atomics do not return in lane order (the in-bucket ranks rely on
    // it): the plain sort
    uint32_t* t0 = scratch_t<uint32_t>(S_LSEG1, n);
    uint32_t* t1 = scratch_t<uint32_t>(S_LSEG2, n);
    int32_t* tp = scratch_t<int32_t>(S_LSEG3, n);
    radix_argsort<int32_t>(keys, n, false, perm, t0, t1, tp, nullptr, s);
    return;
  }
  const uint32_t bits = code_bits_for(n);
  const uint32_t B = 1u << bits;
  // model -> knots
  float* dP = scratch_t<float>(S_LMODEL, MLP_P + KNOTS + 8);
  uint32_t* knots = reinterpret_cast<uint32_t*>(dP + MLP_P + 3);
  {
    // through the pinned buffer: a copy from pageable memory would wait
    // for the stream's earlier work before returning
    float* hp = static_cast<float*>(pinned(sizeof(float) * MLP_P));
    std::memcpy(hp, hm.p, sizeof(float) * MLP_P);
    NFH_CHECK(hipMemcpyAsync(dP, hp, sizeof(float) * MLP_P, hipMemcpyHostToDevice, s));
  }
  {
    ProfScope prof("learned_knots", s);
    hipLaunchKernelGGL(k_cdf_knots, dim3(1), dim3(1024), 0, s, dP, float(B) * float(1 << KN_UNIT),
                       knots);
    NFH_LAUNCH_CHECK();
  }
  LearnedMap map;
  map.knots = knots;
  map.wmin = hm.wmin;
  map.span = hm.span;
  static_assert(KNOTS << KN_FRAC == 1 << 21, "knot position has 21 bits");
  map.shl = uint32_t(__builtin_clz(hm.span));
  map.mul = uint32_t((uint64_t(1) << 53) / ((uint64_t(hm.span) + 1u) << map.shl));
  map.maxcode = B - 1;
  // per-chain code histograms and bin starts of the two code passes, tile
  // counters, large / huge bucket counts; the large-bucket lists hold up to
  // B entries each (at most one per bucket)
  constexpr size_t CH = size_t(NCHAIN) * 256;
  const size_t head = 4 * CH + 2 * NCHAIN + 2;
  uint32_t* meta =
      scratch_t<uint32_t>(S_LBUCKET, head + (B + 1) + 2 * size_t(B) + 2 * (size_t(B) + 1));
  uint32_t* hist = meta;            // [2][NCHAIN][256]
  uint32_t* bins = meta + 2 * CH;   // [2][NCHAIN][256]
  uint32_t* ctr = meta + 4 * CH;    // [2][NCHAIN] tile counters
  uint32_t* nlh = ctr + 2 * NCHAIN; // large buckets, huge buckets
  uint32_t* bstart = meta + head;
  uint32_t* large = bstart + (B + 1);
  uint32_t* huge = large + B;
  uint32_t* cst = huge + B;
  uint32_t* unequal = cst + (B + 1);
  fill_u32(meta, 0, head, s);
  const bool two = bits > 8;
  uint32_t* w0 = scratch_t<uint32_t>(S_L0, n);
  uint32_t* i0 = scratch_t<uint32_t>(S_L1, n);
  uint32_t* w1 = scratch_t<uint32_t>(S_L2, n);
  uint32_t* i1 = scratch_t<uint32_t>(S_L3, n);
  const size_t tiles = (n + SortGeom<int32_t>::TILE - 1) / SortGeom<int32_t>::TILE;
  uint64_t* lb = lookback_words(tiles * 256, s);
  if (hm.hints & NFH_CDF_HINT_SORTED)
    code_passes<true>(keys, n, map, two, hist, bins, ctr, lb, w0, i0, w1, i1, s);
  else
    code_passes<false>(keys, n, map, two, hist, bins, ctr, lb, w0, i0, w1, i1, s);
  const uint32_t* W = two ? w1 : w0;
  const uint32_t* I = two ? i1 : i0;
  {
    ProfScope prof("learned_bounds", s);
    hipLaunchKernelGGL(k_bucket_bounds, dim3((B + 1 + 255) / 256), dim3(256), 0, s, W, n, map, B,
                       bstart);
    NFH_LAUNCH_CHECK();
  }
  {
    ProfScope prof("learned_bucket_sort", s, double(n) * 12.0);
    hipLaunchKernelGGL(k_bucket_sort, dim3(B), dim3(LS_BLOCK), 0, s, W, I, bstart, perm, large,
                       nlh);
    NFH_LAUNCH_CHECK();
  }
  // over-capacity buckets on the device: copied in chunks (equal keys are
  // then done), the others sorted by one workgroup each; the free
  // ping-pong pair is the temporary of those global-memory passes
  {
    ProfScope prof("learned_large_prep", s);
    hipLaunchKernelGGL(k_large_prep, dim3(1), dim3(LS_PREP_BLOCK), 0, s, bstart, large, nlh,
                       cst, unequal);
    NFH_LAUNCH_CHECK();
  }
  {
    ProfScope prof("learned_large_copy", s);
    hipLaunchKernelGGL(k_large_copy, dim3(LS_COPY_GRID), dim3(256), 0, s, W, I, bstart, large,
                       nlh, cst, perm, unequal);
    NFH_LAUNCH_CHECK();
  }
  {
    ProfScope prof("learned_large_sort", s);
    uint32_t* tw = W == w1 ? w0 : w1;
    uint32_t* ti = W == w1 ? i0 : i1;
    hipLaunchKernelGGL(k_large_sort, dim3(LS_LARGE_GRID), dim3(LS_BLOCK), 0, s,
                       const_cast<uint32_t*>(W), const_cast<uint32_t*>(I), tw, ti, bstart, large,
                       nlh, unequal, perm, huge, nlh + 1);
    NFH_LAUNCH_CHECK();
  }
  uint32_t counts[2];
  NFH_CHECK(hipMemcpyAsync(counts, nlh, 8, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  const uint32_t nl = counts[0], nh = counts[1];
  std::vector<uint32_t> st;
  if (stats || nh > 0) {
    st.resize(B + 1);
    NFH_CHECK(hipMemcpyAsync(st.data(), bstart, (B + 1) * 4, hipMemcpyDeviceToHost, s));
    NFH_CHECK(hipStreamSynchronize(s));
  }
  if (stats) {
    uint64_t nonempty = 0, mx = 0;
    for (uint32_t b = 0; b < B; ++b) {
      const uint64_t sz = st[b + 1] - st[b];
      nonempty += sz > 0;
      mx = std::max(mx, sz);
    }
    stats[0] = nonempty;
    stats[1] = mx;
    stats[2] = nl;
    stats[3] = nh;
  }
  if (nh == 0) return;
  // overflow fix-up for the buckets too large for one workgroup (and not
  // all equal): they are concatenated and sorted by one plain radix
  // argsort.  Buckets cover disjoint, increasing key ranges and keep their
  // rows in input order, so the stable order of the concatenation is each
  // bucket's stable order.
  std::vector<uint32_t> lst(nh);
  NFH_CHECK(hipMemcpyAsync(lst.data(), huge, nh * 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  std::sort(lst.begin(), lst.end());
  uint32_t total = 0;
  std::vector<uint32_t> fseg;  // {lo, m, offset in the concatenation}
  for (uint32_t k = 0; k < nh; ++k) {
    const uint32_t lo = st[lst[k]], m = st[lst[k] + 1] - lo;
    fseg.insert(fseg.end(), {lo, m, total});
    total += m;
  }
  int32_t* sk = scratch_t<int32_t>(S_LSEG1, total);
  int32_t* sp = scratch_t<int32_t>(S_LSEG2, total);
  uint32_t* t0 = scratch_t<uint32_t>(S_LSEG3, total);
  uint32_t* t1 = scratch_t<uint32_t>(S_LSEG4, total);
  int32_t* tp = scratch_t<int32_t>(S_LTRAIN, total);
  for (size_t q = 0; q < fseg.size(); q += 3) {
    hipLaunchKernelGGL(k_seg_keys, dim3(grid_for(fseg[q + 1], 256, 4096)), dim3(256), 0, s, W,
                       fseg[q], fseg[q + 1], sk + fseg[q + 2]);
    NFH_LAUNCH_CHECK();
  }
  radix_argsort<int32_t>(sk, total, false, sp, t0, t1, tp, nullptr, s);
  for (size_t q = 0; q < fseg.size(); q += 3) {
    hipLaunchKernelGGL(k_seg_gather, dim3(grid_for(fseg[q + 1], 256, 4096)), dim3(256), 0, s,
                       sp + fseg[q + 2], I, fseg[q], fseg[q + 1], fseg[q + 2], perm);
    NFH_LAUNCH_CHECK();
  }
}

// ------------------------------------------------------------ host: training
__global__ void k_train_sample(const int32_t* __restrict__ keys, uint32_t n, uint32_t S,
                               int32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S) out[i] = keys[uint32_t((uint64_t(i) * n) / S)];
}

// Training pairs from the sorted sample: x = normalised word, y = midpoint
// of the key's empirical CDF step (equal keys share one target).
__global__ void k_train_pairs(const int32_t* __restrict__ sorted, uint32_t S, uint32_t wmin,
                              float inv_span, float* __restrict__ x, float* __restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  const int32_t k = sorted[i];
  uint32_t lo = 0, hi = i;  // first index of k
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (sorted[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t a = i + 1, b = S;  // one past the last index of k
  while (a < b) {
    const uint32_t mid = (a + b) / 2;
    if (sorted[mid] <= k)
      a = mid + 1;
    else
      b = mid;
  }
  const uint32_t w = uint32_t(k) ^ 0x80000000u;
  x[i] = float(double(w - wmin) * double(inv_span));
  y[i] = float((double(lo) + double(a)) * 0.5 / double(S));
}

// Initial network: the piecewise-linear interpolant of the sample's
// empirical CDF through 33 knots chosen greedily (each new knot at the
// sample point farthest, vertically, from the current interpolant), written
// as layer 1 = ReLU ramps at the knots, layer 2 = identity, layer 3 = slope
// changes.  Gaps between clusters get knots at their edges.
// The alternative start (training flag 1): 32 ramps at uniformly spaced
// knots with only the first one weighted, i.e. the linear CDF f(x) = x;
// training then has to place every slope change itself.
static void init_uniform(float* P) {
  for (int q = 0; q < MLP_P; ++q) P[q] = 0.f;
  for (int k = 0; k < MLP_H; ++k) {
    P[MLP_W1 + k] = 1.f;
    P[MLP_B1 + k] = -float(k) / MLP_H;
    P[MLP_W2 + k * MLP_H + k] = 1.f;
  }
  P[MLP_W3] = 1.f;
}

static void init_params(const std::vector<int32_t>& sorted, uint32_t wmin, uint32_t span,
                        float* P) {
  for (int q = 0; q < MLP_P; ++q) P[q] = 0.f;
  const size_t S = sorted.size();
  std::vector<double> x(S), y(S);
  for (size_t i = 0; i < S; ++i) {
    x[i] = double((uint32_t(sorted[i]) ^ 0x80000000u) - wmin) / double(span);
    y[i] = (double(i) + 0.5) / double(S);
  }
  std::vector<size_t> kn = {0, S - 1};
  auto worst = [&](size_t a, size_t b, double* err) {
    size_t arg = a;
    *err = -1.0;
    const double dx = x[b] - x[a];
    for (size_t i = a + 1; i < b; ++i) {
      const double t = dx > 0 ? (x[i] - x[a]) / dx : 0.0;
      const double e = std::fabs(y[i] - (y[a] + t * (y[b] - y[a])));
      if (e > *err) {
        *err = e;
        arg = i;
      }
    }
    return arg;
  };
  while (int(kn.size()) < MLP_H + 1) {
    double best = -1.0;
    size_t bi = 0, pos = 0;
    for (size_t s = 0; s + 1 < kn.size(); ++s) {
      double e;
      const size_t a = worst(kn[s], kn[s + 1], &e);
      if (e > best) {
        best = e;
        bi = a;
        pos = s + 1;
      }
    }
    if (best <= 0.0) break;
    kn.insert(kn.begin() + long(pos), bi);
  }
  float prev = 0.f;
  for (int k = 0; k < MLP_H; ++k) {
    P[MLP_W2 + k * MLP_H + k] = 1.f;  // identity second layer
    if (k + 1 >= int(kn.size())) {
      P[MLP_B1 + k] = -1.f;
      continue;
    }
    const size_t a = kn[size_t(k)], b = kn[size_t(k) + 1];
    P[MLP_W1 + k] = 1.f;
    P[MLP_B1 + k] = -float(x[a]);
    const double dx = std::max(x[b] - x[a], 1e-7);
    const float slope = float(std::min((y[b] - y[a]) / dx, 1e6));
    P[MLP_W3 + k] = slope - prev;
    prev = slope;
  }
  P[MLP_B3] = float(y[0]);
}

float cdf_train_i32(const int32_t* keys, uint32_t n, uint32_t steps, uint32_t flags,
                    CdfModel* model, hipStream_t s) {
  const uint32_t S = std::min<uint32_t>(n, LS_SAMPLE);
  // strided sample, sorted by the plain radix argsort
  int32_t* smp = scratch_t<int32_t>(S_LTRAIN, size_t(S) * 6 + 4096);
  int32_t* sorted = smp + S;
  int32_t* sperm = sorted + S;
  uint32_t* t0 = reinterpret_cast<uint32_t*>(sperm + S);
  uint32_t* t1 = t0 + S;
  int32_t* tp = reinterpret_cast<int32_t*>(t1 + S);
  hipLaunchKernelGGL(k_train_sample, dim3((S + 255) / 256), dim3(256), 0, s, keys, n, S, smp);
  NFH_LAUNCH_CHECK();
  radix_argsort<int32_t>(smp, S, false, sperm, t0, t1, tp, sorted, s);
  // the plain kernel's digit histograms over the whole column
  uint32_t* hist = scratch_t<uint32_t>(S_LBUCKET, 4 * 256);
  radix_histograms_i32(keys, n, hist, s);
  std::vector<int32_t> hs(S), hu(S);
  std::vector<uint32_t> hh(4 * 256);
  NFH_CHECK(hipMemcpyAsync(hs.data(), sorted, size_t(S) * 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipMemcpyAsync(hu.data(), smp, size_t(S) * 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipMemcpyAsync(hh.data(), hist, 4 * 256 * 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  const uint32_t wmin = uint32_t(hs.front()) ^ 0x80000000u;
  const uint32_t wmax = uint32_t(hs.back()) ^ 0x80000000u;
  model->wmin = wmin;
  model->span = wmax - wmin;
  // sortedness of the column in row order, from the strided sample
  size_t asc = 0;
  for (size_t i = 1; i < size_t(S); ++i) asc += hu[i - 1] <= hu[i];
  model->hints = (S > 1 && 4 * asc >= 3 * size_t(S - 1)) ? NFH_CDF_HINT_SORTED : 0u;
  if (model->span == 0) {
    for (int q = 0; q < MLP_P; ++q) model->p[q] = 0.f;
    return 0.f;
  }
  const float inv_span = float(1.0 / double(model->span));
  // exact CDF points at the top-byte boundaries inside [wmin, wmax]
  std::vector<float> hx, hy;
  uint64_t cum = 0;
  for (int b = 0; b < 255; ++b) {
    cum += hh[3 * 256 + b];
    const uint32_t wb = uint32_t(b + 1) << 24;
    if (wb <= wmin || wb > wmax) continue;
    hx.push_back(float(double(wb - wmin) / double(model->span)));
    hy.push_back(float(double(cum) / double(n)));
  }
  const uint32_t T = S + uint32_t(hx.size());
  float* xs = scratch_t<float>(S_LSEG0, size_t(T) * 2 + MLP_P * 5 + 64);
  float* ys = xs + T;
  float* P = ys + T;
  float* best = P + MLP_P;
  float* loss = best + MLP_P;  // [0] last loss, [1] best loss, [2] step scale
  hipLaunchKernelGGL(k_train_pairs, dim3((S + 255) / 256), dim3(256), 0, s, sorted, S, wmin,
                     inv_span, xs, ys);
  NFH_LAUNCH_CHECK();
  if (!hx.empty()) {
    NFH_CHECK(hipMemcpyAsync(xs + S, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, s));
    NFH_CHECK(hipMemcpyAsync(ys + S, hy.data(), hy.size() * 4, hipMemcpyHostToDevice, s));
  }
  std::vector<float> p0(MLP_P);
  if (flags & 1u)
    init_uniform(p0.data());
  else
    init_params(hs, wmin, model->span, p0.data());
  NFH_CHECK(hipMemcpyAsync(P, p0.data(), MLP_P * 4, hipMemcpyHostToDevice, s));
  NFH_CHECK(hipMemcpyAsync(best, p0.data(), MLP_P * 4, hipMemcpyHostToDevice, s));
  fill_u32(loss + 1, 0x7F800000u, 1, s);  // best loss = +inf
  fill_u32(loss + 2, 0x3F800000u, 1, s);  // step scale = 1.0f
  const uint32_t nb = std::min<uint32_t>(TG_GRID, (T + TG_BLOCK - 1) / TG_BLOCK);
  float* part = scratch_t<float>(S_LSEG1, size_t(nb) * (MLP_P + 1));
  const float step = 1e-3f;  // initial step length; backtracking adapts it
  for (uint32_t it = 0; it <= steps; ++it) {
    // the last round only evaluates (and keeps the best parameters)
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TG_BLOCK), 0, s, P, xs, ys, T, 1.f / float(T),
                       part, nullptr);
    NFH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(256), 0, s, P, part, nb,
                       it < steps ? step : 0.f, nullptr, loss, best, loss + 1, loss + 2);
    NFH_LAUNCH_CHECK();
  }
  float hl = 0.f;
  NFH_CHECK(hipMemcpyAsync(model->p, best, MLP_P * 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipMemcpyAsync(&hl, loss + 1, 4, hipMemcpyDeviceToHost, s));
  NFH_CHECK(hipStreamSynchronize(s));
  return hl;
}

// Gradient of the mean squared error at params P over (x, y): the same
// kernels as training, for checking against the oracle.
void mlp_grad_device(const float* P, const float* x, const float* y, uint32_t S, float* grad,
                     float* loss, float* pred, hipStream_t s) {
  const uint32_t nb = std::min<uint32_t>(TG_GRID, (S + TG_BLOCK - 1) / TG_BLOCK);
  float* part = scratch_t<float>(S_LSEG1, size_t(nb) * (MLP_P + 1));
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TG_BLOCK), 0, s, P, x, y, S, 1.f / float(S), part,
                     pred);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(256), 0, s, nullptr, part, nb, 0.f, grad, loss,
                     nullptr, nullptr, nullptr);
  NFH_LAUNCH_CHECK();
}

}  // namespace nfh

using namespace nfh;

// Additive (no reference counterpart): train the CDF model on a column.
// flags bit 0: start from the linear CDF instead of the greedy
// piecewise-linear fit.  Returns 0, -1 null, -3 n > i32::MAX, -6 device
// error.  *loss_out (may be null): final mean squared CDF error on the
// training points.
NF_EXPORT int32_t nfh_cdf_train_i32(const int32_t* keys, size_t n, uint32_t steps,
                                    uint32_t flags, CdfModel* model_out, float* loss_out) {
  if (!keys || !model_out) return NULL_OR_ABORT;
  if (n == 0) return MALFORMED;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  const float l = cdf_train_i32(k, uint32_t(n), steps, flags, model_out, st.s);
  if (loss_out) *loss_out = l;
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

// Additive: stable ascending argsort with a trained model; the permutation
// equals nf_sort_perm_i32(descending = 0).  stats (optional, 4 words):
// non-empty buckets, largest bucket, buckets over the LDS capacity, buckets
// re-sorted by the overflow fix-up.
NF_EXPORT int32_t nfh_sort_perm_learned_i32(const int32_t* keys, size_t n,
                                            const CdfModel* model, int32_t* perm_out,
                                            uint64_t* stats) {
  if (n == 0) return OK;
  if (!keys || !model || !perm_out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  int32_t* p = st.out(perm_out, n, S_OUT0);
  learned_argsort_i32(k, uint32_t(n), model, p, st.s, stats);
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

// Additive: mean-squared-error gradient of the CDF network (training
// kernels) at params[1153] over (x, y); for checking against the oracle.
NF_EXPORT int32_t nfh_mlp_grad(const float* params, const float* x, const float* y, size_t n,
                               float* grad_out, float* loss_out, float* pred_out) {
  if (n == 0) return OK;
  if (!params || !x || !y || !grad_out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const float* P = st.in(params, MLP_P, S_IN0);
  const float* dx = st.in(x, n, S_IN1);
  const float* dy = st.in(y, n, S_IN2);
  float* g = st.out(grad_out, MLP_P, S_OUT0);
  float* l = loss_out ? st.out(loss_out, 1, S_OUT1) : nullptr;
  float* pr = pred_out ? st.out(pred_out, n, S_OUT2) : nullptr;
  mlp_grad_device(P, dx, dy, uint32_t(n), g, l, pr, st.s);
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}
