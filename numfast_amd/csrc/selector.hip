// selector.hip -- learned group-by strategy selector.
//
// Features (k_sel_features, one workgroup): a strided sample of up to 4096
// keys gives 64 numbers: a 56-bin histogram of the sample over
// [min, max] (fractions) and 8 summaries -- log2(n) / 32, log2(sample span)
// / 32, distinct fraction, largest key share, share of rows whose key occurs
// at least 4 times, fraction of ascending neighbour pairs, log2(distinct) /
// 16, and a constant 1.
//
// Classifier: a 64 -> 64 (ReLU) -> 5 fp32 MLP with softmax cross-entropy,
// trained on the device: k_sel_grad (forward + backward, one sample per
// thread, per-workgroup partial gradients in a fixed order) and k_sel_adam
// (reduction + Adam).  Class c is group-by strategy c + 1 (dense,
// partitioned, sort-based, hash-scatter, partitioned with half the
// partitions; groupagg.hip).  A wrong pick costs time only: every strategy
// returns the same result.
//
// Parameter layout (floats): W1[64][64] (row = hidden unit), b1[64],
// W2[5][64], b2[5].
#include <cmath>

#include "device.hpp"
#include "runtime.hpp"

namespace nfh {

constexpr int SF = 64;  // features
constexpr int SH = 64;  // hidden
constexpr int SC = 5;   // classes
constexpr int S_W1 = 0, S_B1 = SF * SH, S_W2 = S_B1 + SH, S_B2 = S_W2 + SC * SH,
              S_P = S_B2 + SC;
constexpr int SEL_SAMPLE = 4096;
constexpr int SEL_HBINS = 56;
constexpr int SEL_TBL = 8192;

__global__ __launch_bounds__(1024) void k_sel_features(const int32_t* __restrict__ keys,
                                                       uint32_t n, float* __restrict__ out) {
  __shared__ int32_t sk[SEL_SAMPLE];
  __shared__ unsigned long long tb[SEL_TBL];
  __shared__ uint32_t tc[SEL_TBL];
  __shared__ uint32_t hist[SEL_HBINS];
  __shared__ int32_t s_mn[16], s_mx[16];
  __shared__ uint32_t s_u[16][5];
  const int tid = threadIdx.x;
  const uint32_t S = n < SEL_SAMPLE ? n : SEL_SAMPLE;
  for (int i = tid; i < SEL_TBL; i += blockDim.x) {
    tb[i] = 0;
    tc[i] = 0;
  }
  if (tid < SEL_HBINS) hist[tid] = 0;
  int32_t mn = INT32_MAX, mx = INT32_MIN;
  for (uint32_t j = tid; j < S; j += blockDim.x) {
    const int32_t k = keys[uint32_t((uint64_t(j) * n) / S)];
    sk[j] = k;
    mn = k < mn ? k : mn;
    mx = k > mx ? k : mx;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const int32_t a = __shfl_xor(mn, off), b = __shfl_xor(mx, off);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if (lane_id() == 0) {
    s_mn[wave_id()] = mn;
    s_mx[wave_id()] = mx;
  }
  __syncthreads();
  mn = INT32_MAX;
  mx = INT32_MIN;
  for (int w = 0; w < int(blockDim.x / WAVE); ++w) {
    mn = s_mn[w] < mn ? s_mn[w] : mn;
    mx = s_mx[w] > mx ? s_mx[w] : mx;
  }
  const uint64_t span = uint64_t(int64_t(mx) - int64_t(mn)) + 1;
  uint32_t asc = 0;
  for (uint32_t j = tid; j < S; j += blockDim.x) {
    const int32_t k = sk[j];
    const uint32_t bin = uint32_t((uint64_t(int64_t(k) - int64_t(mn)) * SEL_HBINS) / span);
    atomicAdd(&hist[bin], 1u);
    if (j + 1 < S && k <= sk[j + 1]) ++asc;
    // insert into the sample's hash set (u64 slot: 1 << 32 | key; 0 empty)
    const unsigned long long v = (1ull << 32) | uint32_t(k);
    uint32_t h = (uint32_t(k) * 0x9E3779B1u) >> (32 - 13);
    for (;;) {
      const unsigned long long old = atomicCAS(&tb[h], 0ull, v);
      if (old == 0ull || old == v) {
        atomicAdd(&tc[h], 1u);
        break;
      }
      h = (h + 1) & (SEL_TBL - 1);
    }
  }
  __syncthreads();
  uint32_t distinct = 0, maxc = 0, hot = 0;
  for (int i = tid; i < SEL_TBL; i += blockDim.x) {
    const uint32_t c = tc[i];
    distinct += c > 0;
    maxc = c > maxc ? c : maxc;
    hot += c >= 4 ? c : 0;
  }
  uint32_t vals[4] = {asc, distinct, maxc, hot};
  for (int q = 0; q < 4; ++q) {
    uint32_t x = vals[q];
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t o = __shfl_xor(x, off);
      x = q == 2 ? (o > x ? o : x) : x + o;
    }
    if (lane_id() == 0) s_u[wave_id()][q] = x;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t tot[4] = {0, 0, 0, 0};
    for (int w = 0; w < int(blockDim.x / WAVE); ++w)
      for (int q = 0; q < 4; ++q)
        tot[q] = q == 2 ? (s_u[w][q] > tot[q] ? s_u[w][q] : tot[q]) : tot[q] + s_u[w][q];
    const float fs = float(S);
    out[SEL_HBINS + 0] = log2f(float(n)) / 32.f;
    out[SEL_HBINS + 1] = log2f(float(span)) / 32.f;
    out[SEL_HBINS + 2] = float(tot[1]) / fs;
    out[SEL_HBINS + 3] = float(tot[2]) / fs;
    out[SEL_HBINS + 4] = float(tot[3]) / fs;
    out[SEL_HBINS + 5] = S > 1 ? float(tot[0]) / float(S - 1) : 1.f;
    out[SEL_HBINS + 6] = log2f(float(tot[1])) / 16.f;
    out[SEL_HBINS + 7] = 1.f;
  }
  if (tid < SEL_HBINS) out[tid] = float(hist[tid]) / float(S);
}

// ------------------------------------------------------------ classifier
constexpr int SG_BLOCK = 64;  // one wave
constexpr int SG_GRID = 64;

// Forward + backward over (X[N][64], y[N]) of the mean cross-entropy
// (classifier), or, with Tt (targets[N][SC]) non-null, of the mean of
// 1/2 sum_c (z_c - Tt_c)^2 (cost model: z_c predicts strategy c's centred
// log time and the pick is the smallest).  Workgroups loop over chunks of
// 64 samples; per-chunk sums in a fixed order, accumulated in registers;
// part[blockIdx.x * (S_P + 1) ...] gets the workgroup's gradient and loss
// share.
__global__ __launch_bounds__(SG_BLOCK) void k_sel_grad(const float* __restrict__ P,
                                                       const float* __restrict__ X,
                                                       const int32_t* __restrict__ Y,
                                                       const float* __restrict__ Tt,
                                                       uint32_t N, float inv_n,
                                                       float* __restrict__ part) {
  __shared__ float s_x[SG_BLOCK][SF + 1];
  __shared__ float s_h[SG_BLOCK][SH + 1];
  __shared__ float s_dh[SG_BLOCK][SH + 1];
  __shared__ float s_dz[SG_BLOCK][SC];
  __shared__ float s_l[SG_BLOCK];
  const int t = threadIdx.x;
  float acc_w1[SH];  // column t of the W1 gradient
#pragma unroll
  for (int m = 0; m < SH; ++m) acc_w1[m] = 0.f;
  float acc_b1 = 0.f, acc_w2[SC], acc_b2 = 0.f, acc_l = 0.f;
#pragma unroll
  for (int c = 0; c < SC; ++c) acc_w2[c] = 0.f;
  for (uint32_t c0 = blockIdx.x * SG_BLOCK; c0 < N; c0 += gridDim.x * SG_BLOCK) {
    const uint32_t s = c0 + t;
    const bool valid = s < N;
    for (int j = 0; j < SF; ++j) s_x[t][j] = valid ? X[size_t(s) * SF + j] : 0.f;
    float z[SC];
#pragma unroll
    for (int c = 0; c < SC; ++c) z[c] = P[S_B2 + c];
#pragma unroll 1
    for (int i = 0; i < SH; ++i) {
      float a = P[S_B1 + i];
#pragma unroll 16
      for (int j = 0; j < SF; ++j) a = fmaf(P[S_W1 + i * SF + j], s_x[t][j], a);
      const float h = a > 0.f ? a : 0.f;
      s_h[t][i] = h;
#pragma unroll
      for (int c = 0; c < SC; ++c) z[c] = fmaf(P[S_W2 + c * SH + i], h, z[c]);
    }
    float dz[SC];
    if (Tt != nullptr) {  // cost model (uniform over the launch)
      float l = 0.f;
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const float d = valid ? z[c] - Tt[size_t(s) * SC + c] : 0.f;
        dz[c] = d * inv_n;
        l = fmaf(0.5f * d, d, l);
        s_dz[t][c] = dz[c];
      }
      s_l[t] = l * inv_n;
    } else {
      float zm = z[0];
#pragma unroll
      for (int c = 1; c < SC; ++c) zm = fmaxf(zm, z[c]);
      float e[SC], se = 0.f;
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        e[c] = expf(z[c] - zm);
        se += e[c];
      }
      const int y = valid ? Y[s] : 0;
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const float p = e[c] / se;
        dz[c] = valid ? (p - (c == y ? 1.f : 0.f)) * inv_n : 0.f;
        s_dz[t][c] = dz[c];
      }
      s_l[t] = valid ? (logf(se) + zm - z[y]) * inv_n : 0.f;
    }
#pragma unroll 1
    for (int i = 0; i < SH; ++i) {
      float g = 0.f;
#pragma unroll
      for (int c = 0; c < SC; ++c) g = fmaf(P[S_W2 + c * SH + i], dz[c], g);
      s_dh[t][i] = s_h[t][i] > 0.f ? g : 0.f;
    }
    __syncthreads();
    // column t of dW1 = sum_s dh[s][m] x[s][t]; b1[t]; W2[c][t]; b2[c]
#pragma unroll 1
    for (int q = 0; q < SG_BLOCK; ++q) {
      const float xv = s_x[q][t];
#pragma unroll
      for (int m = 0; m < SH; ++m) acc_w1[m] = fmaf(s_dh[q][m], xv, acc_w1[m]);
      acc_b1 += s_dh[q][t];
#pragma unroll
      for (int c = 0; c < SC; ++c) acc_w2[c] = fmaf(s_dz[q][c], s_h[q][t], acc_w2[c]);
      if (t < SC) acc_b2 += s_dz[q][t];
      if (t == 0) acc_l += s_l[q];
    }
    __syncthreads();
  }
  float* o = part + size_t(blockIdx.x) * (S_P + 1);
#pragma unroll
  for (int m = 0; m < SH; ++m) o[S_W1 + m * SF + t] = acc_w1[m];
  o[S_B1 + t] = acc_b1;
#pragma unroll
  for (int c = 0; c < SC; ++c) o[S_W2 + c * SH + t] = acc_w2[c];
  if (t < SC) o[S_B2 + t] = acc_b2;
  if (t == 0) o[S_P] = acc_l;
}

// Sum of the partials in workgroup order; with lr > 0 one Adam step
// (beta1 0.9, beta2 0.999, eps 1e-8; bias corrections bc1, bc2).
__global__ __launch_bounds__(256) void k_sel_adam(float* __restrict__ P, float* __restrict__ m1,
                                                  float* __restrict__ m2,
                                                  const float* __restrict__ part, uint32_t nb,
                                                  float lr, float bc1, float bc2,
                                                  float* __restrict__ grad_out,
                                                  float* __restrict__ loss_out) {
  for (int q = threadIdx.x; q <= S_P; q += blockDim.x) {
    float g = 0.f;
    for (uint32_t b = 0; b < nb; ++b) g += part[size_t(b) * (S_P + 1) + q];
    if (q == S_P) {
      if (loss_out) *loss_out = g;
      continue;
    }
    if (grad_out) grad_out[q] = g;
    if (lr > 0.f) {
      const float a = 0.9f * m1[q] + 0.1f * g;
      const float v = 0.999f * m2[q] + 0.001f * g * g;
      m1[q] = a;
      m2[q] = v;
      P[q] -= lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
    }
  }
}

// Pick per sample: the largest score (classifier) or, with lowest, the
// smallest predicted cost.
__global__ void k_sel_predict(const float* __restrict__ P, const float* __restrict__ X,
                              uint32_t N, bool lowest, int32_t* __restrict__ cls) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  float z[SC];
  for (int c = 0; c < SC; ++c) z[c] = P[S_B2 + c];
  for (int i = 0; i < SH; ++i) {
    float a = P[S_B1 + i];
    for (int j = 0; j < SF; ++j) a = fmaf(P[S_W1 + i * SF + j], X[size_t(s) * SF + j], a);
    const float h = a > 0.f ? a : 0.f;
    for (int c = 0; c < SC; ++c) z[c] = fmaf(P[S_W2 + c * SH + i], h, z[c]);
  }
  int best = 0;
  for (int c = 1; c < SC; ++c)
    if (lowest ? z[c] < z[best] : z[c] > z[best]) best = c;
  cls[s] = best;
}

}  // namespace nfh

using namespace nfh;

// Additive (no reference counterpart): the selector's 64 features of a
// device or host int32 key column.  0, -1, -3 or -6.
NF_EXPORT int32_t nfh_groupby_features_i32(const int32_t* keys, size_t n, float* out64) {
  if (n == 0 || !keys || !out64) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const int32_t* k = st.in(keys, n, S_IN0);
  float* o = st.out(out64, SF, S_OUT0);
  hipLaunchKernelGGL(k_sel_features, dim3(1), dim3(1024), 0, st.s, k, uint32_t(n), o);
  NFH_LAUNCH_CHECK();
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

static int32_t sel_grad(const float* params, const float* X, const int32_t* y, const float* T,
                        size_t n, float* grad_out, float* loss_out) {
  if (n == 0 || !params || !X || (!y && !T) || !grad_out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const float* P = st.in(params, S_P, S_IN0);
  const float* dx = st.in(X, n * SF, S_IN1);
  const int32_t* dy = y ? st.in(y, n, S_IN2) : nullptr;
  const float* dt = T ? st.in(T, n * SC, S_IN3) : nullptr;
  float* g = st.out(grad_out, S_P, S_OUT0);
  float* l = loss_out ? st.out(loss_out, 1, S_OUT1) : nullptr;
  const uint32_t nb = std::min<uint32_t>(SG_GRID, uint32_t((n + SG_BLOCK - 1) / SG_BLOCK));
  float* part = scratch_t<float>(S_WORK0, size_t(nb) * (S_P + 1));
  hipLaunchKernelGGL(k_sel_grad, dim3(nb), dim3(SG_BLOCK), 0, st.s, P, dx, dy, dt, uint32_t(n),
                     1.f / float(n), part);
  NFH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sel_adam, dim3(1), dim3(256), 0, st.s, nullptr, nullptr, nullptr, part, nb,
                     0.f, 1.f, 1.f, g, l);
  NFH_LAUNCH_CHECK();
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

static int32_t sel_train(const float* X, const int32_t* y, const float* T, size_t n,
                         uint32_t steps, float lr, float* params_inout, float* loss_out) {
  if (n == 0 || !params_inout || !X || (!y && !T)) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const float* dx = st.in(X, n * SF, S_IN1);
  const int32_t* dy = y ? st.in(y, n, S_IN2) : nullptr;
  const float* dt = T ? st.in(T, n * SC, S_IN3) : nullptr;
  float* P = st.out(params_inout, S_P, S_OUT0, /*copy_in=*/true);
  float* mm = scratch_t<float>(S_WORK1, 2 * S_P + 8);
  fill_u32(mm, 0, 2 * S_P, st.s);
  float* l = loss_out ? st.out(loss_out, 1, S_OUT1) : nullptr;
  const uint32_t nb = std::min<uint32_t>(SG_GRID, uint32_t((n + SG_BLOCK - 1) / SG_BLOCK));
  float* part = scratch_t<float>(S_WORK0, size_t(nb) * (S_P + 1));
  double b1 = 1.0, b2 = 1.0;
  for (uint32_t it = 0; it < steps; ++it) {
    b1 *= 0.9;
    b2 *= 0.999;
    hipLaunchKernelGGL(k_sel_grad, dim3(nb), dim3(SG_BLOCK), 0, st.s, P, dx, dy, dt,
                       uint32_t(n), 1.f / float(n), part);
    NFH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sel_adam, dim3(1), dim3(256), 0, st.s, P, mm, mm + S_P, part, nb, lr,
                       float(1.0 - b1), float(1.0 - b2), nullptr, l);
    NFH_LAUNCH_CHECK();
  }
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

static int32_t sel_predict(const float* params, const float* X, size_t n, bool lowest,
                           int32_t* cls_out) {
  if (n == 0) return OK;
  if (!params || !X || !cls_out) return NULL_OR_ABORT;
  if (n > size_t(INT32_MAX)) return MALFORMED;
  NFH_GUARD_BEGIN
  Stage st;
  const float* P = st.in(params, S_P, S_IN0);
  const float* dx = st.in(X, n * SF, S_IN1);
  int32_t* c = st.out(cls_out, n, S_OUT0);
  hipLaunchKernelGGL(k_sel_predict, dim3((n + 255) / 256), dim3(256), 0, st.s, P, dx,
                     uint32_t(n), lowest, c);
  NFH_LAUNCH_CHECK();
  st.finish();
  return OK;
  NFH_GUARD_END(int32_t)
}

// Mean cross-entropy gradient of the selector at params[4420] over
// (X[n][64], y[n]) by the training kernels (grad_out[4420]).
NF_EXPORT int32_t nfh_selector_grad(const float* params, const float* X, const int32_t* y,
                                    size_t n, float* grad_out, float* loss_out) {
  return sel_grad(params, X, y, nullptr, n, grad_out, loss_out);
}

// Train the selector in place: `steps` full-batch Adam steps (learning
// rate lr) from params_inout[4420] on (X[n][64], y[n] in 0..3).
// *loss_out: the loss before the last step.
NF_EXPORT int32_t nfh_selector_train(const float* X, const int32_t* y, size_t n, uint32_t steps,
                                     float lr, float* params_inout, float* loss_out) {
  return sel_train(X, y, nullptr, n, steps, lr, params_inout, loss_out);
}

// Predicted class (0..3) per sample of X[n][64].
NF_EXPORT int32_t nfh_selector_predict(const float* params, const float* X, size_t n,
                                       int32_t* cls_out) {
  return sel_predict(params, X, n, false, cls_out);
}

// Cost model: gradient of the mean 1/2 sum_c (z_c - T_c)^2 over
// (X[n][64], T[n][4]), T = each strategy's centred log time.
NF_EXPORT int32_t nfh_selector_grad_cost(const float* params, const float* X, const float* T,
                                         size_t n, float* grad_out, float* loss_out) {
  return sel_grad(params, X, nullptr, T, n, grad_out, loss_out);
}

// Cost model: train in place (as nfh_selector_train) on T[n][4].
NF_EXPORT int32_t nfh_selector_train_cost(const float* X, const float* T, size_t n,
                                          uint32_t steps, float lr, float* params_inout,
                                          float* loss_out) {
  return sel_train(X, nullptr, T, n, steps, lr, params_inout, loss_out);
}

// Cost model: the strategy class with the smallest predicted cost.
NF_EXPORT int32_t nfh_selector_predict_cost(const float* params, const float* X, size_t n,
                                            int32_t* cls_out) {
  return sel_predict(params, X, n, true, cls_out);
}
