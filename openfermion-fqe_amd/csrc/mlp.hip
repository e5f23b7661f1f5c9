// Learned correction of the Davidson diagonal preconditioner: a 3-layer
// MLP (10 -> 128 -> 128 -> 1, tanh) evaluated per determinant, with its
// training kernels (forward with saved activations, back-propagation,
// weight-gradient reduction, Adam).  FP32 throughout; the weights of a
// workgroup live in LDS (72 KB) and every thread carries one sample, its
// first hidden layer in registers.
//
// Parameter vector: W1 [128][10], b1 [128], W2 [128][128], b2 [128],
// w3 [128], b3 [1] (18049 floats).
//
// Features of determinant I = (Ia, Ib) for Ritz value theta (d = <I|H|I>,
// dmin = min d, ref = reference strings, HOMO / LUMO of the references):
//   d - theta, 1/(|d - theta| + 0.05), d - dmin, theta - dmin,
//   popcount(sa & ~ref_a), popcount(sb & ~ref_b),
//   sa[homo_a], sa[lumo_a], sb[homo_b], sb[lumo_b].
#include "fqehip_common.h"

#include <cmath>

namespace fqehip {
namespace {

constexpr int kF = FQEHIP_MLP_NFEAT;
constexpr int kW = FQEHIP_MLP_WIDTH;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kW * kF;
constexpr int kOffW2 = kOffB1 + kW;
constexpr int kOffB2 = kOffW2 + kW * kW;
constexpr int kOffW3 = kOffB2 + kW;
constexpr int kOffB3 = kOffW3 + kW;
constexpr int kNParams = kOffB3 + 1;
static_assert(kNParams == FQEHIP_MLP_NPARAMS, "parameter layout");

constexpr int kT = 256;

struct FeatureArgs {
  int norb;
  int64_t lena, lenb;
  const int64_t* astr;
  const int64_t* bstr;
  const double* diag;
  double theta, dmin;
  int64_t ref_a, ref_b;
  int homo_a, lumo_a, homo_b, lumo_b;
};

__device__ __forceinline__ void features_of(const FeatureArgs& f, int64_t idx,
                                            float* x) {
  const int64_t ia = idx / f.lenb;
  const int64_t ib = idx - ia * f.lenb;
  const uint64_t sa = static_cast<uint64_t>(f.astr[ia]);
  const uint64_t sb = static_cast<uint64_t>(f.bstr[ib]);
  const double d = f.diag[idx];
  const double delta = d - f.theta;
  x[0] = static_cast<float>(delta);
  x[1] = static_cast<float>(1.0 / (fabs(delta) + 0.05));
  x[2] = static_cast<float>(d - f.dmin);
  x[3] = static_cast<float>(f.theta - f.dmin);
  x[4] = static_cast<float>(__popcll(sa & ~static_cast<uint64_t>(f.ref_a)));
  x[5] = static_cast<float>(__popcll(sb & ~static_cast<uint64_t>(f.ref_b)));
  x[6] = f.homo_a >= 0 ? static_cast<float>((sa >> f.homo_a) & 1) : 0.0f;
  x[7] = f.lumo_a >= 0 ? static_cast<float>((sa >> f.lumo_a) & 1) : 0.0f;
  x[8] = f.homo_b >= 0 ? static_cast<float>((sb >> f.homo_b) & 1) : 0.0f;
  x[9] = f.lumo_b >= 0 ? static_cast<float>((sb >> f.lumo_b) & 1) : 0.0f;
}

__device__ __forceinline__ void load_params(const float* __restrict__ p,
                                            float* sp) {
  for (int i = threadIdx.x; i < kNParams; i += kT) sp[i] = p[i];
  __syncthreads();
}

// y = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3 for one sample; h1 / h2
// (if non-null) receive the hidden activations at stride `ld`.
__device__ __forceinline__ float mlp_eval(const float* sp, const float* x,
                                          float* h1out, float* h2out,
                                          int64_t ld) {
  float h1[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) {
    float a = sp[kOffB1 + j];
#pragma unroll
    for (int f = 0; f < kF; ++f) a += sp[kOffW1 + j * kF + f] * x[f];
    h1[j] = tanhf(a);
    if (h1out) h1out[j * ld] = h1[j];
  }
  float y = sp[kOffB3];
#pragma unroll 1
  for (int j = 0; j < kW; ++j) {
    float a = sp[kOffB2 + j];
    const float4* w = reinterpret_cast<const float4*>(sp + kOffW2 + j * kW);
#pragma unroll
    for (int k = 0; k < kW / 4; ++k) {
      const float4 wv = w[k];
      a += wv.x * h1[4 * k] + wv.y * h1[4 * k + 1] + wv.z * h1[4 * k + 2] +
           wv.w * h1[4 * k + 3];
    }
    const float h2 = tanhf(a);
    if (h2out) h2out[j * ld] = h2;
    y += sp[kOffW3 + j] * h2;
  }
  return y;
}

// Features of selected determinants idx[s] (every determinant, s itself,
// when idx is NULL): x[s][f].
__global__ void __launch_bounds__(kT)
k_mlp_features(FeatureArgs f, int64_t m, const int64_t* __restrict__ idx,
               float* __restrict__ x) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  if (s >= m) return;
  float v[kF];
  features_of(f, idx ? idx[s] : s, v);
#pragma unroll
  for (int q = 0; q < kF; ++q) x[s * kF + q] = v[q];
}

// Training forward: y[s], h1[j][s], h2[j][s] (activations transposed so
// the per-sample stores of a wave are coalesced).
__global__ void __launch_bounds__(kT)
k_mlp_forward(int64_t m, const float* __restrict__ params,
              const float* __restrict__ x, float* __restrict__ y,
              float* __restrict__ h1, float* __restrict__ h2) {
  extern __shared__ __align__(16) float sp[];
  load_params(params, sp);
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  if (s >= m) return;
  float xv[kF];
#pragma unroll
  for (int q = 0; q < kF; ++q) xv[q] = x[s * kF + q];
  y[s] = mlp_eval(sp, xv, h1 ? h1 + s : nullptr, h2 ? h2 + s : nullptr, m);
}

// Per-sample back-propagation of L = sum_s w_s (y_s - t_s)^2:
//   d3_s = 2 w_s (y_s - t_s);  d2[j][s] = d3_s w3_j (1 - h2_js^2);
//   d1[k][s] = (1 - h1_ks^2) sum_j W2[j][k] d2[j][s].
// The loss terms go to lossv[s].
__global__ void __launch_bounds__(kT)
k_mlp_backward(int64_t m, const float* __restrict__ params,
               const float* __restrict__ y, const float* __restrict__ target,
               const float* __restrict__ weight, const float* __restrict__ h1,
               const float* __restrict__ h2, float* __restrict__ d3,
               float* __restrict__ d2, float* __restrict__ d1,
               float* __restrict__ lossv) {
  extern __shared__ __align__(16) float sp[];
  load_params(params, sp);
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  if (s >= m) return;
  const float res = y[s] - target[s];
  const float g3 = 2.0f * weight[s] * res;
  lossv[s] = weight[s] * res * res;
  d3[s] = g3;
  float g2[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) {
    const float h = h2[j * m + s];
    g2[j] = g3 * sp[kOffW3 + j] * (1.0f - h * h);
    d2[j * m + s] = g2[j];
  }
  for (int k = 0; k < kW; ++k) {
    float a = 0.0f;
#pragma unroll
    for (int j = 0; j < kW; ++j) a += sp[kOffW2 + j * kW + k] * g2[j];
    const float h = h1[k * m + s];
    d1[k * m + s] = (1.0f - h * h) * a;
  }
}

// One thread per parameter: its gradient summed over the m samples in a
// fixed order (deterministic); grad[kNParams] gets the loss sum.
__global__ void __launch_bounds__(kT)
k_mlp_grad(int64_t m, const float* __restrict__ x, const float* __restrict__ h1,
           const float* __restrict__ h2, const float* __restrict__ d3,
           const float* __restrict__ d2, const float* __restrict__ d1,
           const float* __restrict__ lossv, float* __restrict__ grad) {
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q > kNParams) return;
  double acc = 0.0;
  if (q < kOffB1) {                      // W1[j][f]
    const int j = q / kF, f = q % kF;
    for (int64_t s = 0; s < m; ++s) acc += d1[j * m + s] * x[s * kF + f];
  } else if (q < kOffW2) {               // b1[j]
    const int j = q - kOffB1;
    for (int64_t s = 0; s < m; ++s) acc += d1[j * m + s];
  } else if (q < kOffB2) {               // W2[j][k]
    const int j = (q - kOffW2) / kW, k = (q - kOffW2) % kW;
    for (int64_t s = 0; s < m; ++s) acc += d2[j * m + s] * h1[k * m + s];
  } else if (q < kOffW3) {               // b2[j]
    const int j = q - kOffB2;
    for (int64_t s = 0; s < m; ++s) acc += d2[j * m + s];
  } else if (q < kOffB3) {               // w3[j]
    const int j = q - kOffW3;
    for (int64_t s = 0; s < m; ++s) acc += d3[s] * h2[j * m + s];
  } else if (q == kOffB3) {              // b3
    for (int64_t s = 0; s < m; ++s) acc += d3[s];
  } else {                               // loss
    for (int64_t s = 0; s < m; ++s) acc += lossv[s];
  }
  grad[q] = static_cast<float>(acc);
}

__global__ void __launch_bounds__(kT)
k_mlp_adam(int n, float* __restrict__ p, const float* __restrict__ g,
           float* __restrict__ mo, float* __restrict__ vo, float lr, float b1,
           float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const float m = b1 * mo[i] + (1.0f - b1) * g[i];
  const float v = b2 * vo[i] + (1.0f - b2) * g[i] * g[i];
  mo[i] = m;
  vo[i] = v;
  p[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

constexpr size_t kParamLds = ((kNParams + 3) / 4) * 4 * sizeof(float);

int homo_of(int64_t ref, int norb) {
  int h = -1;
  for (int i = 0; i < norb; ++i)
    if ((ref >> i) & 1) h = i;
  return h;
}

int lumo_of(int64_t ref, int norb) {
  for (int i = 0; i < norb; ++i)
    if (!((ref >> i) & 1)) return i;
  return -1;
}

FeatureArgs make_features(int norb, int64_t lena, int64_t lenb,
                          const int64_t* astr, const int64_t* bstr,
                          const double* diag, double theta, double dmin,
                          int64_t ref_a, int64_t ref_b) {
  FeatureArgs f;
  f.norb = norb;
  f.lena = lena;
  f.lenb = lenb;
  f.astr = astr;
  f.bstr = bstr;
  f.diag = diag;
  f.theta = theta;
  f.dmin = dmin;
  f.ref_a = ref_a;
  f.ref_b = ref_b;
  f.homo_a = homo_of(ref_a, norb);
  f.lumo_a = lumo_of(ref_a, norb);
  f.homo_b = homo_of(ref_b, norb);
  f.lumo_b = lumo_of(ref_b, norb);
  return f;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_mlp_features(int norb, int64_t lena, int64_t lenb,
                                   const int64_t* d_astr,
                                   const int64_t* d_bstr,
                                   const double* d_diag, double theta,
                                   double dmin, int64_t ref_a, int64_t ref_b,
                                   int64_t m, const int64_t* d_idx,
                                   float* d_x, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(norb >= 0 && norb <= kMaxNorb && lena >= 0 && lenb >= 0 &&
                       m >= 0, "bad sizes");
  if (m == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_astr && d_bstr && d_diag && d_idx && d_x, "NULL pointer");
  const FeatureArgs f = make_features(norb, lena, lenb, d_astr, d_bstr, d_diag,
                                      theta, dmin, ref_a, ref_b);
  k_mlp_features<<<blocks_for(m, kT), kT, 0, as_stream(stream)>>>(f, m, d_idx,
                                                                  d_x);
  FQEHIP_CHECK_LAUNCH("k_mlp_features");
  return FQEHIP_OK;
}

extern "C" int fqehip_mlp_correction(const float* d_params, int norb,
                                     int64_t lena, int64_t lenb,
                                     const int64_t* d_astr,
                                     const int64_t* d_bstr,
                                     const double* d_diag, double theta,
                                     double dmin, int64_t ref_a,
                                     int64_t ref_b, float* d_xwork,
                                     float* d_corr, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(norb >= 0 && norb <= kMaxNorb && lena >= 0 && lenb >= 0,
                   "bad sizes");
  const int64_t n = lena * lenb;
  if (n == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_params && d_astr && d_bstr && d_diag && d_xwork && d_corr,
                   "NULL pointer");
  const FeatureArgs f = make_features(norb, lena, lenb, d_astr, d_bstr, d_diag,
                                      theta, dmin, ref_a, ref_b);
  hipStream_t st = as_stream(stream);
  k_mlp_features<<<blocks_for(n, kT), kT, 0, st>>>(f, n, nullptr, d_xwork);
  FQEHIP_CHECK_LAUNCH("k_mlp_features");
  k_mlp_forward<<<blocks_for(n, kT), kT, kParamLds, st>>>(
      n, d_params, d_xwork, d_corr, nullptr, nullptr);
  FQEHIP_CHECK_LAUNCH("k_mlp_forward");
  return FQEHIP_OK;
}

extern "C" int64_t fqehip_mlp_train_workspace_floats(int64_t m) {
  return m <= 0 ? 0 : m * (4 * static_cast<int64_t>(kW) + 3);
}

extern "C" int fqehip_mlp_forward(const float* d_params, int64_t m,
                                  const float* d_x, float* d_y, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(m >= 0, "m=%lld", static_cast<long long>(m));
  if (m == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_params && d_x && d_y, "NULL pointer");
  k_mlp_forward<<<blocks_for(m, kT), kT, kParamLds, as_stream(stream)>>>(
      m, d_params, d_x, d_y, nullptr, nullptr);
  FQEHIP_CHECK_LAUNCH("k_mlp_forward");
  return FQEHIP_OK;
}

extern "C" int fqehip_mlp_loss_grad(const float* d_params, int64_t m,
                                    const float* d_x, const float* d_target,
                                    const float* d_weight, float* d_grad,
                                    float* d_workspace, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(m > 0, "m=%lld", static_cast<long long>(m));
  FQEHIP_CHECK_ARG(d_params && d_x && d_target && d_weight && d_grad &&
                       d_workspace, "NULL pointer");
  hipStream_t st = as_stream(stream);
  float* h1 = d_workspace;
  float* h2 = h1 + kW * m;
  float* d2 = h2 + kW * m;
  float* d1 = d2 + kW * m;
  float* y = d1 + kW * m;
  float* d3 = y + m;
  float* lossv = d3 + m;
  const unsigned nb = blocks_for(m, kT);
  k_mlp_forward<<<nb, kT, kParamLds, st>>>(m, d_params, d_x, y, h1, h2);
  FQEHIP_CHECK_LAUNCH("k_mlp_forward");
  k_mlp_backward<<<nb, kT, kParamLds, st>>>(m, d_params, y, d_target, d_weight,
                                            h1, h2, d3, d2, d1, lossv);
  FQEHIP_CHECK_LAUNCH("k_mlp_backward");
  k_mlp_grad<<<blocks_for(kNParams + 1, kT), kT, 0, st>>>(m, d_x, h1, h2, d3,
                                                          d2, d1, lossv, d_grad);
  FQEHIP_CHECK_LAUNCH("k_mlp_grad");
  return FQEHIP_OK;
}

extern "C" int fqehip_mlp_adam(float* d_params, const float* d_grad,
                               float* d_m, float* d_v, int step, float lr,
                               float beta1, float beta2, float eps,
                               void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(step >= 1 && lr > 0.0f, "step=%d lr=%g", step, lr);
  FQEHIP_CHECK_ARG(d_params && d_grad && d_m && d_v, "NULL pointer");
  const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
  const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
  k_mlp_adam<<<blocks_for(kNParams, kT), kT, 0, as_stream(stream)>>>(
      kNParams, d_params, d_grad, d_m, d_v, lr, beta1, beta2, eps, c1, c2);
  FQEHIP_CHECK_LAUNCH("k_mlp_adam");
  return FQEHIP_OK;
}
