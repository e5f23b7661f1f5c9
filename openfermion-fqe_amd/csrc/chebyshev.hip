// Step of the Chebyshev propagator's three-term recurrence on the device.
//
// The apply (fqehip_apply with alpha = 2/a, beta = -1, cin = T_{k-2}) writes
// (2/a) H T_{k-1} - T_{k-2} into T_k; this kernel finishes the term and
// folds it into the series in the same pass:
//   T_k += gamma T_{k-1}           (the spectral shift, gamma = 2 s / a)
//   result += c_k T_k              (FP64, complex128)
//   partials[block] = sum |T_k|^2  (FP64, fixed order, for the stopping test)
// It does nothing once the device flag *done is set, like the apply kernels,
// so a whole series is enqueued without host synchronisation.
#include "fqehip_common.h"

namespace fqehip {
namespace {

constexpr int kChebBlock = 256;   // same blocks as the apply's norm partials

template <typename T>
__global__ void __launch_bounds__(kChebBlock)
k_cheb_update(int64_t n, T* __restrict__ tk, const T* __restrict__ tkm1,
              double gr, double gi, double cr, double ci,
              double* __restrict__ result, double* __restrict__ partials,
              const int32_t* __restrict__ done) {
  __shared__ double red[kChebBlock / 64];
  if (done && *done) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kChebBlock + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    double xr = static_cast<double>(tk[2 * i]);
    double xi = static_cast<double>(tk[2 * i + 1]);
    if (tkm1) {
      const double pr = static_cast<double>(tkm1[2 * i]);
      const double pi = static_cast<double>(tkm1[2 * i + 1]);
      xr += gr * pr - gi * pi;
      xi += gr * pi + gi * pr;
    }
    const T sr = static_cast<T>(xr), si = static_cast<T>(xi);
    tk[2 * i] = sr;
    tk[2 * i + 1] = si;
    xr = static_cast<double>(sr);
    xi = static_cast<double>(si);
    result[2 * i] += cr * xr - ci * xi;
    result[2 * i + 1] += cr * xi + ci * xr;
    sq = xr * xr + xi * xi;
  }
  if (!partials) return;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kChebBlock / 64; ++w) s += red[w];
    partials[blockIdx.x] = s;
  }
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_cheb_update(int precision, int64_t n, void* d_tk,
                                  const void* d_tkm1, double gamma_re,
                                  double gamma_im, double c_re, double c_im,
                                  double* d_result, double* d_norm_partials,
                                  const int32_t* d_done, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(precision == FQEHIP_C128 || precision == FQEHIP_C64,
                   "precision=%d", precision);
  FQEHIP_CHECK_ARG(n >= 0, "n=%lld", static_cast<long long>(n));
  FQEHIP_CHECK_ARG(n == 0 || (d_tk && d_result), "NULL term or result");
  FQEHIP_CHECK_ARG(d_tk != d_tkm1 || n == 0, "T_k aliases T_(k-1)");
  if (n == 0) return FQEHIP_OK;
  hipStream_t st = as_stream(stream);
  const unsigned nb = blocks_for(n, kChebBlock);
  if (precision == FQEHIP_C128)
    k_cheb_update<double><<<nb, kChebBlock, 0, st>>>(
        n, static_cast<double*>(d_tk), static_cast<const double*>(d_tkm1),
        gamma_re, gamma_im, c_re, c_im, d_result, d_norm_partials, d_done);
  else
    k_cheb_update<float><<<nb, kChebBlock, 0, st>>>(
        n, static_cast<float*>(d_tk), static_cast<const float*>(d_tkm1),
        gamma_re, gamma_im, c_re, c_im, d_result, d_norm_partials, d_done);
  FQEHIP_CHECK_LAUNCH("k_cheb_update");
  return FQEHIP_OK;
}
