// (b) dvec gather, (c) pair GEMM on the matrix cores, (d) gather-form
// scatter with the Taylor epilogue, the one-body apply, norm reduction and
// the composed apply / Taylor entry points.
//
// The apply follows the dvec form of the operator
//   H C = sum_p h1'[p] D[p] + sum_p E_p Y[p],   D[p] = E_p C,   Y = V D
// (reference src/fqe/fqe_data.py:646-683).  Both D = E C and the scatter
// sum_p E_p Y[p] are gathers at the target determinant through the same
// pair table (entry (p, I) = sign * (source + 1) with <I|E_p|source> =
// sign), so no kernel needs atomics.  The one-body term is the last row of
// the GEMM's A matrix and lands in plane npair of Y.
#include "fqehip_common.h"

#include <cmath>

namespace fqehip {
namespace {

constexpr int kBlock = 256;

typedef double double4_t __attribute__((ext_vector_type(4)));

// idx -> (row, column) of a lena x lenb matrix without a 64-bit integer
// division: a double estimate corrected by one step.
__device__ __forceinline__ void split_index(int64_t idx, int64_t lenb,
                                            double inv_lenb, int64_t& ia,
                                            int64_t& ib) {
  ia = static_cast<int64_t>(static_cast<double>(idx) * inv_lenb);
  ib = idx - ia * lenb;
  if (ib < 0) {
    --ia;
    ib += lenb;
  } else if (ib >= lenb) {
    ++ia;
    ib -= lenb;
  }
}

template <typename T>
__device__ __forceinline__ T signed_val(int32_t code, T v) {
  return code > 0 ? v : -v;
}

template <typename T> struct vec2_of;
template <> struct vec2_of<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <> struct vec2_of<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <typename T> struct vec4_of;
template <> struct vec4_of<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};
template <> struct vec4_of<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};

// ---------------------------------------------------------------------------
// (b) dvec gather: D[p][re|im][I] for one plane p per blockIdx.y.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dvec_gather(int64_t lena, int64_t lenb, double inv_lenb,
              const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
              const T* __restrict__ coeff, T* __restrict__ dvec,
              const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t n = lena * lenb;
  const int p = blockIdx.y;
  const int32_t* pa = pta + static_cast<int64_t>(p) * lena;
  const int32_t* pb = ptb + static_cast<int64_t>(p) * lenb;
  T* dre = dvec + static_cast<int64_t>(2 * p) * n;
  T* dim = dre + n;
  for (int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < n; idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t ia, ib;
    split_index(idx, lenb, inv_lenb, ia, ib);
    T re = 0, im = 0;
    const int32_t a = pa[ia];
    if (a) {
      const int64_t src = (a > 0 ? a : -a) - 1;
      const T* c = coeff + 2 * (src * lenb + ib);
      re += signed_val(a, c[0]);
      im += signed_val(a, c[1]);
    }
    const int32_t b = pb[ib];
    if (b) {
      const int64_t src = (b > 0 ? b : -b) - 1;
      const T* c = coeff + 2 * (ia * lenb + src);
      re += signed_val(b, c[0]);
      im += signed_val(b, c[1]);
    }
    dre[idx] = re;
    dim[idx] = im;
  }
}

// Row form of the same gather (the one the apply uses): one workgroup per
// alpha string ia writes D[p][re|im][ia, :] for every plane p.
//   - C[ia, :] is staged once in LDS (ROW_LDS; else read from the cache),
//     so the beta term sum over ptb[p][ib] = sign (src + 1) is an LDS
//     gather;
//   - the alpha term is the whole row C[src, :] of the alpha source of
//     (p, ia), one sign for the workgroup: coalesced reads, served by the
//     caches for the second and later targets that share a source;
//   - the planes are written with non-temporal stores, coalesced along ib;
//   - with lenb even a lane takes a column pair (one 32- or 16-byte read
//     of C, one 8-byte read of the table, two plane stores of 2 values).
constexpr int kRowBlock = 512;
constexpr int64_t kRowLdsBytes = 64 * 1024;

// One column pair (ib, ib + 1) of plane p for target row ia: alpha term
// from the source row src (null when the table entry is 0), beta term
// from the LDS row (or the global row own).
template <typename T, bool ROW_LDS>
__device__ __forceinline__ void dvec_pair(
    const T* __restrict__ src, T sa,
    const typename vec2_of<T>::type* __restrict__ row,
    const typename vec2_of<T>::type* __restrict__ own, int2 b, int64_t ib,
    typename vec2_of<T>::type& re, typename vec2_of<T>::type& im) {
  typedef typename vec2_of<T>::type T2;
  typedef typename vec4_of<T>::type T4;
  re = T2{0, 0};
  im = T2{0, 0};
  if (src) {
    const T4 v = *reinterpret_cast<const T4*>(src + 2 * ib);
    re = T2{sa * v[0], sa * v[2]};
    im = T2{sa * v[1], sa * v[3]};
  }
  if (b.x) {
    const int64_t j = (b.x > 0 ? b.x : -b.x) - 1;
    const T2 w = ROW_LDS ? row[j] : own[j];
    re[0] += b.x > 0 ? w[0] : -w[0];
    im[0] += b.x > 0 ? w[1] : -w[1];
  }
  if (b.y) {
    const int64_t j = (b.y > 0 ? b.y : -b.y) - 1;
    const T2 w = ROW_LDS ? row[j] : own[j];
    re[1] += b.y > 0 ? w[0] : -w[0];
    im[1] += b.y > 0 ? w[1] : -w[1];
  }
}

template <typename T, bool ROW_LDS>
__global__ void __launch_bounds__(kRowBlock)
k_dvec_rows(int npair, int64_t lena, int64_t lenb,
            const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
            const T* __restrict__ coeff, T* __restrict__ dvec,
            const int32_t* __restrict__ done) {
  typedef typename vec2_of<T>::type T2;
  extern __shared__ __align__(16) unsigned char row_raw[];
  T2* row = reinterpret_cast<T2*>(row_raw);
  if (done && *done) return;
  // targets of one XCD are a contiguous range: they share alpha sources
  const int64_t ia = xcd_order(blockIdx.x, lena);
  const int64_t n = lena * lenb;
  const T2* c2 = reinterpret_cast<const T2*>(coeff);
  const T2* own = c2 + ia * lenb;
  if (ROW_LDS) {
    for (int64_t ib = threadIdx.x; ib < lenb; ib += kRowBlock) row[ib] = own[ib];
    __syncthreads();
  }
  const int64_t base = ia * lenb;
  // lenb even: every row starts on an even element, and each lane takes a
  // column pair (16-byte plane stores in complex128, 8-byte in complex64)
  const bool pairs = (lenb & 1) == 0;
  for (int p = 0; p < npair; ++p) {
    // uniform over the workgroup
    const int32_t a = __builtin_amdgcn_readfirstlane(pta[static_cast<int64_t>(p) * lena + ia]);
    const T* src = a ? coeff + 2 * static_cast<int64_t>((a > 0 ? a : -a) - 1) * lenb
                     : nullptr;
    const T sa = a > 0 ? T(1) : T(-1);
    const int32_t* pb = ptb + static_cast<int64_t>(p) * lenb;
    T* dre = dvec + static_cast<int64_t>(2 * p) * n + base;
    T* dim = dre + n;
    if (pairs) {
      for (int64_t ib = 2 * threadIdx.x; ib < lenb; ib += 2 * kRowBlock) {
        T2 re, im;
        dvec_pair<T, ROW_LDS>(src, sa, row, own,
                              *reinterpret_cast<const int2*>(pb + ib), ib, re, im);
        __builtin_nontemporal_store(re, reinterpret_cast<T2*>(dre + ib));
        __builtin_nontemporal_store(im, reinterpret_cast<T2*>(dim + ib));
      }
      continue;
    }
    for (int64_t ib = threadIdx.x; ib < lenb; ib += kRowBlock) {
      T re = 0, im = 0;
      if (src) {
        re = sa * src[2 * ib];
        im = sa * src[2 * ib + 1];
      }
      const int32_t b = pb[ib];
      if (b) {
        const int64_t j = (b > 0 ? b : -b) - 1;
        const T2 w = ROW_LDS ? row[j] : own[j];
        re += b > 0 ? w[0] : -w[0];
        im += b > 0 ? w[1] : -w[1];
      }
      __builtin_nontemporal_store(re, dre + ib);
      __builtin_nontemporal_store(im, dim + ib);
    }
  }
}

// ---------------------------------------------------------------------------
// (c) FP64 GEMM Y = A B on v_mfma_f64_16x16x4f64.
// A: m x k (row-major, read through the cache; it is tiny and shared by all
// workgroups).  B: k x ncol, Y: m x ncol, row-major.  A wave owns one 16-
// column tile and RT 16-row tiles; a workgroup is 4 waves = 64 columns.
// MFMA operand layout (16x16x4 f64, 64 lanes): A[r = lane%16][k = lane/16],
// B[k = lane/16][c = lane%16]; result item i of lane l is
// Y[l/16 + 4*i][l%16] (rows interleaved, unlike the f32-output MFMAs).
// ---------------------------------------------------------------------------
constexpr int kGemmRowTiles = 8;

__global__ void __launch_bounds__(kBlock)
k_gemm_f64(int m, int k, const double* __restrict__ a, int lda, int64_t ncol,
           const double* __restrict__ b, double* __restrict__ y,
           int accumulate, const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t col0 = (static_cast<int64_t>(blockIdx.x) * 4 + wave) * 16;
  if (col0 >= ncol) return;
  const int row0 = blockIdx.y * kGemmRowTiles * 16;
  const int64_t col = col0 + (lane & 15);
  const bool colok = col < ncol;
  const int kl = lane >> 4;
  double4_t acc[kGemmRowTiles];
#pragma unroll
  for (int t = 0; t < kGemmRowTiles; ++t) acc[t] = double4_t{0, 0, 0, 0};
  int ntiles = (m - row0 + 15) / 16;
  if (ntiles > kGemmRowTiles) ntiles = kGemmRowTiles;
  for (int k0 = 0; k0 < k; k0 += 4) {
    const int kk = k0 + kl;
    const bool kok = kk < k;
    const double bv = (colok && kok) ? b[static_cast<int64_t>(kk) * ncol + col] : 0.0;
#pragma unroll
    for (int t = 0; t < kGemmRowTiles; ++t) {
      if (t < ntiles) {
        const int r = row0 + t * 16 + (lane & 15);
        const double av = (r < m && kok) ? a[static_cast<int64_t>(r) * lda + kk] : 0.0;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
      }
    }
  }
  if (!colok) return;
#pragma unroll
  for (int t = 0; t < kGemmRowTiles; ++t) {
    if (t < ntiles) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + t * 16 + (lane >> 4) + 4 * i;
        if (r < m) {
          double* dst = y + static_cast<int64_t>(r) * ncol + col;
          *dst = accumulate ? *dst + acc[t][i] : acc[t][i];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// (d) scatter + epilogue (see fqehip_scatter).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void epilogue(int64_t idx, double re, double im,
                                         double alpha_re, double alpha_im,
                                         double beta_re, double beta_im,
                                         const T* __restrict__ cin,
                                         T* __restrict__ out,
                                         double* __restrict__ acc,
                                         double& sq) {
  double vr = alpha_re * re - alpha_im * im;
  double vi = alpha_re * im + alpha_im * re;
  if (beta_re != 0.0 || beta_im != 0.0) {
    const double cr = static_cast<double>(cin[2 * idx]);
    const double ci = static_cast<double>(cin[2 * idx + 1]);
    vr += beta_re * cr - beta_im * ci;
    vi += beta_re * ci + beta_im * cr;
  }
  const T or_ = static_cast<T>(vr);
  const T oi = static_cast<T>(vi);
  out[2 * idx] = or_;
  out[2 * idx + 1] = oi;
  if (acc) {
    acc[2 * idx] += vr;
    acc[2 * idx + 1] += vi;
  }
  // the norm is taken of what was stored
  const double sr = static_cast<double>(or_);
  const double si = static_cast<double>(oi);
  sq += sr * sr + si * si;
}

__device__ __forceinline__ void block_sum_store(double v, double* partials) {
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    partials[blockIdx.x] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_scatter(int npair, int64_t lena, int64_t lenb, double inv_lenb,
          const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
          const T* __restrict__ y, double alpha_re, double alpha_im,
          double beta_re, double beta_im, const T* cin, T* out,
          double* __restrict__ acc, double* __restrict__ partials,
          const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t n = lena * lenb;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (idx < n) {
    int64_t ia, ib;
    split_index(idx, lenb, inv_lenb, ia, ib);
    const T* yh = y + static_cast<int64_t>(2 * npair) * n;
    double re = static_cast<double>(yh[idx]);
    double im = static_cast<double>(yh[n + idx]);
    for (int p = 0; p < npair; ++p) {
      const T* yre = y + static_cast<int64_t>(2 * p) * n;
      const int32_t a = pta[static_cast<int64_t>(p) * lena + ia];
      if (a) {
        const int64_t off = ((a > 0 ? a : -a) - 1) * lenb + ib;
        re += signed_val(a, static_cast<double>(yre[off]));
        im += signed_val(a, static_cast<double>(yre[n + off]));
      }
      const int32_t b = ptb[static_cast<int64_t>(p) * lenb + ib];
      if (b) {
        const int64_t off = ia * lenb + (b > 0 ? b : -b) - 1;
        re += signed_val(b, static_cast<double>(yre[off]));
        im += signed_val(b, static_cast<double>(yre[n + off]));
      }
    }
    epilogue(idx, re, im, alpha_re, alpha_im, beta_re, beta_im, cin, out, acc,
             sq);
  }
  if (partials) block_sum_store(sq, partials);
}

// One-body apply: out = alpha * sum_p h1[p] (E_p C) + beta * cin.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_apply1(int npair, int64_t lena, int64_t lenb, double inv_lenb,
         const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
         const double* __restrict__ h1, const T* __restrict__ coeff,
         double alpha_re, double alpha_im, double beta_re, double beta_im,
         const T* cin, T* out, double* __restrict__ acc,
         double* __restrict__ partials, const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t n = lena * lenb;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (idx < n) {
    int64_t ia, ib;
    split_index(idx, lenb, inv_lenb, ia, ib);
    double re = 0.0, im = 0.0;
    for (int p = 0; p < npair; ++p) {
      const double hr = h1[2 * p], hi = h1[2 * p + 1];
      double sr = 0.0, si = 0.0;
      const int32_t a = pta[static_cast<int64_t>(p) * lena + ia];
      if (a) {
        const int64_t off = 2 * (((a > 0 ? a : -a) - 1) * lenb + ib);
        sr += signed_val(a, static_cast<double>(coeff[off]));
        si += signed_val(a, static_cast<double>(coeff[off + 1]));
      }
      const int32_t b = ptb[static_cast<int64_t>(p) * lenb + ib];
      if (b) {
        const int64_t off = 2 * (ia * lenb + (b > 0 ? b : -b) - 1);
        sr += signed_val(b, static_cast<double>(coeff[off]));
        si += signed_val(b, static_cast<double>(coeff[off + 1]));
      }
      re += hr * sr - hi * si;
      im += hr * si + hi * sr;
    }
    epilogue(idx, re, im, alpha_re, alpha_im, beta_re, beta_im, cin, out, acc,
             sq);
  }
  if (partials) block_sum_store(sq, partials);
}

// Chunked apply (fqehip_apply_panels): the scatter of one panel of output
// pairs [p0, p0 + np) into an FP64 accumulator, oacc (+)= sum_p E_p Y[p];
// the h1' row (p == npair) is added without excitation.  Y planes of the
// panel: 2n values per pair (re | im halves), as in the full Y.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_scatter_panel(int p0, int np, int npair, int64_t lena, int64_t lenb,
                double inv_lenb, const int32_t* __restrict__ pta,
                const int32_t* __restrict__ ptb, const T* __restrict__ y,
                double* __restrict__ oacc, int first,
                const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t n = lena * lenb;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int64_t ia, ib;
  split_index(idx, lenb, inv_lenb, ia, ib);
  double re = first ? 0.0 : oacc[2 * idx];
  double im = first ? 0.0 : oacc[2 * idx + 1];
  for (int j = 0; j < np; ++j) {
    const int p = p0 + j;
    const T* yre = y + static_cast<int64_t>(2 * j) * n;
    if (p == npair) {
      re += static_cast<double>(yre[idx]);
      im += static_cast<double>(yre[n + idx]);
      continue;
    }
    const int32_t a = pta[static_cast<int64_t>(p) * lena + ia];
    if (a) {
      const int64_t off = ((a > 0 ? a : -a) - 1) * lenb + ib;
      re += signed_val(a, static_cast<double>(yre[off]));
      im += signed_val(a, static_cast<double>(yre[n + off]));
    }
    const int32_t b = ptb[static_cast<int64_t>(p) * lenb + ib];
    if (b) {
      const int64_t off = ia * lenb + (b > 0 ? b : -b) - 1;
      re += signed_val(b, static_cast<double>(yre[off]));
      im += signed_val(b, static_cast<double>(yre[n + off]));
    }
  }
  oacc[2 * idx] = re;
  oacc[2 * idx + 1] = im;
}

// Final pass of the chunked apply: out = alpha oacc + beta cin, with the
// epilogue of fqehip_scatter (acc, norm partials).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_panel_epilogue(int64_t n, const double* __restrict__ oacc, double alpha_re,
                 double alpha_im, double beta_re, double beta_im, const T* cin,
                 T* out, double* __restrict__ acc, double* __restrict__ partials,
                 const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (idx < n)
    epilogue(idx, oacc[2 * idx], oacc[2 * idx + 1], alpha_re, alpha_im,
             beta_re, beta_im, cin, out, acc, sq);
  if (partials) block_sum_store(sq, partials);
}

// Single-workgroup reduction in a fixed order; optional convergence test.
__global__ void __launch_bounds__(kBlock)
k_norm_finalize(const double* __restrict__ partials, int64_t np,
                double* __restrict__ norm_out, double accuracy, int term,
                int32_t* done, int32_t* nterms) {
  __shared__ double red[kBlock];
  if (done && *done) return;
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < np; k += kBlock) s += partials[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nrm = sqrt(red[0]);
    norm_out[0] = nrm;
    if (done) {
      nterms[0] = term;
      if (nrm < accuracy) *done = 1;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_from_c128(int64_t n2, const double* __restrict__ in, T* __restrict__ out) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       k < n2; k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[k] = static_cast<T>(in[k]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_to_c128(int64_t n, const T* __restrict__ in, double sr, double si,
          double* __restrict__ out) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       k < n; k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double a = static_cast<double>(in[2 * k]);
    const double b = static_cast<double>(in[2 * k + 1]);
    out[2 * k] = sr * a - si * b;
    out[2 * k + 1] = sr * b + si * a;
  }
}

__global__ void k_status_init(int32_t* status) {
  if (threadIdx.x == 0) {
    status[0] = 0;
    status[1] = 0;
  }
}

}  // namespace

// Defined in pair_gemm.hip.
bool pair_gemm_lds_ok(int m, int k, int64_t ncol, const void* b,
                      const void* y);
int launch_pair_gemm_lds(int m, int k, const double* a, int lda, int64_t ncol,
                         const double* b, double* y, int accumulate,
                         hipStream_t st, const int32_t* done);

namespace {

unsigned grid_stride_blocks(int64_t n) {
  const unsigned b = blocks_for(n, kBlock);
  return b < 65536u ? (b == 0 ? 1u : b) : 65536u;
}

int check_precision(int precision) {
  FQEHIP_CHECK_ARG(precision == FQEHIP_C128 || precision == FQEHIP_C64,
                   "unknown precision %d", precision);
  return FQEHIP_OK;
}

template <typename T>
int launch_gather(int npair, int64_t lena, int64_t lenb, const int32_t* pta,
                  const int32_t* ptb, const void* coeff, void* dvec,
                  hipStream_t st, const int32_t* done) {
  const int64_t n = lena * lenb;
  if (n == 0 || npair == 0) return FQEHIP_OK;
  if (lena <= 0x7fffffff) {
    const int64_t row = lenb * 2 * static_cast<int64_t>(sizeof(T));
    if (row <= kRowLdsBytes)
      k_dvec_rows<T, true><<<static_cast<unsigned>(lena), kRowBlock, row, st>>>(
          npair, lena, lenb, pta, ptb, static_cast<const T*>(coeff),
          static_cast<T*>(dvec), done);
    else
      k_dvec_rows<T, false><<<static_cast<unsigned>(lena), kRowBlock, 0, st>>>(
          npair, lena, lenb, pta, ptb, static_cast<const T*>(coeff),
          static_cast<T*>(dvec), done);
    FQEHIP_CHECK_LAUNCH("k_dvec_rows");
    return FQEHIP_OK;
  }
  unsigned bx = blocks_for(n, kBlock);
  if (bx > 8192u) bx = 8192u;
  dim3 grid(bx, npair);
  k_dvec_gather<T><<<grid, kBlock, 0, st>>>(lena, lenb, 1.0 / lenb, pta, ptb,
                                           static_cast<const T*>(coeff),
                                           static_cast<T*>(dvec), done);
  FQEHIP_CHECK_LAUNCH("k_dvec_gather");
  return FQEHIP_OK;
}

template <typename T>
int launch_scatter(int npair, int64_t lena, int64_t lenb, const int32_t* pta,
                   const int32_t* ptb, const void* y, double ar, double ai,
                   double br, double bi, const void* cin, void* out,
                   double* acc, double* partials, const int32_t* done,
                   hipStream_t st) {
  const int64_t n = lena * lenb;
  if (n == 0) return FQEHIP_OK;
  k_scatter<T><<<blocks_for(n, kBlock), kBlock, 0, st>>>(
      npair, lena, lenb, 1.0 / lenb, pta, ptb, static_cast<const T*>(y), ar,
      ai, br, bi, static_cast<const T*>(cin), static_cast<T*>(out), acc,
      partials, done);
  FQEHIP_CHECK_LAUNCH("k_scatter");
  return FQEHIP_OK;
}

template <typename T>
int launch_apply1(int npair, int64_t lena, int64_t lenb, const int32_t* pta,
                  const int32_t* ptb, const double* h1, const void* coeff,
                  double ar, double ai, double br, double bi, const void* cin,
                  void* out, double* acc, double* partials,
                  const int32_t* done, hipStream_t st) {
  const int64_t n = lena * lenb;
  if (n == 0) return FQEHIP_OK;
  k_apply1<T><<<blocks_for(n, kBlock), kBlock, 0, st>>>(
      npair, lena, lenb, 1.0 / lenb, pta, ptb, h1,
      static_cast<const T*>(coeff), ar, ai, br, bi,
      static_cast<const T*>(cin), static_cast<T*>(out), acc, partials, done);
  FQEHIP_CHECK_LAUNCH("k_apply1");
  return FQEHIP_OK;
}

int launch_gemm_f64(int m, int k, const double* a, int64_t ncol,
                    const double* b, double* y, hipStream_t st,
                    const int32_t* done, int lda = -1, int accumulate = 0) {
  if (m == 0 || ncol == 0) return FQEHIP_OK;
  if (pair_gemm_lds_ok(m, k, ncol, b, y))
    return launch_pair_gemm_lds(m, k, a, lda < 0 ? k : lda, ncol, b, y,
                                accumulate, st, done);
  dim3 grid(blocks_for(ncol, 64), (m + kGemmRowTiles * 16 - 1) / (kGemmRowTiles * 16));
  k_gemm_f64<<<grid, kBlock, 0, st>>>(m, k, a, lda < 0 ? k : lda, ncol, b, y,
                                      accumulate, done);
  FQEHIP_CHECK_LAUNCH("k_gemm_f64");
  return FQEHIP_OK;
}

}  // namespace

// Defined in fused.hip.
bool fused_supported(int precision, int npair, int64_t lena, int64_t lenb);
bool split_supported(int npair, int64_t lena, int64_t lenb);
bool split_f64_supported(int npair, int64_t lena, int64_t lenb, int tile);
int launch_fused(int precision, int norb, int npair, int64_t lena, int64_t lenb,
                 const double* a, const int32_t* pta, const int32_t* ptb,
                 const int32_t* fwa, const int32_t* fwb, const int32_t* sla,
                 int nslot, const void* coeff, void* slots, void* rowacc,
                 bool split, int tile, hipStream_t st, const int32_t* done);
// Defined in ozaki.hip.
bool split_oz_supported(int norb, int npair, int64_t lena, int64_t lenb);
int64_t split_oz_table_bytes(int64_t lenb);
int launch_split_oz(int norb, int npair, int64_t lena, int64_t lenb,
                    const double* a, const int32_t* pta, const int32_t* ptb,
                    const int32_t* fwa, const int32_t* sla, int nslot,
                    const double* coeff, const double* rowmax, void* btab,
                    uint32_t* slots, double* rowacc, bool build_tables,
                    hipStream_t st, const int32_t* done);
int launch_row_absmax(int64_t lena, int64_t lenb, const double* coeff,
                      double* rowmax, hipStream_t st, const int32_t* done);
bool beta_oz_shape_ok(int64_t lena, int64_t lenb);
int64_t beta_oz_workspace_bytes(int64_t lena, int64_t lenb);
int launch_beta_gemm_oz(int64_t lena, int64_t lenb, const void* hoz,
                        const double* coeff, const double* rowmax,
                        void* cwork, double* out, hipStream_t st,
                        const int32_t* done);
// Defined in beta_block.hip.
bool beta_gemm_shape_ok(int64_t lena, int64_t lenb, int64_t ldh);
int launch_beta_gemm(int64_t lena, int64_t lenb, int64_t ldh, const void* hhi,
                     const void* hlo, const float* coeff, float* out,
                     hipStream_t st, const int32_t* done);
bool beta_gemm_f64_shape_ok(int64_t lena, int64_t lenb);
int launch_beta_gemm_f64(int64_t lena, int64_t lenb, const double* hcol,
                         const double* coeff, double* out, hipStream_t st,
                         const int32_t* done);
int launch_slot_reduce(int precision, bool p48, int nslot, int64_t lena,
                       int64_t lenb, const void* slots, const void* rowacc,
                       double ar, double ai, double br, double bi,
                       const void* cin, void* out, double* acc,
                       double* partials, hipStream_t st, const int32_t* done);

// Declared in gemm_bf16.hip.
int launch_gemm_bf16x3(int m, int k, const double* a, int64_t ncol,
                       const float* b, float* y, hipStream_t st,
                       const int32_t* done);

namespace {

int launch_gemm(int precision, int m, int k, const double* a, int64_t ncol,
                const void* b, void* y, hipStream_t st, const int32_t* done) {
  if (precision == FQEHIP_C128)
    return launch_gemm_f64(m, k, a, ncol, static_cast<const double*>(b),
                           static_cast<double*>(y), st, done);
  return launch_gemm_bf16x3(m, k, a, ncol, static_cast<const float*>(b),
                            static_cast<float*>(y), st, done);
}

int launch_gather_any(int precision, int npair, int64_t lena, int64_t lenb,
                      const int32_t* pta, const int32_t* ptb,
                      const void* coeff, void* dvec, hipStream_t st,
                      const int32_t* done) {
  if (precision == FQEHIP_C128)
    return launch_gather<double>(npair, lena, lenb, pta, ptb, coeff, dvec, st,
                                 done);
  return launch_gather<float>(npair, lena, lenb, pta, ptb, coeff, dvec, st,
                              done);
}

}  // namespace

}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_dvec_gather(int precision, int npair, int64_t lena,
                                  int64_t lenb, const int32_t* d_ptab_a,
                                  const int32_t* d_ptab_b, const void* d_coeff,
                                  void* d_dvec, void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(npair >= 0 && lena >= 0 && lenb >= 0, "negative size");
  FQEHIP_CHECK_ARG(d_ptab_a && d_ptab_b && d_coeff && d_dvec, "NULL pointer");
  return launch_gather_any(precision, npair, lena, lenb, d_ptab_a, d_ptab_b,
                           d_coeff, d_dvec, as_stream(stream), nullptr);
}

extern "C" int fqehip_pair_gemm(int precision, int m, int k, const double* d_a,
                                int64_t ncol, const void* d_b, void* d_y,
                                void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(m >= 0 && k >= 0 && ncol >= 0, "negative size");
  FQEHIP_CHECK_ARG(d_a && d_b && d_y, "NULL pointer");
  return launch_gemm(precision, m, k, d_a, ncol, d_b, d_y, as_stream(stream),
                     nullptr);
}

extern "C" int64_t fqehip_norm_partials_count(int64_t n) {
  return n > 0 ? static_cast<int64_t>(blocks_for(n, kBlock)) : 0;
}

extern "C" int fqehip_scatter(int precision, int npair, int64_t lena,
                              int64_t lenb, const int32_t* d_ptab_a,
                              const int32_t* d_ptab_b, const void* d_y,
                              double alpha_re, double alpha_im, double beta_re,
                              double beta_im, const void* d_cin, void* d_out,
                              double* d_acc, double* d_norm_partials,
                              const int32_t* d_done, void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(npair >= 0 && lena >= 0 && lenb >= 0, "negative size");
  FQEHIP_CHECK_ARG(d_ptab_a && d_ptab_b && d_y && d_out, "NULL pointer");
  FQEHIP_CHECK_ARG((beta_re == 0.0 && beta_im == 0.0) || d_cin,
                   "beta != 0 needs d_cin");
  if (precision == FQEHIP_C128)
    return launch_scatter<double>(npair, lena, lenb, d_ptab_a, d_ptab_b, d_y,
                                  alpha_re, alpha_im, beta_re, beta_im, d_cin,
                                  d_out, d_acc, d_norm_partials, d_done,
                                  as_stream(stream));
  return launch_scatter<float>(npair, lena, lenb, d_ptab_a, d_ptab_b, d_y,
                               alpha_re, alpha_im, beta_re, beta_im, d_cin,
                               d_out, d_acc, d_norm_partials, d_done,
                               as_stream(stream));
}

extern "C" int fqehip_apply1(int precision, int npair, int64_t lena,
                             int64_t lenb, const int32_t* d_ptab_a,
                             const int32_t* d_ptab_b, const double* d_h1,
                             const void* d_coeff, double alpha_re,
                             double alpha_im, double beta_re, double beta_im,
                             const void* d_cin, void* d_out, double* d_acc,
                             double* d_norm_partials, const int32_t* d_done,
                             void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(npair >= 0 && lena >= 0 && lenb >= 0, "negative size");
  FQEHIP_CHECK_ARG(d_ptab_a && d_ptab_b && d_h1 && d_coeff && d_out,
                   "NULL pointer");
  FQEHIP_CHECK_ARG(d_coeff != d_out, "apply1 cannot run in place");
  FQEHIP_CHECK_ARG((beta_re == 0.0 && beta_im == 0.0) || d_cin,
                   "beta != 0 needs d_cin");
  if (precision == FQEHIP_C128)
    return launch_apply1<double>(npair, lena, lenb, d_ptab_a, d_ptab_b, d_h1,
                                 d_coeff, alpha_re, alpha_im, beta_re, beta_im,
                                 d_cin, d_out, d_acc, d_norm_partials, d_done,
                                 as_stream(stream));
  return launch_apply1<float>(npair, lena, lenb, d_ptab_a, d_ptab_b, d_h1,
                              d_coeff, alpha_re, alpha_im, beta_re, beta_im,
                              d_cin, d_out, d_acc, d_norm_partials, d_done,
                              as_stream(stream));
}

extern "C" int fqehip_norm_finalize(const double* d_partials, int64_t npartials,
                                    double* d_norm_out, double accuracy,
                                    int term, int32_t* d_done,
                                    int32_t* d_nterms, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(d_norm_out && (npartials == 0 || d_partials), "NULL pointer");
  FQEHIP_CHECK_ARG((d_done == nullptr) == (d_nterms == nullptr),
                   "d_done and d_nterms go together");
  k_norm_finalize<<<1, kBlock, 0, as_stream(stream)>>>(
      d_partials, npartials, d_norm_out, accuracy, term, d_done, d_nterms);
  FQEHIP_CHECK_LAUNCH("k_norm_finalize");
  return FQEHIP_OK;
}

extern "C" int fqehip_convert_from_c128(int precision, int64_t n,
                                        const double* d_in, void* d_out,
                                        void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(n >= 0 && (n == 0 || (d_in && d_out)), "bad arguments");
  if (n == 0) return FQEHIP_OK;
  hipStream_t st = as_stream(stream);
  if (precision == FQEHIP_C128)
    k_from_c128<double><<<grid_stride_blocks(2 * n), kBlock, 0, st>>>(
        2 * n, d_in, static_cast<double*>(d_out));
  else
    k_from_c128<float><<<grid_stride_blocks(2 * n), kBlock, 0, st>>>(
        2 * n, d_in, static_cast<float*>(d_out));
  FQEHIP_CHECK_LAUNCH("k_from_c128");
  return FQEHIP_OK;
}

extern "C" int fqehip_convert_to_c128(int precision, int64_t n,
                                      const void* d_in, double scale_re,
                                      double scale_im, double* d_out,
                                      void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(n >= 0 && (n == 0 || (d_in && d_out)), "bad arguments");
  if (n == 0) return FQEHIP_OK;
  hipStream_t st = as_stream(stream);
  if (precision == FQEHIP_C128)
    k_to_c128<double><<<grid_stride_blocks(n), kBlock, 0, st>>>(
        n, static_cast<const double*>(d_in), scale_re, scale_im, d_out);
  else
    k_to_c128<float><<<grid_stride_blocks(n), kBlock, 0, st>>>(
        n, static_cast<const float*>(d_in), scale_re, scale_im, d_out);
  FQEHIP_CHECK_LAUNCH("k_to_c128");
  return FQEHIP_OK;
}

// ---------------------------------------------------------------------------
// Composed apply and Taylor loop.
// ---------------------------------------------------------------------------
namespace {

int check_sector_op(int precision, const fqehip_sector* s,
                    const fqehip_operator* op) {
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(s && op, "NULL sector or operator");
  FQEHIP_CHECK_ARG(s->npair >= 0 && s->lena >= 0 && s->lenb >= 0,
                   "negative sector size");
  FQEHIP_CHECK_ARG(s->ptab_a && s->ptab_b, "sector without pair tables");
  FQEHIP_CHECK_ARG(op->h1 || op->a, "operator has neither A nor h1");
  FQEHIP_CHECK_ARG(op->complex_op == 0 || op->complex_op == 1,
                   "complex_op=%d", op->complex_op);
  FQEHIP_CHECK_ARG(s->tile == 0 || s->tile == FQEHIP_TILE_ROWS4 ||
                       s->tile == FQEHIP_TILE_ROWS8 ||
                       s->tile == FQEHIP_TILE_SPLIT_ALLROWS ||
                       s->tile == FQEHIP_TILE_SPLIT_F64_DTILE ||
                       s->tile == FQEHIP_TILE_OZAKI, "tile=%d", s->tile);
  return FQEHIP_OK;
}

size_t elem_size(int precision) {
  return precision == FQEHIP_C128 ? sizeof(double) : sizeof(float);
}

// The complex128 split apply with Ozaki-digit contractions.
bool ozaki_split(int precision, const fqehip_sector* s) {
  return precision == FQEHIP_C128 && s->hbb_f64 != nullptr &&
         s->tile == FQEHIP_TILE_OZAKI;
}

// Workspace of the fused / split paths: slot rows, row accumulators and,
// for the Ozaki-digit split apply, the beta gather table (lenb x 128
// words), the row maxima of |C| (lena doubles) and, with hbb_oz, the C
// digit planes of the beta-beta GEMM.
int64_t fused_workspace_bytes(int precision, const fqehip_sector* s) {
  const int64_t n = s->lena * s->lenb;
  int64_t b = static_cast<int64_t>(s->nslot + 1) * 2 * n *
              static_cast<int64_t>(elem_size(precision));
  if (ozaki_split(precision, s)) {
    b += split_oz_table_bytes(s->lenb);
    b += s->lena * static_cast<int64_t>(sizeof(double));  // row maxima
    if (s->hbb_oz) b += beta_oz_workspace_bytes(s->lena, s->lenb);
  }
  return b;
}

}  // namespace

extern "C" int fqehip_fused_supported(int precision, const fqehip_sector* s,
                                      const fqehip_operator* op) {
  if (!(s && op && op->a && !op->h1 && !op->complex_op && s->fwd_a &&
        s->fwd_b && s->slot_a))
    return 0;
  if (ozaki_split(precision, s))
    return split_oz_supported(s->norb, s->npair, s->lena, s->lenb) &&
           (s->hbb_oz ? beta_oz_shape_ok(s->lena, s->lenb)
                      : beta_gemm_f64_shape_ok(s->lena, s->lenb));
  if (precision == FQEHIP_C128 && s->hbb_f64)
    return split_f64_supported(s->npair, s->lena, s->lenb, s->tile) &&
           beta_gemm_f64_shape_ok(s->lena, s->lenb);
  return fused_supported(precision, s->npair, s->lena, s->lenb);
}

namespace fqehip {
namespace {

// fqehip_apply_fused_rows; reuse_tables: see apply_reusing_tables.
int fused_rows(int precision, const fqehip_sector* s, const fqehip_operator* op,
               const void* d_coeff, void* d_workspace, const int32_t* d_done,
               void* stream, bool reuse_tables) {
  if (int rc = check_sector_op(precision, s, op)) return rc;
  FQEHIP_CHECK_ARG(fqehip_fused_supported(precision, s, op),
                   "the fused path does not support this sector/operator");
  FQEHIP_CHECK_ARG(d_coeff && d_workspace, "NULL pointer");
  FQEHIP_CHECK_ARG(s->nslot == s->nalpha * (s->norb - s->nalpha),
                   "nslot=%d != nalpha*(norb-nalpha)", s->nslot);
  const int64_t n = s->lena * s->lenb;
  char* ws = static_cast<char*>(d_workspace);
  void* rowacc = ws + static_cast<int64_t>(s->nslot) * 2 * n * elem_size(precision);
  hipStream_t st = as_stream(stream);
  if (ozaki_split(precision, s)) {
    // row maxima of |C| (digit scales), rowacc <- C Hbb^T (digit products
    // when the operator's Hbb digits are set, else FP64), then the alpha
    // rows (digit products)
    void* btab = static_cast<double*>(rowacc) + 2 * n;
    double* rowmax = reinterpret_cast<double*>(static_cast<char*>(btab) +
                                               split_oz_table_bytes(s->lenb));
    int rc = launch_row_absmax(s->lena, s->lenb,
                               static_cast<const double*>(d_coeff), rowmax, st,
                               d_done);
    if (rc) return rc;
    if (s->hbb_oz)
      rc = launch_beta_gemm_oz(s->lena, s->lenb, s->hbb_oz,
                               static_cast<const double*>(d_coeff), rowmax,
                               rowmax + s->lena, static_cast<double*>(rowacc),
                               st, d_done);
    else
      rc = launch_beta_gemm_f64(s->lena, s->lenb, s->hbb_f64,
                                static_cast<const double*>(d_coeff),
                                static_cast<double*>(rowacc), st, d_done);
    if (rc) return rc;
    return launch_split_oz(s->norb, s->npair, s->lena, s->lenb, op->a,
                           s->ptab_a, s->ptab_b, s->fwd_a, s->slot_a, s->nslot,
                           static_cast<const double*>(d_coeff), rowmax, btab,
                           reinterpret_cast<uint32_t*>(ws),
                           static_cast<double*>(rowacc), !reuse_tables, st,
                           d_done);
  }
  if (precision == FQEHIP_C128 && s->hbb_f64 != nullptr) {
    // rowacc <- C Hbb^T (FP64), then the live alpha rows add their part
    int rc = launch_beta_gemm_f64(s->lena, s->lenb, s->hbb_f64,
                                  static_cast<const double*>(d_coeff),
                                  static_cast<double*>(rowacc), st, d_done);
    if (rc) return rc;
    return launch_fused(precision, s->norb, s->npair, s->lena, s->lenb, op->a,
                        s->ptab_a, s->ptab_b, s->fwd_a, s->fwd_b, s->slot_a,
                        s->nslot, d_coeff, ws, rowacc, true, s->tile, st,
                        d_done);
  }
  const bool split = precision == FQEHIP_C64 && s->hbb_hi != nullptr;
  if (split) {
    FQEHIP_CHECK_ARG(s->hbb_lo != nullptr, "hbb_hi set without hbb_lo");
    FQEHIP_CHECK_ARG(beta_gemm_shape_ok(s->lena, s->lenb, s->ldh) &&
                         split_supported(s->npair, s->lena, s->lenb),
                     "the split apply does not support this sector (ldh=%lld)",
                     static_cast<long long>(s->ldh));
    // rowacc <- C Hbb^T, then the alpha rows add their diagonal part
    int rc = launch_beta_gemm(s->lena, s->lenb, s->ldh, s->hbb_hi, s->hbb_lo,
                              static_cast<const float*>(d_coeff),
                              static_cast<float*>(rowacc), st, d_done);
    if (rc) return rc;
  }
  return launch_fused(precision, s->norb, s->npair, s->lena, s->lenb, op->a,
                      s->ptab_a, s->ptab_b, s->fwd_a, s->fwd_b, s->slot_a,
                      s->nslot, d_coeff, ws, rowacc, split, s->tile, st,
                      d_done);
}

}  // namespace
}  // namespace fqehip

extern "C" int fqehip_apply_fused_rows(int precision, const fqehip_sector* s,
                                       const fqehip_operator* op,
                                       const void* d_coeff, void* d_workspace,
                                       const int32_t* d_done, void* stream) {
  clear_error();
  return fused_rows(precision, s, op, d_coeff, d_workspace, d_done, stream,
                    false);
}

extern "C" int fqehip_apply_slot_reduce(int precision, const fqehip_sector* s,
                                        const void* d_workspace,
                                        double alpha_re, double alpha_im,
                                        double beta_re, double beta_im,
                                        const void* d_cin, void* d_out,
                                        double* d_acc, double* d_norm_partials,
                                        const int32_t* d_done, void* stream) {
  clear_error();
  if (int rc = check_precision(precision)) return rc;
  FQEHIP_CHECK_ARG(s && d_workspace && d_out, "NULL pointer");
  FQEHIP_CHECK_ARG((beta_re == 0.0 && beta_im == 0.0) || d_cin,
                   "beta != 0 needs d_cin");
  const int64_t n = s->lena * s->lenb;
  const char* ws = static_cast<const char*>(d_workspace);
  const void* rowacc = ws + static_cast<int64_t>(s->nslot) * 2 * n * elem_size(precision);
  // the Ozaki-digit row kernel writes the slot rows in 48 bits per double
  return launch_slot_reduce(precision, ozaki_split(precision, s), s->nslot,
                            s->lena, s->lenb, ws, rowacc,
                            alpha_re, alpha_im, beta_re, beta_im, d_cin, d_out,
                            d_acc, d_norm_partials, as_stream(stream), d_done);
}

extern "C" int64_t fqehip_apply_workspace_bytes(int precision,
                                               const fqehip_sector* s) {
  if (!s) return -1;
  const int64_t n = s->lena * s->lenb;
  const int64_t unfused = static_cast<int64_t>(2 * s->npair + 1) * 2 * n *
                          static_cast<int64_t>(elem_size(precision));
  const int64_t fused = fused_workspace_bytes(precision, s);
  return unfused > fused ? unfused : fused;
}

// fqehip_apply.  reuse_tables: the Ozaki-digit split apply takes the beta
// gather table and the pair rows' digits from the workspace as an earlier
// apply of the same sector and operator with the same workspace left them
// (they depend on neither the coefficients nor alpha/beta), instead of
// rebuilding them: the later terms of a Taylor series.
int fqehip::apply_reusing_tables(
    int precision, const fqehip_sector* s, const fqehip_operator* op,
    const void* d_coeff, double alpha_re, double alpha_im, double beta_re,
    double beta_im, const void* d_cin, void* d_out, double* d_acc,
    double* d_norm_partials, const int32_t* d_done, void* d_workspace,
    void* stream, bool reuse_tables) {
  if (int rc = check_sector_op(precision, s, op)) return rc;
  FQEHIP_CHECK_ARG(d_coeff && d_out, "NULL coefficient pointer");
  if (op->h1)
    return fqehip_apply1(precision, s->npair, s->lena, s->lenb, s->ptab_a,
                         s->ptab_b, op->h1, d_coeff, alpha_re, alpha_im,
                         beta_re, beta_im, d_cin, d_out, d_acc,
                         d_norm_partials, d_done, stream);
  FQEHIP_CHECK_ARG(d_workspace != nullptr, "NULL workspace");
  const int64_t n = s->lena * s->lenb;
  char* ws = static_cast<char*>(d_workspace);
  if (fqehip_fused_supported(precision, s, op)) {
    int rc = fused_rows(precision, s, op, d_coeff, d_workspace, d_done, stream,
                        reuse_tables);
    if (rc) return rc;
    return fqehip_apply_slot_reduce(precision, s, d_workspace, alpha_re,
                                    alpha_im, beta_re, beta_im, d_cin, d_out,
                                    d_acc, d_norm_partials, d_done, stream);
  }
  void* dvec = ws;
  void* y = ws + static_cast<int64_t>(s->npair) * 2 * n * elem_size(precision);
  hipStream_t st = as_stream(stream);
  int rc = launch_gather_any(precision, s->npair, s->lena, s->lenb, s->ptab_a,
                             s->ptab_b, d_coeff, dvec, st, d_done);
  if (rc) return rc;
  if (op->complex_op)
    rc = launch_gemm(precision, 2 * (s->npair + 1), 2 * s->npair, op->a, n,
                     dvec, y, st, d_done);
  else
    rc = launch_gemm(precision, s->npair + 1, s->npair, op->a, 2 * n, dvec, y,
                     st, d_done);
  if (rc) return rc;
  return fqehip_scatter(precision, s->npair, s->lena, s->lenb, s->ptab_a,
                        s->ptab_b, y, alpha_re, alpha_im, beta_re, beta_im,
                        d_cin, d_out, d_acc, d_norm_partials, d_done, stream);
}

extern "C" int fqehip_apply(int precision, const fqehip_sector* s,
                            const fqehip_operator* op, const void* d_coeff,
                            double alpha_re, double alpha_im, double beta_re,
                            double beta_im, const void* d_cin, void* d_out,
                            double* d_acc, double* d_norm_partials,
                            const int32_t* d_done, void* d_workspace,
                            void* stream) {
  clear_error();
  return apply_reusing_tables(precision, s, op, d_coeff, alpha_re, alpha_im,
                              beta_re, beta_im, d_cin, d_out, d_acc,
                              d_norm_partials, d_done, d_workspace, stream,
                              false);
}

extern "C" int fqehip_taylor(int precision, const fqehip_sector* s,
                             const fqehip_operator* op, double time,
                             double accuracy, int expansion, void* d_term,
                             double* d_result, int32_t* d_status,
                             double* d_partials, double* d_norm,
                             void* d_workspace, void* stream) {
  clear_error();
  if (int rc = check_sector_op(precision, s, op)) return rc;
  FQEHIP_CHECK_ARG(d_term && d_result && d_status && d_partials && d_norm,
                   "NULL pointer");
  FQEHIP_CHECK_ARG(d_workspace != nullptr, "NULL workspace");
  hipStream_t st = as_stream(stream);
  k_status_init<<<1, 64, 0, st>>>(d_status);
  FQEHIP_CHECK_LAUNCH("k_status_init");
  const int64_t n = s->lena * s->lenb;
  const int64_t np = fqehip_norm_partials_count(n);
  int32_t* nterms = d_status;
  int32_t* done = d_status + 1;
  // The one-body apply cannot run in place: its terms ping-pong between
  // d_term and the workspace.
  void* cur = d_term;
  void* nxt = d_workspace;
  for (int order = 1; order < expansion; ++order) {
    // term <- (-i t / order) H term ;  result += term
    const double ar = 0.0, ai = -time / static_cast<double>(order);
    int rc;
    if (op->h1) {
      rc = fqehip_apply1(precision, s->npair, s->lena, s->lenb, s->ptab_a,
                         s->ptab_b, op->h1, cur, ar, ai, 0.0, 0.0, nullptr,
                         nxt, d_result, d_partials, done, stream);
      void* t = cur;
      cur = nxt;
      nxt = t;
    } else {
      rc = apply_reusing_tables(precision, s, op, d_term, ar, ai, 0.0, 0.0,
                                nullptr, d_term, d_result, d_partials, done,
                                d_workspace, stream, order > 1);
    }
    if (rc) return rc;
    rc = fqehip_norm_finalize(d_partials, np, d_norm, accuracy, order, done,
                              nterms, stream);
    if (rc) return rc;
  }
  return FQEHIP_OK;
}

extern "C" int fqehip_apply_panels(int precision, const fqehip_sector* s,
                                   const fqehip_operator* op,
                                   const void* d_coeff, double alpha_re,
                                   double alpha_im, double beta_re,
                                   double beta_im, const void* d_cin,
                                   void* d_out, double* d_acc,
                                   double* d_norm_partials,
                                   const int32_t* d_done, void* d_workspace,
                                   int64_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_sector_op(precision, s, op)) return rc;
  FQEHIP_CHECK_ARG(precision == FQEHIP_C128,
                   "the chunked apply runs in complex128");
  FQEHIP_CHECK_ARG(op->a && !op->h1, "the chunked apply needs a two-body "
                   "operator (fqehip_apply takes one-body operators)");
  FQEHIP_CHECK_ARG(d_coeff && d_out && d_workspace, "NULL pointer");
  FQEHIP_CHECK_ARG((beta_re == 0.0 && beta_im == 0.0) || d_cin,
                   "beta != 0 needs d_cin");
  const int64_t n = s->lena * s->lenb;
  if (n == 0) return FQEHIP_OK;
  const int npair = s->npair;
  const int64_t plane = 2 * n;                 // doubles per pair plane
  const int64_t planes = workspace_bytes / (plane * 8) - 1;   // oacc first
  FQEHIP_CHECK_ARG(planes >= 2,
                   "workspace of %lld bytes holds fewer than 3 planes of %lld "
                   "bytes", static_cast<long long>(workspace_bytes),
                   static_cast<long long>(plane * 8));
  int ki = static_cast<int>(planes / 2 < npair ? planes / 2 : npair);
  if (ki < 1) ki = 1;
  int64_t left = planes - ki;
  int mo = static_cast<int>(left < npair + 1 ? left : npair + 1);
  const bool complex_op = op->complex_op != 0;
  const int w = complex_op ? 2 : 1;            // GEMM rows / cols per pair
  const int64_t ncol = complex_op ? n : 2 * n;
  double* ws = static_cast<double*>(d_workspace);
  double* oacc = ws;
  double* dpan = ws + plane;
  double* ypan = dpan + static_cast<int64_t>(ki) * plane;
  const int lda = w * npair;
  hipStream_t st = as_stream(stream);
  const bool dall = ki >= npair;               // the whole dvec fits
  if (dall) {
    int rc = launch_gather<double>(npair, s->lena, s->lenb, s->ptab_a,
                                   s->ptab_b, d_coeff, dpan, st, d_done);
    if (rc) return rc;
  }
  for (int p0 = 0; p0 < npair + 1; p0 += mo) {
    const int np = (npair + 1 - p0) < mo ? (npair + 1 - p0) : mo;
    for (int q0 = 0; q0 < npair; q0 += ki) {
      const int nq = (npair - q0) < ki ? (npair - q0) : ki;
      if (!dall) {
        int rc = launch_gather<double>(
            nq, s->lena, s->lenb, s->ptab_a + static_cast<int64_t>(q0) * s->lena,
            s->ptab_b + static_cast<int64_t>(q0) * s->lenb, d_coeff, dpan, st,
            d_done);
        if (rc) return rc;
      }
      const double* a = op->a + static_cast<int64_t>(w * p0) * lda + w * q0;
      int rc = launch_gemm_f64(w * np, w * nq, a, ncol,
                               dpan + (dall ? static_cast<int64_t>(q0) * plane : 0),
                               ypan, st, d_done, lda, q0 > 0);
      if (rc) return rc;
    }
    k_scatter_panel<double><<<blocks_for(n, kBlock), kBlock, 0, st>>>(
        p0, np, npair, s->lena, s->lenb, 1.0 / s->lenb, s->ptab_a, s->ptab_b,
        ypan, oacc, p0 == 0, d_done);
    FQEHIP_CHECK_LAUNCH("k_scatter_panel");
  }
  k_panel_epilogue<double><<<blocks_for(n, kBlock), kBlock, 0, st>>>(
      n, oacc, alpha_re, alpha_im, beta_re, beta_im,
      static_cast<const double*>(d_cin), static_cast<double*>(d_out), d_acc,
      d_norm_partials, d_done);
  FQEHIP_CHECK_LAUNCH("k_panel_epilogue");
  return FQEHIP_OK;
}
