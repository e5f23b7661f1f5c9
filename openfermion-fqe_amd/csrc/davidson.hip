// Kernels of the Davidson eigensolver: the Slater-Condon diagonal <I|H|I>,
// blocked FP64 inner products and linear combinations of subspace vectors
// (Gram-Schmidt and Ritz vectors), and the diagonal preconditioner with an
// optional per-determinant correction.  Vectors are complex128, [n]
// interleaved; a block of k vectors is k rows of stride ldv elements.
#include "fqehip_common.h"

namespace fqehip {
namespace {

constexpr int kThreads = 256;
constexpr int kDotTile = 8;      // vectors per block row of k_block_dot
constexpr int kDotBlocks = 512;  // partial sums per vector

// <I|H|I> = sum_i h1[i,i] (na_i + nb_i)
//         + sum_{i != j} (K_ij - J_ij) (na_i na_j + nb_i nb_j)
//         - sum_{i, j} J_ij (na_i nb_j + nb_i na_j)
// with J_ij = h2[i,j,i,j], K_ij = h2[i,j,j,i] (real parts) for
// H = sum h1 E + sum h2[i,j,k,l] a+_{i r} a+_{j e} a_{k r} a_{l e}.
__global__ void __launch_bounds__(kThreads)
k_diagonal(int norb, int64_t lena, int64_t lenb,
           const int64_t* __restrict__ astr, const int64_t* __restrict__ bstr,
           const double* __restrict__ h1d, const double* __restrict__ jm,
           const double* __restrict__ km, double* __restrict__ diag) {
  __shared__ double sj[kMaxNorb * kMaxNorb];
  __shared__ double sk[kMaxNorb * kMaxNorb];
  __shared__ double sh[kMaxNorb];
  for (int x = threadIdx.x; x < norb * norb; x += kThreads) {
    sj[x] = jm[x];
    sk[x] = km[x];
  }
  for (int x = threadIdx.x; x < norb; x += kThreads) sh[x] = h1d[x];
  __syncthreads();
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= lena * lenb) return;
  const int64_t ia = idx / lenb;
  const int64_t ib = idx - ia * lenb;
  const uint64_t sa = static_cast<uint64_t>(astr[ia]);
  const uint64_t sb = static_cast<uint64_t>(bstr[ib]);
  double e = 0.0;
  for (int i = 0; i < norb; ++i) {
    const double ai = static_cast<double>((sa >> i) & 1);
    const double bi = static_cast<double>((sb >> i) & 1);
    if (ai == 0.0 && bi == 0.0) continue;
    e += sh[i] * (ai + bi);
    for (int j = 0; j < norb; ++j) {
      const double aj = static_cast<double>((sa >> j) & 1);
      const double bj = static_cast<double>((sb >> j) & 1);
      const double jij = sj[i * norb + j];
      if (j != i) e += (sk[i * norb + j] - jij) * (ai * aj + bi * bj);
      e -= jij * (ai * bj + bi * aj);
    }
  }
  diag[idx] = e;
}

// Partial sums of out[j] = sum_i conj(v_j[i]) x[i] for a tile of kDotTile
// vectors (blockIdx.y) over a grid-stride share of i (blockIdx.x).
__global__ void __launch_bounds__(kThreads)
k_block_dot(int64_t n, int k, const double* __restrict__ v, int64_t ldv,
            const double* __restrict__ x, double* __restrict__ partials) {
  __shared__ double red[kThreads / 64][2 * kDotTile];
  const int j0 = blockIdx.y * kDotTile;
  double acc[2 * kDotTile];
#pragma unroll
  for (int t = 0; t < 2 * kDotTile; ++t) acc[t] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const double2 xi = reinterpret_cast<const double2*>(x)[i];
#pragma unroll
    for (int t = 0; t < kDotTile; ++t) {
      if (j0 + t < k) {
        const double2 vi = reinterpret_cast<const double2*>(v + 2 * (j0 + t) * ldv)[i];
        acc[2 * t] += vi.x * xi.x + vi.y * xi.y;
        acc[2 * t + 1] += vi.x * xi.y - vi.y * xi.x;
      }
    }
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < 2 * kDotTile; ++t) {
    double s = acc[t];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wave][t] = s;
  }
  __syncthreads();
  if (threadIdx.x < 2 * kDotTile) {
    const int t = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][t];
    const int j = j0 + t / 2;
    if (j < k) partials[(static_cast<int64_t>(j) * gridDim.x + blockIdx.x) * 2 + (t & 1)] = s;
  }
}

// out[j] = sum over the partials of vector j (fixed order: deterministic).
__global__ void __launch_bounds__(64)
k_block_dot_finalize(int nparts, const double* __restrict__ partials,
                     double* __restrict__ out) {
  const int j = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 64) {
    re += partials[(static_cast<int64_t>(j) * nparts + p) * 2];
    im += partials[(static_cast<int64_t>(j) * nparts + p) * 2 + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_down(re, off, 64);
    im += __shfl_down(im, off, 64);
  }
  if (threadIdx.x == 0) {
    out[2 * j] = re;
    out[2 * j + 1] = im;
  }
}

// x[i] = beta x[i] + sum_j c_j v_j[i]  (complex c_j, beta; c on the device)
__global__ void __launch_bounds__(kThreads)
k_block_combine(int64_t n, int k, const double* __restrict__ v, int64_t ldv,
                const double* __restrict__ c, double beta_re, double beta_im,
                double* __restrict__ x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  double re = 0.0, im = 0.0;
  if (beta_re != 0.0 || beta_im != 0.0) {
    const double2 xi = reinterpret_cast<const double2*>(x)[i];
    re = beta_re * xi.x - beta_im * xi.y;
    im = beta_re * xi.y + beta_im * xi.x;
  }
  for (int j = 0; j < k; ++j) {
    const double cr = c[2 * j], ci = c[2 * j + 1];
    const double2 vi = reinterpret_cast<const double2*>(v + 2 * j * ldv)[i];
    re += cr * vi.x - ci * vi.y;
    im += cr * vi.y + ci * vi.x;
  }
  reinterpret_cast<double2*>(x)[i] = make_double2(re, im);
}

// out[j][i] = sum_l c[l][j] v[l][i] for j < m <= kMultiMax: every row v[l]
// is read once for all m outputs (a thick restart's new basis and images)
constexpr int kMultiMax = 16;

__global__ void __launch_bounds__(kThreads)
k_block_combine_multi(int64_t n, int k, const double* __restrict__ v,
                      int64_t ldv, const double* __restrict__ c, int m,
                      double* __restrict__ out, int64_t ldo) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  double re[kMultiMax], im[kMultiMax];
#pragma unroll
  for (int j = 0; j < kMultiMax; ++j) {
    re[j] = 0.0;
    im[j] = 0.0;
  }
  for (int l = 0; l < k; ++l) {
    const double2 vi = reinterpret_cast<const double2*>(v + 2 * l * ldv)[i];
    const double* cl = c + 2 * static_cast<int64_t>(l) * m;
#pragma unroll
    for (int j = 0; j < kMultiMax; ++j) {
      if (j < m) {
        const double cr = cl[2 * j], ci = cl[2 * j + 1];
        re[j] += cr * vi.x - ci * vi.y;
        im[j] += cr * vi.y + ci * vi.x;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kMultiMax; ++j)
    if (j < m)
      reinterpret_cast<double2*>(out + 2 * j * ldo)[i] = make_double2(re[j], im[j]);
}

// t[i] = -(1 / (d_i - theta) + corr_i) r[i], |d_i - theta| floored at eps
__global__ void __launch_bounds__(kThreads)
k_precondition(int64_t n, const double* __restrict__ r,
               const double* __restrict__ d, double theta, double eps,
               const float* __restrict__ corr, double* __restrict__ t) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  double den = d[i] - theta;
  if (fabs(den) < eps) den = den < 0.0 ? -eps : eps;
  double p = 1.0 / den;
  if (corr) p += static_cast<double>(corr[i]);
  const double2 ri = reinterpret_cast<const double2*>(r)[i];
  reinterpret_cast<double2*>(t)[i] = make_double2(-p * ri.x, -p * ri.y);
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_diagonal(int norb, int64_t lena, int64_t lenb,
                               const int64_t* d_astr, const int64_t* d_bstr,
                               const double* d_h1diag, const double* d_jmat,
                               const double* d_kmat, double* d_diag,
                               void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(norb >= 0 && norb <= kMaxNorb, "norb=%d", norb);
  FQEHIP_CHECK_ARG(lena >= 0 && lenb >= 0, "negative sector size");
  const int64_t n = lena * lenb;
  if (n == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_astr && d_bstr && d_diag, "NULL pointer");
  FQEHIP_CHECK_ARG(norb == 0 || (d_h1diag && d_jmat && d_kmat),
                   "NULL integral pointer");
  k_diagonal<<<blocks_for(n, kThreads), kThreads, 0, as_stream(stream)>>>(
      norb, lena, lenb, d_astr, d_bstr, d_h1diag, d_jmat, d_kmat, d_diag);
  FQEHIP_CHECK_LAUNCH("k_diagonal");
  return FQEHIP_OK;
}

extern "C" int64_t fqehip_block_dot_workspace_bytes(int k) {
  return k <= 0 ? 0 : static_cast<int64_t>(k) * kDotBlocks * 2 * sizeof(double);
}

extern "C" int fqehip_block_dot(int64_t n, int k, const double* d_v,
                                int64_t ldv, const double* d_x, double* d_out,
                                double* d_workspace, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(n >= 0 && k >= 0 && ldv >= n, "n=%lld k=%d ldv=%lld",
                   static_cast<long long>(n), k, static_cast<long long>(ldv));
  if (k == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_v && d_x && d_out && d_workspace, "NULL pointer");
  hipStream_t st = as_stream(stream);
  dim3 grid(kDotBlocks, (k + kDotTile - 1) / kDotTile);
  k_block_dot<<<grid, kThreads, 0, st>>>(n, k, d_v, ldv, d_x, d_workspace);
  FQEHIP_CHECK_LAUNCH("k_block_dot");
  k_block_dot_finalize<<<k, 64, 0, st>>>(kDotBlocks, d_workspace, d_out);
  FQEHIP_CHECK_LAUNCH("k_block_dot_finalize");
  return FQEHIP_OK;
}

extern "C" int fqehip_block_combine(int64_t n, int k, const double* d_v,
                                    int64_t ldv, const double* d_c,
                                    double beta_re, double beta_im,
                                    double* d_x, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(n >= 0 && k >= 0 && (k == 0 || ldv >= n),
                   "n=%lld k=%d ldv=%lld", static_cast<long long>(n), k,
                   static_cast<long long>(ldv));
  if (n == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_x && (k == 0 || (d_v && d_c)), "NULL pointer");
  k_block_combine<<<blocks_for(n, kThreads), kThreads, 0, as_stream(stream)>>>(
      n, k, d_v, ldv, d_c, beta_re, beta_im, d_x);
  FQEHIP_CHECK_LAUNCH("k_block_combine");
  return FQEHIP_OK;
}

extern "C" int fqehip_block_combine_multi(int64_t n, int k, const double* d_v,
                                          int64_t ldv, const double* d_c, int m,
                                          double* d_out, int64_t ldo,
                                          void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(n >= 0 && k >= 0 && m >= 0 && m <= kMultiMax &&
                       (k == 0 || ldv >= n) && (m == 0 || ldo >= n),
                   "n=%lld k=%d m=%d ldv=%lld ldo=%lld (m <= %d)",
                   static_cast<long long>(n), k, m, static_cast<long long>(ldv),
                   static_cast<long long>(ldo), kMultiMax);
  if (n == 0 || m == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_out && (k == 0 || (d_v && d_c)), "NULL pointer");
  k_block_combine_multi<<<blocks_for(n, kThreads), kThreads, 0,
                          as_stream(stream)>>>(n, k, d_v, ldv, d_c, m, d_out,
                                               ldo);
  FQEHIP_CHECK_LAUNCH("k_block_combine_multi");
  return FQEHIP_OK;
}

extern "C" int fqehip_davidson_precondition(int64_t n, const double* d_r,
                                            const double* d_diag, double theta,
                                            double eps, const float* d_corr,
                                            double* d_t, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(n >= 0 && eps > 0.0, "n=%lld eps=%g",
                   static_cast<long long>(n), eps);
  if (n == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_r && d_diag && d_t, "NULL pointer");
  k_precondition<<<blocks_for(n, kThreads), kThreads, 0, as_stream(stream)>>>(
      n, d_r, d_diag, theta, eps, d_corr, d_t);
  FQEHIP_CHECK_LAUNCH("k_precondition");
  return FQEHIP_OK;
}
