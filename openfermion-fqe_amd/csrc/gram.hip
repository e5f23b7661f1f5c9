// Complex Gram matrix G[p][q] = sum_i conj(A[p][i]) B[q][i] of two blocks of
// k complex128 vectors (k rows, n interleaved complex entries each) on the
// FP64 matrix cores: the 2-RDM's dvec overlap D2^H D
// (reference src/fqe/fqe_data.py:1767-1782, tensordot of dvec2.conj() and
// dvec over the determinant axes).
//
// With A = X + iY and B = U + iV:
//   Re G = X U^T + Y V^T,   Im G = X V^T - Y U^T,
// four v_mfma_f64_16x16x4f64 per 4-element step of i.  A wave owns one
// (16 p) x (16 q) tile for one chunk of i; the chunk partials land in a
// workspace [nchunk][k][k] and a second kernel sums them in a fixed order.
#include "fqehip_common.h"

namespace fqehip {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kGWaves = 4;
constexpr int kGChunk = 4096;            // i per chunk

// grid.x: tile pairs / kGWaves, grid.y: chunks of i
__global__ void __launch_bounds__(kGWaves * 64)
k_gram_f64(int64_t n, int k, const double* __restrict__ a,
           const double* __restrict__ b, double* __restrict__ part) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int tiles = (k + 15) / 16;
  const int pair = blockIdx.x * kGWaves + wave;
  if (pair >= tiles * tiles) return;
  const int tp = pair / tiles;
  const int tq = pair - tp * tiles;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kGChunk;
  const int64_t i1 = i0 + kGChunk < n ? i0 + kGChunk : n;
  // f64 MFMA operands: A[r = l%16][kk = l/16], B[kk = l/16][c = l%16]
  const int p = tp * 16 + (lane & 15);
  const int q = tq * 16 + (lane & 15);
  const bool pok = p < k;
  const bool qok = q < k;
  const double2* ar = reinterpret_cast<const double2*>(a + 2 * static_cast<int64_t>(pok ? p : 0) * n);
  const double2* br = reinterpret_cast<const double2*>(b + 2 * static_cast<int64_t>(qok ? q : 0) * n);
  double4_t re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
  // every lane runs the same trip count (the MFMAs are wave-wide); a lane
  // whose i is past the chunk contributes zeros
  for (int64_t ib = i0; ib < i1; ib += 4) {
    const int64_t i = ib + (lane >> 4);
    const bool iok = i < i1;
    const double2 x = (pok && iok) ? ar[i] : make_double2(0.0, 0.0);
    const double2 u = (qok && iok) ? br[i] : make_double2(0.0, 0.0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, u.x, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, u.y, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, u.y, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, u.x, im, 0, 0, 0);
  }
  // f64 MFMA layout: item j of lane l is row l/16 + 4 j, column l%16
  double* dst = part + static_cast<int64_t>(blockIdx.y) * k * k * 2;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int pr = tp * 16 + (lane >> 4) + 4 * j;
    const int qc = tq * 16 + (lane & 15);
    if (pr < k && qc < k) {
      dst[(static_cast<int64_t>(pr) * k + qc) * 2] = re[j];
      dst[(static_cast<int64_t>(pr) * k + qc) * 2 + 1] = im[j];
    }
  }
}

__global__ void __launch_bounds__(256)
k_gram_reduce(int64_t kk2, int nchunk, const double* __restrict__ part,
              double* __restrict__ out) {
  const int64_t x = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (x >= kk2) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[static_cast<int64_t>(c) * kk2 + x];
  out[x] = s;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int64_t fqehip_gram_workspace_bytes(int64_t n, int k) {
  if (n <= 0 || k <= 0) return 0;
  const int64_t nchunk = (n + kGChunk - 1) / kGChunk;
  return nchunk * k * k * 2 * static_cast<int64_t>(sizeof(double));
}

extern "C" int fqehip_gram(int64_t n, int k, const double* d_a,
                           const double* d_b, double* d_out,
                           double* d_workspace, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(n >= 0 && k >= 0 && k <= 4096, "n=%lld k=%d",
                   static_cast<long long>(n), k);
  if (k == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_out != nullptr, "NULL output");
  hipStream_t st = as_stream(stream);
  const int64_t kk2 = static_cast<int64_t>(k) * k * 2;
  if (n == 0) {
    // an empty sum: zero the output with a kernel (no memset nodes)
    k_gram_reduce<<<blocks_for(kk2, 256), 256, 0, st>>>(kk2, 0, nullptr, d_out);
    FQEHIP_CHECK_LAUNCH("k_gram_reduce");
    return FQEHIP_OK;
  }
  FQEHIP_CHECK_ARG(d_a && d_b && d_workspace, "NULL pointer");
  const int tiles = (k + 15) / 16;
  const int64_t nchunk = (n + kGChunk - 1) / kGChunk;
  FQEHIP_CHECK_ARG(nchunk <= 65535, "n=%lld too large for one Gram pass",
                   static_cast<long long>(n));
  dim3 grid(static_cast<unsigned>((tiles * tiles + kGWaves - 1) / kGWaves),
            static_cast<unsigned>(nchunk));
  k_gram_f64<<<grid, kGWaves * 64, 0, st>>>(n, k, d_a, d_b, d_workspace);
  FQEHIP_CHECK_LAUNCH("k_gram_f64");
  k_gram_reduce<<<blocks_for(kk2, 256), 256, 0, st>>>(
      kk2, static_cast<int>(nchunk), d_workspace, d_out);
  FQEHIP_CHECK_LAUNCH("k_gram_reduce");
  return FQEHIP_OK;
}
