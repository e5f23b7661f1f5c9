// (c) Mixed-precision pair GEMM: Y (FP32) = A B with A (FP64 on input) and
// B (FP32) each split into two bfloat16 terms, x = x_hi + x_lo, and the
// three leading products A_hi B_hi + A_hi B_lo + A_lo B_hi accumulated in
// FP32 on v_mfma_f32_16x16x32_bf16.  The dropped A_lo B_lo term and the
// 16-bit split leave a relative error of order 2^-16 per product.
#include "fqehip_common.h"

namespace fqehip {

namespace {

constexpr int kBlock = 256;
constexpr int kRowTiles = 8;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even float -> bfloat16 (finite inputs).
__device__ __forceinline__ __bf16 to_bf16(float x) {
  return static_cast<__bf16>(x);  // round to nearest even (v_cvt_pk_bf16_f32)
}

__device__ __forceinline__ float from_bf16(__bf16 h) {
  const uint16_t b = __builtin_bit_cast(uint16_t, h);
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

__device__ __forceinline__ void split2(float x, __bf16& hi, __bf16& lo) {
  hi = to_bf16(x);
  lo = to_bf16(x - from_bf16(hi));
}

// Same tiling as the FP64 kernel: a wave owns 16 columns and up to
// kRowTiles row tiles; 4 waves per workgroup.  Operand layout of the
// 16x16x32 bf16 MFMA: A[r = lane%16][k = 8*(lane/16) + 0..7],
// B[k = 8*(lane/16) + 0..7][c = lane%16]; result item i of lane l is
// Y[4*(l/16) + i][l%16].
__global__ void __launch_bounds__(kBlock)
k_gemm_bf16x3(int m, int k, const double* __restrict__ a, int64_t ncol,
              const float* __restrict__ b, float* __restrict__ y,
              const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t col0 = (static_cast<int64_t>(blockIdx.x) * 4 + wave) * 16;
  if (col0 >= ncol) return;
  const int row0 = blockIdx.y * kRowTiles * 16;
  const int64_t col = col0 + (lane & 15);
  const bool colok = col < ncol;
  const int kq = 8 * (lane >> 4);
  int ntiles = (m - row0 + 15) / 16;
  if (ntiles > kRowTiles) ntiles = kRowTiles;
  floatx4 acc[kRowTiles];
#pragma unroll
  for (int t = 0; t < kRowTiles; ++t) acc[t] = floatx4{0, 0, 0, 0};
  for (int k0 = 0; k0 < k; k0 += 32) {
    bf16x8 bhi, blo;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = k0 + kq + j;
      const float v = (colok && kk < k) ? b[static_cast<int64_t>(kk) * ncol + col] : 0.0f;
      __bf16 h, l;
      split2(v, h, l);
      bhi[j] = h;
      blo[j] = l;
    }
#pragma unroll
    for (int t = 0; t < kRowTiles; ++t) {
      if (t < ntiles) {
        const int r = row0 + t * 16 + (lane & 15);
        bf16x8 ahi, alo;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int kk = k0 + kq + j;
          const double v = (r < m && kk < k) ? a[static_cast<int64_t>(r) * k + kk] : 0.0;
          // two-term split of the FP64 value: hi from the double, lo from
          // the remainder
          const __bf16 h = to_bf16(static_cast<float>(v));
          const __bf16 l = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
          ahi[j] = h;
          alo[j] = l;
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc[t], 0, 0, 0);
      }
    }
  }
  if (!colok) return;
#pragma unroll
  for (int t = 0; t < kRowTiles; ++t) {
    if (t < ntiles) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + t * 16 + 4 * (lane >> 4) + i;
        if (r < m) y[static_cast<int64_t>(r) * ncol + col] = acc[t][i];
      }
    }
  }
}

}  // namespace

int launch_gemm_bf16x3(int m, int k, const double* a, int64_t ncol,
                       const float* b, float* y, hipStream_t st,
                       const int32_t* done) {
  if (m == 0 || ncol == 0) return FQEHIP_OK;
  dim3 grid(blocks_for(ncol, 64), (m + kRowTiles * 16 - 1) / (kRowTiles * 16));
  k_gemm_bf16x3<<<grid, kBlock, 0, st>>>(m, k, a, ncol, b, y, done);
  FQEHIP_CHECK_LAUNCH("k_gemm_bf16x3");
  return FQEHIP_OK;
}

}  // namespace fqehip
