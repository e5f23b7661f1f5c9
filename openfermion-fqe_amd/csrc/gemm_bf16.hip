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

// The same product with A resident in LDS (m <= 144, the pair matrix of a
// real operator): each workgroup splits A once into bfloat16 hi/lo MFMA
// fragments (fragment (t, s, part) = 64 lanes x 8 bf16, 1 KB, one
// ds_read_b128 per lane), and its waves then walk 32-column panels of B
// (16 from 8 row tiles on) persistently.  A lane loads the 8 consecutive k of its column for each
// k step of 32 (eight 4-byte loads, 64 contiguous bytes per k row across
// the 16 lanes of a column tile), two steps ahead of the MFMAs, and splits
// them in registers.  Rows of B past k are
// clamped to k - 1 (they meet zero A fragments) and columns past ncol to
// column 0 (outputs never stored).
constexpr int kBxWaves = 4;
constexpr int kBxThreads = 64 * kBxWaves;
// column tiles per wave panel: 2 (32 columns), 1 from 8 row tiles on
template <int RT>
constexpr int bx_col_tiles() { return RT >= 8 ? 1 : 2; }
constexpr int kBxSteps = 5;                    // k <= 160
constexpr int kBxMaxRowTiles = 9;              // m <= 144

template <int RT, int KS>
__global__ void __launch_bounds__(kBxThreads)
k_gemm_bf16x3_lds(int m, int k, const double* __restrict__ a, int64_t ncol,
                  const float* __restrict__ b, float* __restrict__ y,
                  const int32_t* __restrict__ done) {
  constexpr int kBxColTiles = bx_col_tiles<RT>();
  extern __shared__ bf16x8 afr[];              // [RT][KS][2][64]
  if (done && *done) return;
  const int tid = threadIdx.x;
  for (int e = tid; e < RT * KS * 64; e += kBxThreads) {
    const int l = e & 63;
    const int ts = e >> 6;
    const int t = ts / KS;
    const int sk = ts - t * KS;
    const int r = 16 * t + (l & 15);
    bf16x8 hi, lo;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = 32 * sk + 8 * (l >> 4) + j;
      const double v = (r < m && kk < k) ? a[static_cast<int64_t>(r) * k + kk] : 0.0;
      const __bf16 h = to_bf16(static_cast<float>(v));
      hi[j] = h;
      lo[j] = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
    }
    afr[(ts * 2) * 64 + l] = hi;
    afr[(ts * 2 + 1) * 64 + l] = lo;
  }
  __syncthreads();
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int kq = 8 * (lane >> 4);
  const int64_t npanel = (ncol + 16 * kBxColTiles - 1) / (16 * kBxColTiles);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBxWaves;
  for (int64_t pan = static_cast<int64_t>(blockIdx.x) * kBxWaves + wave;
       pan < npanel; pan += stride) {
    const int64_t cbase = pan * 16 * kBxColTiles + (lane & 15);
    float bv[KS][kBxColTiles][8];
    auto load_step = [&](int sk) {
#pragma unroll
      for (int ct = 0; ct < kBxColTiles; ++ct) {
        const int64_t c = cbase + 16 * ct;
        const float* bc = b + (c < ncol ? c : 0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int kk0 = 32 * sk + kq + j;
          const int kk = kk0 < k ? kk0 : k - 1;
          bv[sk][ct][j] = __builtin_nontemporal_load(bc + static_cast<int64_t>(kk) * ncol);
        }
      }
    };
    // two k steps of B in flight ahead of the MFMAs
#pragma unroll
    for (int sk = 0; sk < KS && sk < 2; ++sk) load_step(sk);
    floatx4 acc[RT][kBxColTiles];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int ct = 0; ct < kBxColTiles; ++ct) acc[t][ct] = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int sk = 0; sk < KS; ++sk) {
      bf16x8 bhi[kBxColTiles], blo[kBxColTiles];
#pragma unroll
      for (int ct = 0; ct < kBxColTiles; ++ct)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          __bf16 h, l;
          split2(bv[sk][ct][j], h, l);
          bhi[ct][j] = h;
          blo[ct][j] = l;
        }
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        const bf16x8 ahi = afr[((t * KS + sk) * 2) * 64 + lane];
        const bf16x8 alo = afr[((t * KS + sk) * 2 + 1) * 64 + lane];
#pragma unroll
        for (int ct = 0; ct < kBxColTiles; ++ct) {
          acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi[ct], acc[t][ct], 0, 0, 0);
          acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo[ct], acc[t][ct], 0, 0, 0);
          acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi[ct], acc[t][ct], 0, 0, 0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      if (sk + 2 < KS) load_step(sk + 2);
    }
#pragma unroll
    for (int ct = 0; ct < kBxColTiles; ++ct) {
      const int64_t c = cbase + 16 * ct;
      if (c >= ncol) continue;
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = 16 * t + 4 * (lane >> 4) + i;
          if (r < m)
            __builtin_nontemporal_store(acc[t][ct][i], y + static_cast<int64_t>(r) * ncol + c);
        }
    }
  }
}

template <int RT, int KS>
int launch_bx(int m, int k, const double* a, int64_t ncol, const float* b,
              float* y, hipStream_t st, const int32_t* done) {
  constexpr int kBxColTiles = bx_col_tiles<RT>();
  const size_t lds = static_cast<size_t>(RT) * KS * 2 * 64 * sizeof(bf16x8);
  const int64_t npanel = (ncol + 16 * kBxColTiles - 1) / (16 * kBxColTiles);
  int64_t blocks = (npanel + kBxWaves - 1) / kBxWaves;
  if (blocks > 2048) blocks = 2048;
  k_gemm_bf16x3_lds<RT, KS><<<static_cast<unsigned>(blocks), kBxThreads, lds, st>>>(
      m, k, a, ncol, b, y, done);
  FQEHIP_CHECK_LAUNCH("k_gemm_bf16x3_lds");
  return FQEHIP_OK;
}

template <int RT>
int launch_bx_rt(int m, int k, const double* a, int64_t ncol, const float* b,
                 float* y, hipStream_t st, const int32_t* done) {
  switch ((k + 31) / 32) {
    case 1: return launch_bx<RT, 1>(m, k, a, ncol, b, y, st, done);
    case 2: return launch_bx<RT, 2>(m, k, a, ncol, b, y, st, done);
    case 3: return launch_bx<RT, 3>(m, k, a, ncol, b, y, st, done);
    case 4: return launch_bx<RT, 4>(m, k, a, ncol, b, y, st, done);
    default: return launch_bx<RT, 5>(m, k, a, ncol, b, y, st, done);
  }
}

int launch_bx_any(int m, int k, const double* a, int64_t ncol, const float* b,
                  float* y, hipStream_t st, const int32_t* done) {
  switch ((m + 15) / 16) {
    case 1: return launch_bx_rt<1>(m, k, a, ncol, b, y, st, done);
    case 2: return launch_bx_rt<2>(m, k, a, ncol, b, y, st, done);
    case 3: return launch_bx_rt<3>(m, k, a, ncol, b, y, st, done);
    case 4: return launch_bx_rt<4>(m, k, a, ncol, b, y, st, done);
    case 5: return launch_bx_rt<5>(m, k, a, ncol, b, y, st, done);
    case 6: return launch_bx_rt<6>(m, k, a, ncol, b, y, st, done);
    case 7: return launch_bx_rt<7>(m, k, a, ncol, b, y, st, done);
    case 8: return launch_bx_rt<8>(m, k, a, ncol, b, y, st, done);
    default: return launch_bx_rt<9>(m, k, a, ncol, b, y, st, done);
  }
}

}  // namespace

int launch_gemm_bf16x3(int m, int k, const double* a, int64_t ncol,
                       const float* b, float* y, hipStream_t st,
                       const int32_t* done) {
  if (m == 0 || ncol == 0) return FQEHIP_OK;
  if (k > 0 && (m + 15) / 16 <= kBxMaxRowTiles && (k + 31) / 32 <= kBxSteps)
    return launch_bx_any(m, k, a, ncol, b, y, st, done);
  dim3 grid(blocks_for(ncol, 64), (m + kRowTiles * 16 - 1) / (kRowTiles * 16));
  k_gemm_bf16x3<<<grid, kBlock, 0, st>>>(m, k, a, ncol, b, y, done);
  FQEHIP_CHECK_LAUNCH("k_gemm_bf16x3");
  return FQEHIP_OK;
}

}  // namespace fqehip
