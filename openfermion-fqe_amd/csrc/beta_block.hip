// Beta-beta block of a pair-symmetric real operator and its product with
// the CI matrix (mixed-precision split apply).
//
// For A[p][q] = A[q][p] (p, q < npair) the operator
//   H = sum_pq A[p][q] E_p E_q + sum_q A[npair][q] E_q,   E_p = E^a_p + E^b_p
// splits exactly into (alpha and beta pair operators commute)
//   H C = sum_p E^a_p ( [A (D^a + 2 D^b)][p] + A[npair][p] C )  +  C Hbb^T,
//   Hbb = sum_pq A[p][q] E^b_p E^b_q + sum_q A[npair][q] E^b_q  (lenb x lenb),
// so the per-row kernel only scatters alpha excitations and every beta
// excitation of the output becomes one dense GEMM with Hbb.  The reference
// makes the same same-spin / opposite-spin split in its lm kernels
// (src/fqe/lib/fqe_data.c:668-720 and :879-996).
//
//   k_beta_block:  Hbb column by column from the beta forward table
//                  (one workgroup per source string, column in LDS);
//   k_beta_split:  transpose and split into bfloat16 hi + lo, zero-padded
//                  to ldh x ldh (ldh a multiple of 128);
//   k_beta_gemm_bf16x3: out[Ia][Ib'] = sum_Ib C[Ia][Ib] Hbb[Ib'][Ib] on
//                  v_mfma_f32_16x16x32_bf16 with the three-product split;
//   k_beta_gemm_f64: the same product for complex128 C on
//                  v_mfma_f64_16x16x4f64, from the FP64 columns of Hbb.
#include "fqehip_common.h"

namespace fqehip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kBlockThreads = 256;

__device__ __forceinline__ __bf16 to_bf16(float x) {
  return static_cast<__bf16>(x);  // round to nearest even (v_cvt_pk_bf16_f32)
}

__device__ __forceinline__ float from_bf16(__bf16 h) {
  return __uint_as_float(static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h))
                         << 16);
}

__device__ __forceinline__ int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

// Column Ib of Hbb: LDS accumulator of lenb doubles, built in a fixed
// order so that every run gives the same bits.  For each q in turn the
// intermediate string s = E_q Ib is uniform over the workgroup; the
// off-diagonal pairs p move one electron of s, so they reach distinct
// targets t != s and are added by one thread each; the diagonal pairs
// (t = s) and the one-body term of q are summed by wave 0 in a fixed lane
// order (per-lane strided sums, then an xor butterfly) and added to s.
__global__ void __launch_bounds__(kBlockThreads)
k_beta_block(int npair, int64_t lenb, const int32_t* __restrict__ fwd,
             const double* __restrict__ amat, double* __restrict__ hcol) {
  extern __shared__ double col[];
  double* diag = col + lenb;                  // [npair] diagonal terms of q
  const int64_t ib = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  for (int64_t x = tid; x < lenb; x += kBlockThreads) col[x] = 0.0;
  __syncthreads();
  const int K = npair;
  for (int q = 0; q < K; ++q) {
    const int32_t f = fwd[static_cast<int64_t>(q) * lenb + ib];
    if (!f) continue;                         // uniform over the workgroup
    const int64_t s = iabs(f) - 1;
    const double sf = f > 0 ? 1.0 : -1.0;
    for (int p = tid; p < K; p += kBlockThreads) {
      const int32_t g = fwd[static_cast<int64_t>(p) * lenb + s];
      double v = 0.0;
      if (g) {
        const double c = (g > 0 ? sf : -sf) * amat[static_cast<int64_t>(p) * K + q];
        const int64_t t = iabs(g) - 1;
        if (t == s) v = c;
        else col[t] += c;
      }
      diag[p] = v;
    }
    __syncthreads();
    if (tid < 64) {
      double v = 0.0;
      for (int p = lane; p < K; p += 64) v += diag[p];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
      if (lane == 0) col[s] += v + sf * amat[static_cast<int64_t>(npair) * K + q];
    }
    __syncthreads();
  }
  double* dst = hcol + ib * lenb;
  for (int64_t x = tid; x < lenb; x += kBlockThreads) dst[x] = col[x];
}

// hhi/hlo[r][c] = split(Hbb[r][c]) = split(hcol[c][r]); zero outside lenb.
__global__ void __launch_bounds__(256)
k_beta_split(int64_t lenb, int64_t ldh, const double* __restrict__ hcol,
             __bf16* __restrict__ hhi, __bf16* __restrict__ hlo) {
  __shared__ double tile[32][33];
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 32;   // Hbb rows
  const int64_t c0 = static_cast<int64_t>(blockIdx.x) * 32;   // Hbb columns
  const int tx = threadIdx.x & 31;
  const int ty = threadIdx.x >> 5;                            // 0..7
  for (int y = ty; y < 32; y += 8) {
    const int64_t c = c0 + y;
    const int64_t r = r0 + tx;
    tile[y][tx] = (c < lenb && r < lenb) ? hcol[c * lenb + r] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const double v = tile[tx][y];                             // Hbb[r0+y][c0+tx]
    const __bf16 h = to_bf16(static_cast<float>(v));
    const __bf16 l = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
    const int64_t o = (r0 + y) * ldh + c0 + tx;
    hhi[o] = h;
    hlo[o] = l;
  }
}

// GEMM tile: 128 rows (alpha strings) x 128 columns (target beta strings)
// per workgroup of 4 waves in a 2 x 2 arrangement, 64 x 64 per wave, K
// (source beta strings) in steps of 32.  The complex FP32 C tile is split
// into bf16 hi/lo of its real and imaginary parts in LDS (the A operand);
// the Hbb operand is read straight from the split copies in global memory
// (lane l: row col0 + l%16, k 8*(l/16)..+7, 16 contiguous bytes).
constexpr int kGT = 128;                  // tile rows and columns
constexpr int kGK = 32;                   // K step
constexpr int kGStride = kGK + 8;         // bf16 per LDS row (conflict-free)
constexpr int kXcds = 8;                  // XCDs of an MI355X
constexpr int kGroupRows = 8;             // tile rows per group of the order

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p,
                                                       int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0,
                                           static_cast<int>(bytes), 0x00020000);
}

// XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so
// workgroup L runs on XCD L % 8.  Each XCD takes one contiguous run of
// tiles in a grouped order (kGroupRows tile rows at a time), so the C rows
// and Hbb rows it reads stay within its own L2.  Returns false for the
// padding workgroups past ta * tb.
__device__ __forceinline__ bool xcd_tile(int ta, int tb, int per_xcd,
                                         int* tile_a, int* tile_b) {
  const int xcd = static_cast<int>(blockIdx.x % kXcds);
  const int t = xcd * per_xcd + static_cast<int>(blockIdx.x / kXcds);
  if (t >= ta * tb) return false;
  const int grp = t / (kGroupRows * tb);
  const int rows_here = (ta - grp * kGroupRows) < kGroupRows
                            ? (ta - grp * kGroupRows) : kGroupRows;
  const int within = t - grp * kGroupRows * tb;
  *tile_a = grp * kGroupRows + within % rows_here;
  *tile_b = within / rows_here;
  return true;
}

__global__ void __launch_bounds__(256, 2)
k_beta_gemm_bf16x3(int64_t lena, int64_t lenb, int64_t ldh,
                   const __bf16* __restrict__ hhi,
                   const __bf16* __restrict__ hlo,
                   const float* __restrict__ coeff, float* __restrict__ out,
                   const int32_t* __restrict__ done, int ta, int tb,
                   int per_xcd) {
  // two buffers of the split C tile (hr, lr, hi, li): the next K step is
  // stored while the current one feeds the MFMAs (one barrier per step)
  __shared__ __align__(16) __bf16 sa[2][4][kGT * kGStride];
  if (done && *done) return;
  int tile_a, tile_b;
  if (!xcd_tile(ta, tb, per_xcd, &tile_a, &tile_b)) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int64_t a0 = static_cast<int64_t>(tile_a) * kGT;
  const int64_t b0 = static_cast<int64_t>(tile_b) * kGT;
  const int wa = (wave >> 1) * 64;        // wave row offset in the tile
  const int wb = (wave & 1) * 64;         // wave column offset
  const __amdgpu_buffer_rsrc_t rc = rsrc(coeff, lena * lenb * 2 * sizeof(float));
  const __amdgpu_buffer_rsrc_t rh = rsrc(hhi, ldh * ldh * 2);
  const __amdgpu_buffer_rsrc_t rl = rsrc(hlo, ldh * ldh * 2);

  // C loader: thread -> tile row tid/2, 16 complex values from k 16*(tid%2)
  const int lr = tid >> 1;
  const int lk = 16 * (tid & 1);
  const int64_t arow = a0 + lr;
  const bool rowok = arow < lena;
  float cv[32];
  auto load_c = [&](int64_t k0) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int64_t k = k0 + lk + 2 * v;
      const int off = (rowok && k < lenb)
                          ? static_cast<int>((arow * lenb + k) * 2 * sizeof(float)) : -1;
      u32x4 x;
      if (k + 1 < lenb) {
        x = __builtin_amdgcn_raw_buffer_load_b128(rc, off, 0, 0);
      } else {
        // last value of a row: a 16-byte load would reach into the next row
        const u32x2 y = __builtin_amdgcn_raw_buffer_load_b64(rc, off, 0, 0);
        x = u32x4{y[0], y[1], 0u, 0u};
      }
      cv[4 * v + 0] = __uint_as_float(x[0]);
      cv[4 * v + 1] = __uint_as_float(x[1]);
      cv[4 * v + 2] = __uint_as_float(x[2]);
      cv[4 * v + 3] = __uint_as_float(x[3]);
    }
  };
  auto store_c = [&](int buf) {
    bf16x8 hr0, lr0, hi0, li0, hr1, lr1, hi1, li1;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float re = cv[2 * j], im = cv[2 * j + 1];
      const __bf16 h = to_bf16(re), g = to_bf16(im);
      const __bf16 l = to_bf16(re - from_bf16(h)), m = to_bf16(im - from_bf16(g));
      if (j < 8) { hr0[j] = h; lr0[j] = l; hi0[j] = g; li0[j] = m; }
      else { hr1[j - 8] = h; lr1[j - 8] = l; hi1[j - 8] = g; li1[j - 8] = m; }
    }
    const int o = lr * kGStride + lk;
    *reinterpret_cast<bf16x8*>(&sa[buf][0][o]) = hr0;
    *reinterpret_cast<bf16x8*>(&sa[buf][0][o + 8]) = hr1;
    *reinterpret_cast<bf16x8*>(&sa[buf][1][o]) = lr0;
    *reinterpret_cast<bf16x8*>(&sa[buf][1][o + 8]) = lr1;
    *reinterpret_cast<bf16x8*>(&sa[buf][2][o]) = hi0;
    *reinterpret_cast<bf16x8*>(&sa[buf][2][o + 8]) = hi1;
    *reinterpret_cast<bf16x8*>(&sa[buf][3][o]) = li0;
    *reinterpret_cast<bf16x8*>(&sa[buf][3][o + 8]) = li1;
  };
  // Hbb operand registers for one K step: 4 column tiles x (hi, lo)
  bf16x8 bh[4], bl[4];
  auto load_h = [&](int64_t k0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t row = b0 + wb + 16 * c + (lane & 15);  // < ldh (padded)
      const int off = static_cast<int>((row * ldh + k0 + 8 * (lane >> 4)) * 2);
      const u32x4 xh = __builtin_amdgcn_raw_buffer_load_b128(rh, off, 0, 0);
      const u32x4 xl = __builtin_amdgcn_raw_buffer_load_b128(rl, off, 0, 0);
      bh[c] = __builtin_bit_cast(bf16x8, xh);
      bl[c] = __builtin_bit_cast(bf16x8, xl);
    }
  };

  floatx4 accr[4][4], acci[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      accr[r][c] = floatx4{0, 0, 0, 0};
      acci[r][c] = floatx4{0, 0, 0, 0};
    }

  const int64_t nk = (lenb + kGK - 1) / kGK;
  load_c(0);
  store_c(0);
  if (nk > 1) load_c(kGK);                // step 1 in flight
  __syncthreads();
  for (int64_t ks = 0; ks < nk; ++ks) {
    const int64_t k0 = ks * kGK;
    const int cb = static_cast<int>(ks & 1);
    load_h(k0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = (wa + 16 * r + (lane & 15)) * kGStride + 8 * (lane >> 4);
      const bf16x8 ahr = *reinterpret_cast<const bf16x8*>(&sa[cb][0][o]);
      const bf16x8 alr = *reinterpret_cast<const bf16x8*>(&sa[cb][1][o]);
      const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(&sa[cb][2][o]);
      const bf16x8 ali = *reinterpret_cast<const bf16x8*>(&sa[cb][3][o]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        accr[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahr, bh[c], accr[r][c], 0, 0, 0);
        acci[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bh[c], acci[r][c], 0, 0, 0);
        accr[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahr, bl[c], accr[r][c], 0, 0, 0);
        acci[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bl[c], acci[r][c], 0, 0, 0);
        accr[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alr, bh[c], accr[r][c], 0, 0, 0);
        acci[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ali, bh[c], acci[r][c], 0, 0, 0);
      }
    }
    // step ks+1 into the other buffer (its readers finished at the last
    // barrier), step ks+2 in flight, then one barrier
    if (ks + 1 < nk) {
      store_c(cb ^ 1);
      if (ks + 2 < nk) load_c(k0 + 2 * kGK);
    }
    __syncthreads();
  }
  // f32 MFMA layout: item i of lane l is row 4*(l/16)+i, column l%16
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t a = a0 + wa + 16 * r + 4 * (lane >> 4) + i;
        const int64_t b = b0 + wb + 16 * c + (lane & 15);
        if (a < lena && b < lenb) {
          float2* d = reinterpret_cast<float2*>(out + (a * lenb + b) * 2);
          *d = make_float2(accr[r][c][i], acci[r][c][i]);
        }
      }
}

// FP64 GEMM tile: 128 rows (alpha strings) x 128 columns (target beta
// strings) per workgroup of 8 waves in a 4 x 2 arrangement, 32 x 64 per
// wave (2 x 4 MFMA tiles, real and imaginary accumulators), K (source beta
// strings) in steps of 16.  Both operands are staged through LDS: the C
// tile as separate real / imaginary planes [k][row] (the A operand), the
// Hbb columns as [k][col] (the B operand, hcol[Ib][Ib'] = Hbb[Ib'][Ib] is
// already K-major).  The next K step's global loads are in flight during
// the MFMAs of the current one.
constexpr int kFK = 16;                   // K step
constexpr int kFS = kGT + 16;             // doubles per LDS row (padding)

__global__ void __launch_bounds__(512, 1)
k_beta_gemm_f64(int64_t lena, int64_t lenb, const double* __restrict__ hcol,
                const double* __restrict__ coeff, double* __restrict__ out,
                const int32_t* __restrict__ done, int ta, int tb,
                int per_xcd) {
  // two buffers of each tile: the next K step is stored while the current
  // one feeds the MFMAs (one barrier per step)
  __shared__ __align__(16) double cre[2][kFK * kFS];
  __shared__ __align__(16) double cim[2][kFK * kFS];
  __shared__ __align__(16) double hb[2][kFK * kFS];
  if (done && *done) return;
  int tile_a, tile_b;
  if (!xcd_tile(ta, tb, per_xcd, &tile_a, &tile_b)) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int64_t a0 = static_cast<int64_t>(tile_a) * kGT;
  const int64_t b0 = static_cast<int64_t>(tile_b) * kGT;
  const int wa = (wave >> 1) * 32;        // wave row offset in the tile
  const int wb = (wave & 1) * 64;         // wave column offset
  const __amdgpu_buffer_rsrc_t rc = rsrc(coeff, lena * lenb * 2 * sizeof(double));
  const __amdgpu_buffer_rsrc_t rh = rsrc(hcol, lenb * lenb * sizeof(double));

  // C loader: thread -> tile row tid/4, complex k 4*(tid%4)..+3
  const int cr = tid >> 2;
  const int ck = 4 * (tid & 3);
  const int64_t arow = a0 + cr;
  const bool rowok = arow < lena;
  // Hbb loader: thread -> k row tid/32, columns 4*(tid%32)..+3
  const int hk = tid >> 5;
  const int hn = 4 * (tid & 31);
  u32x4 cv[4];
  uint64_t hv[4];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + ck + j;
      const int off = (rowok && k < lenb)
                          ? static_cast<int>((arow * lenb + k) * 2 * sizeof(double)) : -1;
      cv[j] = __builtin_amdgcn_raw_buffer_load_b128(rc, off, 0, 0);
    }
    const int64_t k = k0 + hk;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t n = b0 + hn + j;
      const int off = (k < lenb && n < lenb)
                          ? static_cast<int>((k * lenb + n) * sizeof(double)) : -1;
      const u32x2 y = __builtin_amdgcn_raw_buffer_load_b64(rh, off, 0, 0);
      hv[j] = static_cast<uint64_t>(y[0]) | (static_cast<uint64_t>(y[1]) << 32);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t re = static_cast<uint64_t>(cv[j][0]) | (static_cast<uint64_t>(cv[j][1]) << 32);
      const uint64_t im = static_cast<uint64_t>(cv[j][2]) | (static_cast<uint64_t>(cv[j][3]) << 32);
      cre[buf][(ck + j) * kFS + cr] = __longlong_as_double(static_cast<long long>(re));
      cim[buf][(ck + j) * kFS + cr] = __longlong_as_double(static_cast<long long>(im));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      hb[buf][hk * kFS + hn + j] = __longlong_as_double(static_cast<long long>(hv[j]));
  };

  double4_t accr[2][4], acci[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      accr[r][c] = double4_t{0, 0, 0, 0};
      acci[r][c] = double4_t{0, 0, 0, 0};
    }

  const int64_t nk = (lenb + kFK - 1) / kFK;
  load(0);
  store(0);
  if (nk > 1) load(kFK);                  // step 1 in flight
  __syncthreads();
  for (int64_t ks = 0; ks < nk; ++ks) {
    const int cb = static_cast<int>(ks & 1);
    // f64 MFMA operands: A[r = l%16][k = l/16], B[k = l/16][c = l%16]
#pragma unroll
    for (int kq = 0; kq < kFK / 4; ++kq) {
      const int krow = (4 * kq + (lane >> 4)) * kFS;
      double ar[2], ai[2], bv[4];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        ar[r] = cre[cb][krow + wa + 16 * r + (lane & 15)];
        ai[r] = cim[cb][krow + wa + 16 * r + (lane & 15)];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = hb[cb][krow + wb + 16 * c + (lane & 15)];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          accr[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[r], bv[c], accr[r][c], 0, 0, 0);
          acci[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[r], bv[c], acci[r][c], 0, 0, 0);
        }
    }
    // step ks+1 into the other buffer (its readers finished at the last
    // barrier), step ks+2 in flight, then one barrier
    if (ks + 1 < nk) {
      store(cb ^ 1);
      if (ks + 2 < nk) load((ks + 2) * kFK);
    }
    __syncthreads();
  }
  // f64 MFMA layout: item i of lane l is row l/16 + 4 i, column l%16
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t a = a0 + wa + 16 * r + (lane >> 4) + 4 * i;
        const int64_t b = b0 + wb + 16 * c + (lane & 15);
        if (a < lena && b < lenb) {
          double2* d = reinterpret_cast<double2*>(out + (a * lenb + b) * 2);
          *d = make_double2(accr[r][c][i], acci[r][c][i]);
        }
      }
}

}  // namespace

bool beta_gemm_f64_shape_ok(int64_t lena, int64_t lenb) {
  const int64_t lim = int64_t(1) << 31;
  return lena * lenb * 2 * static_cast<int64_t>(sizeof(double)) < lim &&
         lenb * lenb * static_cast<int64_t>(sizeof(double)) < lim;
}

int launch_beta_gemm_f64(int64_t lena, int64_t lenb, const double* hcol,
                         const double* coeff, double* out, hipStream_t st,
                         const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  const int ta = static_cast<int>((lena + kGT - 1) / kGT);
  const int tb = static_cast<int>((lenb + kGT - 1) / kGT);
  const int per_xcd = (ta * tb + kXcds - 1) / kXcds;
  k_beta_gemm_f64<<<static_cast<unsigned>(per_xcd * kXcds), 512, 0, st>>>(
      lena, lenb, hcol, coeff, out, done, ta, tb, per_xcd);
  FQEHIP_CHECK_LAUNCH("k_beta_gemm_f64");
  return FQEHIP_OK;
}

bool beta_gemm_shape_ok(int64_t lena, int64_t lenb, int64_t ldh) {
  const int64_t lim = int64_t(1) << 31;
  return ldh >= lenb && ldh % kGT == 0 && ldh * ldh * 2 < lim &&
         lena * lenb * 2 * static_cast<int64_t>(sizeof(float)) < lim;
}

int launch_beta_gemm(int64_t lena, int64_t lenb, int64_t ldh, const void* hhi,
                     const void* hlo, const float* coeff, float* out,
                     hipStream_t st, const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  const int ta = static_cast<int>((lena + kGT - 1) / kGT);
  const int tb = static_cast<int>((lenb + kGT - 1) / kGT);
  const int per_xcd = (ta * tb + kXcds - 1) / kXcds;
  k_beta_gemm_bf16x3<<<static_cast<unsigned>(per_xcd * kXcds), 256, 0, st>>>(
      lena, lenb, ldh, static_cast<const __bf16*>(hhi),
      static_cast<const __bf16*>(hlo), coeff, out, done, ta, tb, per_xcd);
  FQEHIP_CHECK_LAUNCH("k_beta_gemm_bf16x3");
  return FQEHIP_OK;
}

}  // namespace fqehip

using namespace fqehip;

extern "C" int64_t fqehip_beta_block_ld(int64_t lenb) {
  return lenb <= 0 ? 0 : ((lenb + kGT - 1) / kGT) * kGT;
}

extern "C" int fqehip_beta_block(int npair, int64_t lenb, const int32_t* d_fwd_b,
                                 const double* d_a, double* d_hcol,
                                 uint16_t* d_hhi, uint16_t* d_hlo, int64_t ldh,
                                 void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(npair >= 0 && lenb >= 0, "npair=%d lenb=%lld", npair,
                   static_cast<long long>(lenb));
  FQEHIP_CHECK_ARG(ldh == fqehip_beta_block_ld(lenb),
                   "ldh=%lld, expected %lld", static_cast<long long>(ldh),
                   static_cast<long long>(fqehip_beta_block_ld(lenb)));
  FQEHIP_CHECK_ARG((lenb + npair) * static_cast<int64_t>(sizeof(double)) <= 160 * 1024,
                   "lenb=%lld exceeds the LDS column of the beta block",
                   static_cast<long long>(lenb));
  if (lenb == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_a && d_hcol && d_hhi && d_hlo, "NULL pointer");
  FQEHIP_CHECK_ARG(npair == 0 || d_fwd_b, "NULL beta forward table");
  hipStream_t st = as_stream(stream);
  // npair == 0 (no orbitals) leaves the LDS column zero: Hbb = 0
  k_beta_block<<<static_cast<unsigned>(lenb), kBlockThreads,
                 (lenb + npair) * sizeof(double), st>>>(npair, lenb, d_fwd_b, d_a,
                                                        d_hcol);
  FQEHIP_CHECK_LAUNCH("k_beta_block");
  dim3 grid(static_cast<unsigned>(ldh / 32), static_cast<unsigned>(ldh / 32));
  k_beta_split<<<grid, 256, 0, st>>>(lenb, ldh, d_hcol,
                                      reinterpret_cast<__bf16*>(d_hhi),
                                      reinterpret_cast<__bf16*>(d_hlo));
  FQEHIP_CHECK_LAUNCH("k_beta_split");
  return FQEHIP_OK;
}

extern "C" int fqehip_beta_gemm(int64_t lena, int64_t lenb,
                                const uint16_t* d_hhi, const uint16_t* d_hlo,
                                int64_t ldh, const void* d_coeff, void* d_out,
                                const int32_t* d_done, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(lena >= 0 && lenb >= 0, "negative size");
  FQEHIP_CHECK_ARG(beta_gemm_shape_ok(lena, lenb, ldh),
                   "ldh=%lld does not fit lenb=%lld (multiple of %d needed)",
                   static_cast<long long>(ldh), static_cast<long long>(lenb), kGT);
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_hhi && d_hlo && d_coeff && d_out, "NULL pointer");
  return launch_beta_gemm(lena, lenb, ldh, d_hhi, d_hlo,
                          static_cast<const float*>(d_coeff),
                          static_cast<float*>(d_out), as_stream(stream), d_done);
}

extern "C" int fqehip_beta_gemm_f64(int64_t lena, int64_t lenb,
                                    const double* d_hcol, const void* d_coeff,
                                    void* d_out, const int32_t* d_done,
                                    void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(lena >= 0 && lenb >= 0, "negative size");
  FQEHIP_CHECK_ARG(beta_gemm_f64_shape_ok(lena, lenb),
                   "lena=%lld lenb=%lld too large for 32-bit offsets",
                   static_cast<long long>(lena), static_cast<long long>(lenb));
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_hcol && d_coeff && d_out, "NULL pointer");
  return launch_beta_gemm_f64(lena, lenb, d_hcol,
                              static_cast<const double*>(d_coeff),
                              static_cast<double*>(d_out), as_stream(stream),
                              d_done);
}
