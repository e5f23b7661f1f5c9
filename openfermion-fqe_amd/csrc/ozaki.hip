// Ozaki-digit variant of the complex128 split apply: the same operands,
// layouts and scatter as the FP64 split path (complex128 C, FP64 row
// accumulators; the slot rows in 48 bits per double, 12 bytes per complex
// value, stream_store_p48), with the pair contraction Y = A' D' of the alpha
// row kernel computed as exact integer products of 8-bit digits on the
// matrix cores (v_mfma_i32_16x16x64_i8, int32 accumulation) instead of
// FP64 matrix-core products.
//
// Digits.  Per alpha string ia, A' (the live pair rows of the operator) is
// scaled row by row by a power of two so that |A' 2^ea| < 2^38, rounded to
// the integer z and split exactly into five balanced signed digits
// a_0..a_4 in [-128, 127] of weights 2^32, 2^24, ..., 1 (from the low end:
// d = ((z + 128) mod 256) - 128, z = (z - d) / 256).
// D' (gathered from C) is scaled by one power of two per alpha string,
// 2^ed with ed from a bound on |D'| (the largest row maximum of |C| over the
// alpha sources of ia, plus twice the row maximum of row ia), so that
// |D' 2^ed| < 2^46, and rounded to a 48-bit integer y.  The six bytes of y
// (two's complement) are the digits d_0..d_5 of weights 1 .. 2^40.
// d_0..d_4 are stored as unsigned bytes minus 128 (one XOR), d_5 as the
// signed top byte, so no carry propagation is needed; the 128 offsets are
// added back exactly, for every digit pair, through the row sums of the A'
// integers.
//
// Products.  Of the 30 digit pairs, the 20 with i - t >= 0 (weight >= 2^32
// of a maximum of 2^72) are formed, six int32 accumulators per output (one
// per i - t); the dropped pairs weigh at most 2^-40 of the largest product
// each, and since the offsets of all pairs are corrected, they enter as
// sums of zero-mean terms: about 1e-11 of the result when the bound is a
// few times the gathered values, and still below 1e-9 for vectors whose
// weight sits on a few determinants.  Every int32 sum is exact (at most
// 5 * 128 * 128 * 128 < 2^24 per accumulator).  The result is assembled in
// FP64: Y = 2^-(ea+ed) (2^32 sum_m 2^(8m) acc_m + 128 (1 + 2^8 + 2^16 +
// 2^24 + 2^32) sum_q z_q).
//
// Reference: the opposite-spin and alpha same-spin terms of
// lm_apply_array12_same_spin_opt / lm_apply_array12_diff_spin_opt
// (src/fqe/lib/fqe_data.c:668-720, :879-996), in complex128.
#include "fqehip_common.h"
#include "rows_common.h"

namespace fqehip {
namespace {

using namespace rows;

typedef int int4v __attribute__((ext_vector_type(4)));

constexpr int kOzKPad = 128;               // gathered rows (pairs + C), padded
constexpr int kOzKSteps = kOzKPad / 64;    // k steps of the i8 MFMA
constexpr int kOzADig = 5;                 // digits of A'
constexpr int kOzDDig = 6;                 // digits 0..5 of the D' integers
constexpr int kOzARow = kOzKPad + 16;      // bytes per LDS row of a digit plane
constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52: rint via one add

// rowmax[I] = max over the row of max(|Re C[I][.]|, |Im C[I][.]|).
__global__ void __launch_bounds__(256)
k_row_absmax(int64_t lena, int64_t lenb, const double* __restrict__ coeff,
             double* __restrict__ rowmax, const int32_t* __restrict__ done) {
  __shared__ double red[4];
  if (done && *done) return;
  for (int64_t r = blockIdx.x; r < lena; r += gridDim.x) {
    const double* row = coeff + r * lenb * 2;
    double m = 0.0;
    for (int64_t k = threadIdx.x; k < 2 * lenb; k += 256) m = fmax(m, fabs(row[k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      rowmax[r] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
  }
}

// Digits 0..5 of four 48-bit integers y_j held as the (lo, hi) words of
// kMagic + y_j: out[i] byte j = digit i of y_j (bytes 0..3 of lo, bytes 0..1
// of hi); digits 0..4 are unsigned bytes offset by -128 (one XOR), digit 5
// is the signed top byte.  A two-stage byte transpose with v_perm_b32
// (perm(a, b, sel): result byte k = byte sel_k of the 8 bytes b0..b3 a0..a3).
__device__ __forceinline__ void digits_of_four(const uint32_t (&lo)[4],
                                               const uint32_t (&hi)[4],
                                               uint32_t (&out)[6]) {
  // x01 = [lo0.b0, lo1.b0, lo0.b1, lo1.b1], y01 = [lo0.b2, lo1.b2, lo0.b3, lo1.b3]
  const uint32_t x01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x05010400u);
  const uint32_t y01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x07030602u);
  const uint32_t x23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x05010400u);
  const uint32_t y23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x07030602u);
  const uint32_t z01 = __builtin_amdgcn_perm(hi[1], hi[0], 0x05010400u);
  const uint32_t z23 = __builtin_amdgcn_perm(hi[3], hi[2], 0x05010400u);
  out[0] = __builtin_amdgcn_perm(x23, x01, 0x05040100u) ^ 0x80808080u;
  out[1] = __builtin_amdgcn_perm(x23, x01, 0x07060302u) ^ 0x80808080u;
  out[2] = __builtin_amdgcn_perm(y23, y01, 0x05040100u) ^ 0x80808080u;
  out[3] = __builtin_amdgcn_perm(y23, y01, 0x07060302u) ^ 0x80808080u;
  out[4] = __builtin_amdgcn_perm(z23, z01, 0x05040100u) ^ 0x80808080u;
  out[5] = __builtin_amdgcn_perm(z23, z01, 0x07060302u);
}

__device__ __forceinline__ void split_words(double v, uint32_t& lo,
                                            uint32_t& hi) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
  lo = static_cast<uint32_t>(b);
  hi = static_cast<uint32_t>(b >> 32);
}

__device__ __forceinline__ double as_double(uint32_t lo, uint32_t hi) {
  return __longlong_as_double(static_cast<long long>(
      static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32)));
}

// Offset correction constant: the -128 offsets of data digits 0..4 over
// every digit pair, 128 (1 + 2^8 + 2^16 + 2^24 + 2^32) per unit of the
// operator integers' sum.
constexpr double kOffsetUnit = 128.0 * 4311810305.0;
// Invalid alpha gather offset, in 16-byte units: (kNoRow16 + column) * 16
// lies past the end of C (lena * lenb * 16 <= 0x7ff00000 is a launch
// requirement) and below 2^31 (lenb < 2^16).
constexpr uint32_t kNoRow16 = 0x7ff0000u;
// Beta gather table entry without an excitation (global-memory row): a
// byte offset past the end of every row of C.
constexpr uint32_t kNoCol = 0x7ffffff0u;

// Beta gather table of the Ozaki row kernel, column-major and padded to
// kOzKPad entries per column: tab[jb][q] = byte offset of the source column
// of beta excitation q of column jb within its row (16 (|ptb| - 1)), with
// the sign of the excitation in bit 31; `none` where there is none
// (ptb == 0, or q >= npair).
__global__ void __launch_bounds__(256)
k_beta_table_oz(int npair, int64_t lenb, const int32_t* __restrict__ ptb,
                uint32_t none, uint32_t* __restrict__ tab,
                const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t total = lenb * kOzKPad;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
       i < total; i += static_cast<int64_t>(gridDim.x) * 256) {
    const int64_t jb = i / kOzKPad;
    const int q = static_cast<int>(i % kOzKPad);
    uint32_t w = none;
    if (q < npair) {
      const int32_t b = ptb[static_cast<int64_t>(q) * lenb + jb];
      if (b)
        w = (b < 0 ? 0x80000000u : 0u) |
            (static_cast<uint32_t>(iabs(b) - 1) << 4);
    }
    tab[i] = w;
  }
}

// Digits of one quarter row (32 consecutive k) of an operator row, by a
// group of 4 consecutive lanes (one quarter each): the row's scale 2^ea with
// |x 2^ea| < 2^38 from the group's maximum, the exact balanced digits of the
// integers z = rint(x 2^ea) (from the low end, d = ((z + 128) mod 256) -
// 128, z = (z - d) / 256; w[t] holds digit t, weight 2^(8(4 - t)), of k
// 4x..4x+3 in word x), and the row sum of the z.
__device__ __forceinline__ void quarter_row_digits(const double (&v)[32],
                                                   uint32_t (&w)[kOzADig][8],
                                                   int& ea, double& zsum) {
  double amax = 0.0;
#pragma unroll
  for (int j = 0; j < 32; ++j) amax = fmax(amax, fabs(v[j]));
  amax = fmax(amax, __shfl_xor(amax, 1));
  amax = fmax(amax, __shfl_xor(amax, 2));
  ea = 0;
  if (amax > 0.0) {
    int e;
    (void)frexp(amax, &e);
    ea = 38 - e;
  }
  zsum = 0.0;
#pragma unroll
  for (int t = 0; t < kOzADig; ++t)
#pragma unroll
    for (int x = 0; x < 8; ++x) w[t][x] = 0u;
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    int64_t z = static_cast<int64_t>(rint(ldexp(v[j], ea)));
    zsum += static_cast<double>(z);
#pragma unroll
    for (int t = kOzADig - 1; t >= 0; --t) {
      const int64_t d = t ? ((z + 128) & 255) - 128 : z;
      z = (z - d) >> 8;
      w[t][j >> 2] |= (static_cast<uint32_t>(d) & 0xffu) << (8 * (j & 3));
    }
  }
  zsum += __shfl_xor(zsum, 1);
  zsum += __shfl_xor(zsum, 2);
}

// A' digits of every pair row p < npair (the same for every alpha string:
// A[p][0..npair-1] and the h1' entry A[npair][p] at k = npair, zero beyond),
// once per apply: pdig[p][t][k] bytes, pea[p] = ea, pzs[p] = sum of z.
__global__ void __launch_bounds__(256)
k_pair_digits_oz(int npair, const double* __restrict__ amat,
                 uint8_t* __restrict__ pdig, int32_t* __restrict__ pea,
                 double* __restrict__ pzs, const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int p = blockIdx.x * 64 + (threadIdx.x >> 2);
  const int k0 = (threadIdx.x & 3) * 32;
  const int K = npair;
  const bool ok = p < npair;
  double v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int k = k0 + j;
    double x = 0.0;
    if (ok) {
      if (k < K) x = amat[static_cast<int64_t>(p) * K + k];
      else if (k == K) x = amat[static_cast<int64_t>(npair) * K + p];
    }
    v[j] = x;
  }
  uint32_t w[kOzADig][8];
  int ea;
  double zsum;
  quarter_row_digits(v, w, ea, zsum);
  if (!ok) return;
#pragma unroll
  for (int t = 0; t < kOzADig; ++t) {
    uint4* dst = reinterpret_cast<uint4*>(pdig + (static_cast<int64_t>(p) * kOzADig + t) * kOzKPad + k0);
    dst[0] = make_uint4(w[t][0], w[t][1], w[t][2], w[t][3]);
    dst[1] = make_uint4(w[t][4], w[t][5], w[t][6], w[t][7]);
  }
  if ((threadIdx.x & 3) == 0) {
    pea[p] = ea;
    pzs[p] = zsum;
  }
}

constexpr int kRowThreads = 512;           // 8 waves, 16 columns each
constexpr int kRowCols = 128;              // columns per chunk
constexpr int64_t kDstNone = -1;           // padding row: no store
constexpr int64_t kDstRowacc = -2;         // folded row: into rowacc

// Alpha rows of the split apply (see k_split_cols_f64 in fused.hip for the
// operands), one alpha string per block at a time.  Set-up: the live pair
// rows' A' digits go to LDS and, with ROW_LDS, row ia of C too (plus one
// zero entry at index lenb, the target of absent beta excitations); then
// each wave owns 16 columns of a 128-column chunk, gathers their D' values
// (lane l: column l%16, k = 64 s + 16 (l/16) + j) four k values at a time
// with the next four's loads in flight - the alpha part from global memory
// (absent sources read out of range: no memory access), the beta part from
// the LDS row (random 16-byte reads, which cost about a seventh of a
// random global gather) or, without ROW_LDS, from global memory - forms
// the six digit operands in registers and runs the 20 digit products
// (weight classes m = i - t = 0..5) of each of the four 16-row tiles, the
// real and imaginary parts on the same A' operands.  The next chunk's beta
// table words load during the products.  i8 MFMA layouts: lane l holds
// A[row l%16][k 16 (l/16) + j] and B[k 16 (l/16) + j][col l%16] (byte j of
// its four words); output item e of lane l is [row 4 (l/16) + e][col l%16].
// Work items: alpha strings 0 .. nwhole - 1 whole (in XCD order), then each
// of the remaining strings as two items, its first and second half of the
// column chunks (so a last, partial round of strings over the CUs takes
// half as long).
template <bool ROW_LDS>
__global__ void __launch_bounds__(kRowThreads)
k_split_rows_oz(int norb, int compressed, int npair, int64_t lena,
                int64_t nwhole, int64_t lenb, const double* __restrict__ amat,
                const int32_t* __restrict__ pta, const uint32_t* __restrict__ btab,
                const uint8_t* __restrict__ pdig, const int32_t* __restrict__ pea,
                const double* __restrict__ pzs,
                const int32_t* __restrict__ fwa, const int32_t* __restrict__ sla,
                int nslot, const double* __restrict__ coeff,
                const double* __restrict__ rowmax, uint32_t* __restrict__ slots,
                double* __restrict__ rowacc, const int32_t* __restrict__ done) {
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  const int Kg = K + 1;                     // gathered rows: pairs + C
  uint8_t* adig = reinterpret_cast<uint8_t*>(smem);  // [5][kLiveRows][kOzARow]
  double* gsa = reinterpret_cast<double*>(adig + kOzADig * kLiveRows * kOzARow);
  uint32_t* goff = reinterpret_cast<uint32_t*>(gsa + kOzKPad);  // [kOzKPad]
  double* drow = reinterpret_cast<double*>(goff + kOzKPad);     // [kOzKPad]
  double* rscale = drow + kOzKPad;          // [kLiveRows] +-2^(32 - ea - ed)
  double* corr = rscale + kLiveRows;        // [kLiveRows] offset correction
  // [kLiveRows] destination of a live row: word offset of its slot row,
  // kDstRowacc (the folded row) or kDstNone (padding)
  int64_t* rdst = reinterpret_cast<int64_t*>(corr + kLiveRows);
  double* red = reinterpret_cast<double*>(rdst + kLiveRows);  // [8] block scratch
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(red + 8);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  int32_t* lpair = rt.ssgn + kRowTabs;      // [kLiveRows] pair of live row
  uint8_t* crow = reinterpret_cast<uint8_t*>(lpair + kLiveRows);  // [lenb + 1] x 16 B

  const int nchunks = static_cast<int>((lenb + kRowCols - 1) / kRowCols);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(double));
  const __amdgpu_buffer_rsrc_t rtb = buffer_of(btab, lenb * kOzKPad * 4);
  const int kq = lane >> 4;                 // k group of the lane
  const int cl = lane & 15;                 // column within the wave's 16

  if (ROW_LDS && tid == 0)
    *reinterpret_cast<u32x4*>(crow + 16 * lenb) = u32x4{0u, 0u, 0u, 0u};

  // beta table words of the lane's 32 k values in a chunk:
  // tw[s][v][u] for k = 64 s + 16 kq + 4 v + u
  u32x4 tw[kOzKSteps][4];
  auto load_table = [&](int chunk) {
    const int64_t jb = static_cast<int64_t>(chunk) * kRowCols + wave * 16 + cl;
    const int base = static_cast<int>(jb < lenb ? jb : 0) * (kOzKPad * 4) + 64 * kq;
#pragma unroll
    for (int s = 0; s < kOzKSteps; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        tw[s][v] = __builtin_amdgcn_raw_buffer_load_b128(rtb, base + 256 * s + 16 * v, 0, 0);
  };
  const int64_t nitems = nwhole + 2 * (lena - nwhole);
  const int hchunks = (nchunks + 1) / 2;    // chunks of a first half
  auto first_chunk = [&](int64_t it) {
    return (it >= nwhole && ((it - nwhole) & 1)) ? hchunks : 0;
  };
  load_table(first_chunk(blockIdx.x));

  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int64_t ia = it < nwhole ? xcd_order(it, nwhole) : nwhole + ((it - nwhole) >> 1);
    const int c0 = first_chunk(it);
    const int c1 = (it < nwhole || c0 > 0) ? nchunks : hchunks;
    const int next_c0 = first_chunk(it + gridDim.x);
    load_row_tables<double>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                            kRowThreads, K);
    if (ROW_LDS) {
      const double* src = coeff + ia * lenb * 2;
      for (int64_t j = tid; j < lenb; j += kRowThreads)
        *reinterpret_cast<double2*>(crow + 16 * j) =
            *reinterpret_cast<const double2*>(src + 2 * j);
    }
    __syncthreads();
    if (wave == 0) {
      // compact the live rows: pairs exciting ia to another string, then
      // the folded row npair
      int base = 0;
      for (int p0 = 0; p0 < K; p0 += 64) {
        const int p = p0 + lane;
        const bool live = p < K && rt.sbase[p] >= 0;
        const uint64_t m = __ballot(live);
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (live && pos < kLiveRows) lpair[pos] = p;
        base += __popcll(m);
      }
      if (lane >= base) lpair[lane] = lane == base ? npair : -1;
    } else if (wave == 1) {
      // bound on |D'|: alpha sources (sign 1) plus twice the beta part,
      // which gathers within row ia
      double m = 0.0;
      for (int q = lane; q < Kg; q += 64) {
        const int32_t off = rt.aoff[q];
        if (off >= 0) m = fmax(m, rowmax[off / (2 * lenb)]);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
      if (lane == 0) red[0] = m + 2.0 * rowmax[ia];
    } else if (tid >= 256 && tid < 256 + kOzKPad) {
      const int k = tid - 256;
      double v = 0.0;
      if (k < Kg) {
        for (int i = 0; i < norb; ++i) {
          const int dp = compressed ? i + i * (i + 1) / 2 : i * (norb + 1);
          if (rt.sbase[dp] == -2)
            v += k < K ? amat[static_cast<int64_t>(dp) * K + k]
                       : amat[static_cast<int64_t>(npair) * K + dp];
        }
      }
      drow[k] = v;
    }
    __syncthreads();
    int ed = 0;
    {
      const double bound = red[0];
      if (bound > 0.0) {
        int e;
        (void)frexp(bound, &e);
        ed = 46 - e;                        // |D' 2^ed| < 2^46
      }
    }
    const double dsc = ldexp(1.0, ed);
    if (tid >= 256 && tid < 256 + kOzKPad) {
      const int q = tid - 256;
      const int32_t off = q < Kg ? rt.aoff[q] : -1;
      goff[q] = off >= 0 ? static_cast<uint32_t>(off) >> 1 : kNoRow16;
      gsa[q] = off >= 0 ? static_cast<double>(rt.asgn[q]) * dsc : 0.0;
    } else if (tid < 256) {
      // live pair rows: A' digits of the apply's pair table (thread -> live
      // row tid/4, k in [32 (tid%4), +32)); padding rows zero; the folded
      // row is formed below
      const int m = tid >> 2;
      const int k0 = (tid & 3) * 32;
      const int p = lpair[m];
      if (p != npair) {
        uint4 w[kOzADig][2];
#pragma unroll
        for (int t = 0; t < kOzADig; ++t) {
          if (p >= 0) {
            const uint4* src = reinterpret_cast<const uint4*>(
                pdig + (static_cast<int64_t>(p) * kOzADig + t) * kOzKPad + k0);
            w[t][0] = src[0];
            w[t][1] = src[1];
          } else {
            w[t][0] = w[t][1] = make_uint4(0u, 0u, 0u, 0u);
          }
        }
#pragma unroll
        for (int t = 0; t < kOzADig; ++t) {
          uint4* dst = reinterpret_cast<uint4*>(adig + (t * kLiveRows + m) * kOzARow + k0);
          dst[0] = w[t][0];
          dst[1] = w[t][1];
        }
        if ((tid & 3) == 0) {
          // Y = 2^(32 - ea - ed) t + 2^-(ea + ed) corr (powers of two: exact),
          // times the sign of the scatter to the slot row
          const double sc = p >= 0 ? ldexp(1.0, -(pea[p] + ed)) * rt.ssgn[p] : 0.0;
          corr[m] = p >= 0 ? kOffsetUnit * pzs[p] * sc : 0.0;
          rscale[m] = 4294967296.0 * sc;
          rdst[m] = p >= 0 ? 3 * (rt.sbase[p] >> 1) : kDstNone;
        }
      }
    } else if (wave == 6) {
      // the folded row (sum of the rows of ia's occupied diagonal pairs):
      // lanes 4 g .. 4 g + 3 of group g < 4 take its quarters
      const uint64_t fm = __ballot(lpair[lane] == npair);
      const int m = fm ? __ffsll(static_cast<unsigned long long>(fm)) - 1 : -1;
      if (m >= 0 && lane < 4) {
        const int k0 = lane * 32;
        double v[32];
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = drow[k0 + j];
        uint32_t w[kOzADig][8];
        int ea;
        double zsum;
        quarter_row_digits(v, w, ea, zsum);
#pragma unroll
        for (int t = 0; t < kOzADig; ++t) {
          uint4* dst = reinterpret_cast<uint4*>(adig + (t * kLiveRows + m) * kOzARow + k0);
          dst[0] = make_uint4(w[t][0], w[t][1], w[t][2], w[t][3]);
          dst[1] = make_uint4(w[t][4], w[t][5], w[t][6], w[t][7]);
        }
        if (lane == 0) {
          const double sc = ldexp(1.0, -(ea + ed));
          corr[m] = kOffsetUnit * zsum * sc;
          rscale[m] = 4294967296.0 * sc;
          rdst[m] = kDstRowacc;
        }
      }
    }
    __syncthreads();
    // row ia of C in global memory (beta gathers without ROW_LDS; offsets
    // past the row read 0)
    const __amdgpu_buffer_rsrc_t rrow = buffer_of(coeff + ia * lenb * 2, lenb * 16);
    const double dsc2 = 2.0 * dsc;

    for (int chunk = c0; chunk < c1; ++chunk) {
      const int64_t jb = static_cast<int64_t>(chunk) * kRowCols + wave * 16 + cl;
      const bool colok = jb < lenb;
      // padding columns gather column 0 (valid memory) and store nothing
      const uint32_t ib = static_cast<uint32_t>(colok ? jb : 0);
      int4v dd[2][kOzDDig][kOzKSteps];      // [re, im][digit][k step]
      // batch b: k = 64 (b / 4) + 16 kq + 4 (b % 4) + u, u = 0..3
      // batch b: digits of D' for k = 64 (b / 4) + 16 kq + 4 (b % 4) + u
      auto digits = [&](int b, const u32x4 (&A)[4], const u32x4 (&B)[4]) {
        const int s = b >> 2, v = b & 3;
        const int q0 = 64 * s + 16 * kq + 4 * v;
        uint32_t lre[4], hre[4], lim[4], him[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t sg = tw[s][v][u] & 0x80000000u;
          const double sa = gsa[q0 + u];
          const double vre = fma(sa, as_double(A[u][0], A[u][1]),
                                 fma(dsc2, as_double(B[u][0], B[u][1] ^ sg), kMagic));
          const double vim = fma(sa, as_double(A[u][2], A[u][3]),
                                 fma(dsc2, as_double(B[u][2], B[u][3] ^ sg), kMagic));
          split_words(vre, lre[u], hre[u]);
          split_words(vim, lim[u], him[u]);
        }
        uint32_t ore[kOzDDig], oim[kOzDDig];
        digits_of_four(lre, hre, ore);
        digits_of_four(lim, him, oim);
#pragma unroll
        for (int i = 0; i < kOzDDig; ++i) {
          // the digit words exist here, before the next batch's loads
          // (otherwise the conversions sink to the MFMAs and every
          // batch's loaded values stay live at once)
          asm volatile("" : "+v"(ore[i]), "+v"(oim[i]));
          dd[0][i][s][v] = static_cast<int>(ore[i]);
          dd[1][i][s][v] = static_cast<int>(oim[i]);
        }
        asm volatile("" ::: "memory");
      };
      auto alpha_load = [&](int b, u32x4 (&A)[4]) {
        const int q0 = 64 * (b >> 2) + 16 * kq + 4 * (b & 3);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          A[u] = __builtin_amdgcn_raw_buffer_load_b128(
              rc, static_cast<int>((goff[q0 + u] + ib) << 4), 0, 0);
      };
      if constexpr (ROW_LDS) {
        // global alpha gathers two batches ahead; the beta gathers from
        // the LDS row at use
        u32x4 la[3][4];
        alpha_load(0, la[0]);
        alpha_load(1, la[1]);
#pragma unroll
        for (int b = 0; b < 4 * kOzKSteps; ++b) {
          if (b + 2 < 4 * kOzKSteps) alpha_load(b + 2, la[(b + 2) % 3]);
          const int s = b >> 2, v = b & 3;
          u32x4 lb[4];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            lb[u] = *reinterpret_cast<const u32x4*>(crow + (tw[s][v][u] & 0x7fffffffu));
          digits(b, la[b % 3], lb);
        }
      } else {
        u32x4 la[2][4], lb[2][4];
        auto issue = [&](int b) {
          const int s = b >> 2, v = b & 3;
          alpha_load(b, la[b & 1]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            lb[b & 1][u] = __builtin_amdgcn_raw_buffer_load_b128(
                rrow, static_cast<int>(tw[s][v][u] & 0x7fffffffu), 0, 0);
        };
        issue(0);
#pragma unroll
        for (int b = 0; b < 4 * kOzKSteps; ++b) {
          if (b + 1 < 4 * kOzKSteps) issue(b + 1);
          digits(b, la[b & 1], lb[b & 1]);
        }
      }
      // the next chunk's table (the next item's first chunk after the last)
      load_table(chunk + 1 < c1 ? chunk + 1 : next_c0);
      // the products at a raised wave priority: the SIMD's other wave,
      // gathering, issues in the gaps instead of competing for issue with
      // the MFMA stream (measured: 4.21 -> 4.13 ms at norb 14; priority
      // during the gather instead: no gain)
      __builtin_amdgcn_s_setprio(1);
#pragma unroll 1
      for (int r = 0; r < kLiveRows / 16; ++r) {
        int4v acc[2][6];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int m = 0; m < 6; ++m) acc[c][m] = int4v{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < kOzKSteps; ++s) {
          int4v a[kOzADig];
#pragma unroll
          for (int t = 0; t < kOzADig; ++t)
            a[t] = *reinterpret_cast<const int4v*>(
                adig + (t * kLiveRows + 16 * r + cl) * kOzARow + 64 * s + 16 * kq);
          // pairs (data digit i, operator digit t) with m = i - t >= 0
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < kOzDDig; ++i)
#pragma unroll
              for (int t = 0; t < kOzADig; ++t)
                if (t <= i)
                  acc[c][i - t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                      a[t], dd[c][i][s], acc[c][i - t], 0, 0, 0);
        }
        if (colok) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = 16 * r + 4 * kq + e;
            double y[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              // classes 5 and 4, and 3 and 2, merged in int32: with K <= 128,
              // |acc5| <= 65 * 64 * 128 < 2^19.1, |acc4| < 2^21.1,
              // |acc3| < 2^22.1, |acc2| < 2^22.7, so both merges stay below
              // 2^31
              const int32_t m54 = acc[c][5][e] * 256 + acc[c][4][e];
              const int32_t m32 = acc[c][3][e] * 256 + acc[c][2][e];
              double t = static_cast<double>(m54);
              t = fma(t, 65536.0, static_cast<double>(m32));
              t = fma(t, 256.0, static_cast<double>(acc[c][1][e]));
              t = fma(t, 256.0, static_cast<double>(acc[c][0][e]));
              y[c] = fma(t, rscale[row], corr[row]);
            }
            const int64_t dst = rdst[row];
            if (dst >= 0) {
              stream_store_p48(slots + dst + 3 * jb, y[0], y[1]);
            } else if (dst == kDstRowacc) {
              double2* d = reinterpret_cast<double2*>(rowacc + ia * lenb * 2 + 2 * jb);
              double2 o = *d;
              o.x += y[0];
              o.y += y[1];
              *d = o;
            }
          }
        }
      }
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
  }
}

size_t split_rows_oz_lds_bytes(bool row_lds, int64_t lenb) {
  return static_cast<size_t>(kOzADig) * kLiveRows * kOzARow +
         kOzKPad * (sizeof(double) + sizeof(uint32_t)) +
         (kOzKPad + 3 * kLiveRows + 8) * sizeof(double) +
         kRowTabs * (sizeof(int64_t) + 3 * sizeof(int32_t)) +
         kLiveRows * sizeof(int32_t) +
         (row_lds ? static_cast<size_t>(lenb + 1) * 16 : 0);
}

constexpr size_t kLdsLimit = 160 * 1024;

}  // namespace

// Alpha rows of the Ozaki-digit split apply (requirements as for the FP64
// split kernel: npair + 1 <= 128, norb^2/4 + 1 <= 64 live rows, 32-bit byte
// offsets); rowmax: the row maxima of |C| (launch_row_absmax first).
bool split_oz_supported(int norb, int npair, int64_t lena, int64_t lenb) {
  return npair + 1 <= kOzKPad && norb * norb / 4 + 1 <= kLiveRows &&
         lenb < (1 << 16) &&
         lena * lenb * 16 <= static_cast<int64_t>(kNoRow16) * 16;
}

// Bytes of the tables of launch_split_oz: the beta gather table (lenb x
// 128 words), then the pair rows' A' digits (128 x 5 x 128 bytes), their
// scale exponents (128 int32) and integer sums (128 doubles).
constexpr int64_t kPairDigitBytes = kOzKPad * (kOzADig * kOzKPad + 4 + 8);
int64_t split_oz_table_bytes(int64_t lenb) {
  return lenb * kOzKPad * 4 + kPairDigitBytes;
}

// build_tables false: btab already holds the beta table and the pair rows'
// digits of this sector and operator (an earlier call with the same btab).
int launch_split_oz(int norb, int npair, int64_t lena, int64_t lenb,
                    const double* a, const int32_t* pta, const int32_t* ptb,
                    const int32_t* fwa, const int32_t* sla, int nslot,
                    const double* coeff, const double* rowmax, void* btab,
                    uint32_t* slots, double* rowacc, bool build_tables,
                    hipStream_t st, const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  if (!split_oz_supported(norb, npair, lena, lenb)) {
    set_error("the Ozaki split apply needs npair + 1 <= %d, norb^2/4 + 1 <= %d, "
              "lenb < 65536 and 16 lena lenb <= %lld",
              kOzKPad, kLiveRows, static_cast<long long>(kNoRow16) * 16);
    return FQEHIP_ERR_INVALID;
  }
  uint32_t* tab = static_cast<uint32_t*>(btab);
  uint8_t* pdig = reinterpret_cast<uint8_t*>(tab + lenb * kOzKPad);
  int32_t* pea = reinterpret_cast<int32_t*>(pdig + kOzKPad * kOzADig * kOzKPad);
  double* pzs = reinterpret_cast<double*>(pea + kOzKPad);
  const bool row_lds = split_rows_oz_lds_bytes(true, lenb) <= kLdsLimit;
  if (build_tables) {
    k_pair_digits_oz<<<(npair + 63) / 64, 256, 0, st>>>(npair, a, pdig, pea, pzs, done);
    FQEHIP_CHECK_LAUNCH("k_pair_digits_oz");
    const int64_t entries = lenb * kOzKPad;
    const unsigned tgrid = static_cast<unsigned>(std::min<int64_t>((entries + 255) / 256, 4096));
    // absent beta excitations: the zero entry after the LDS row, or past the
    // end of the row in global memory
    const uint32_t none = row_lds ? static_cast<uint32_t>(16 * lenb) : kNoCol;
    k_beta_table_oz<<<tgrid, 256, 0, st>>>(npair, lenb, ptb, none, tab, done);
    FQEHIP_CHECK_LAUNCH("k_beta_table_oz");
  }
  const int compressed = npair == norb * (norb + 1) / 2 ? 1 : 0;
  // one workgroup per CU at a time: strings beyond the last full round over
  // the CUs run as column halves when they fill at most half a round
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_error("launch_split_oz: no device attributes");
    return FQEHIP_ERR_HIP;
  }
  const int64_t rem = ncu > 0 ? lena % ncu : 0;
  const int64_t nchunks = (lenb + kRowCols - 1) / kRowCols;
  const int64_t nwhole = (nchunks >= 2 && 2 * rem <= ncu) ? lena - rem : lena;
  const int64_t nitems = nwhole + 2 * (lena - nwhole);
  const unsigned grid = static_cast<unsigned>(nitems < 65536 ? nitems : 65536);
  const size_t lds = split_rows_oz_lds_bytes(row_lds, lenb);
  if (row_lds)
    k_split_rows_oz<true><<<grid, kRowThreads, lds, st>>>(
        norb, compressed, npair, lena, nwhole, lenb, a, pta, tab, pdig, pea,
        pzs, fwa, sla, nslot, coeff, rowmax, slots, rowacc, done);
  else
    k_split_rows_oz<false><<<grid, kRowThreads, lds, st>>>(
        norb, compressed, npair, lena, nwhole, lenb, a, pta, tab, pdig, pea,
        pzs, fwa, sla, nslot, coeff, rowmax, slots, rowacc, done);
  FQEHIP_CHECK_LAUNCH("k_split_rows_oz");
  return FQEHIP_OK;
}

}  // namespace fqehip

// ---------------------------------------------------------------------------
// Beta-beta half of the Ozaki-digit split apply: rowacc[Ia][Ib'] =
// sum_Ib C[Ia][Ib] Hbb[Ib'][Ib] as digit products on the int8 matrix cores.
//
//   C: each row Ia scaled by 2^ec (|C 2^ec| < 2^46 from the row maximum,
//      shared by the real and imaginary parts), rounded to 48-bit integers;
//      digits 0..5 (bytes), 0..4 offset by -128 as in the alpha rows.
//      M index of the GEMM = 2 Ia + (0 real | 1 imaginary).
//   Hbb: each column Ib' scaled by 2^eh (|Hbb 2^eh| < 2^38), five balanced
//      digits, built once per operator (fqehip_beta_block_digits) together
//      with the column scales and the offset corrections.
// Products with i - t >= 0 (20 pairs), six int32 accumulators per output
// (exact for lenb < 26214: at most 5 * 128 * 128 * lenb < 2^31).
// ---------------------------------------------------------------------------
namespace fqehip {
namespace {

constexpr int kBgTM = 128;                  // tile rows (M)
constexpr int kBgTN = 64;                   // tile columns (N)
constexpr int kBgK = 64;                    // K step
constexpr int kBgXcds = 8;                  // XCDs of an MI355X
constexpr int kBgGroup = 8;                 // row tiles per group of the order

// Tile of workgroup blockIdx.x in an order that keeps each XCD on its own
// contiguous run of tiles, grouped kBgGroup row tiles at a time (workgroup
// L runs on XCD L % 8).  False for padding workgroups.
__device__ __forceinline__ bool oz_tile(int tm, int tn, int per_xcd, int* im,
                                        int* in) {
  const int x = static_cast<int>(blockIdx.x % kBgXcds);
  const int t = x * per_xcd + static_cast<int>(blockIdx.x / kBgXcds);
  if (t >= tm * tn) return false;
  const int g = t / (kBgGroup * tn);
  const int left = tm - g * kBgGroup;
  const int rows = left < kBgGroup ? left : kBgGroup;
  const int w = t - g * kBgGroup * tn;
  *im = g * kBgGroup + w % rows;
  *in = w / rows;
  return true;
}

// C digits: thread -> (row Ia, four consecutive columns); rows of ldk
// bytes per (digit, M row); cscale[Ia] = 2^-ec.
__global__ void __launch_bounds__(256)
k_c_digits(int64_t lena, int64_t lenb, int64_t ldk,
           const double* __restrict__ coeff, const double* __restrict__ rowmax,
           uint32_t* __restrict__ cdig, double* __restrict__ cscale,
           const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int64_t groups = ldk / 4;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= lena * groups) return;
  const int64_t ia = idx / groups;
  const int64_t g = idx - ia * groups;
  int ec = 0;
  const double mx = rowmax[ia];
  if (mx > 0.0) {
    int e;
    (void)frexp(mx, &e);
    ec = 46 - e;
  }
  if (g == 0) cscale[ia] = ldexp(1.0, -ec);
  const double sc = ldexp(1.0, ec);
  uint32_t lre[4], hre[4], lim[4], him[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t k = 4 * g + u;
    double xr = 0.0, xi = 0.0;
    if (k < lenb) {
      const double2 v = reinterpret_cast<const double2*>(coeff)[ia * lenb + k];
      xr = v.x;
      xi = v.y;
    }
    split_words(fma(xr, sc, kMagic), lre[u], hre[u]);
    split_words(fma(xi, sc, kMagic), lim[u], him[u]);
  }
  uint32_t ore[6], oim[6];
  digits_of_four(lre, hre, ore);
  digits_of_four(lim, him, oim);
  const int64_t rows = 2 * lena;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    cdig[((i * rows) + 2 * ia) * groups + g] = ore[i];
    cdig[((i * rows) + 2 * ia + 1) * groups + g] = oim[i];
  }
}

// Hbb column maxima: colmax[Ib'] = max_Ib |Hbb[Ib'][Ib]| (hcol[Ib][Ib']).
__global__ void __launch_bounds__(256)
k_hbb_colmax(int64_t lenb, const double* __restrict__ hcol,
             double* __restrict__ colmax) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (c >= lenb) return;
  double m = 0.0;
  for (int64_t k = 0; k < lenb; ++k) m = fmax(m, fabs(hcol[k * lenb + c]));
  colmax[c] = m;
}

// Hbb digits: thread -> column Ib'; five exact balanced digits of every K
// value's integer z (as for A' in the alpha rows), packed four K values per
// word; then the column's scale 2^-eh and offset correction
// 128 (1 + 2^8 + 2^16 + 2^24 + 2^32) sum_K z.
__global__ void __launch_bounds__(256)
k_hbb_digits(int64_t lenb, int64_t ldk, const double* __restrict__ hcol,
             const double* colmax, uint32_t* __restrict__ hdig,
             double* __restrict__ hscale, double* hcorr) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (c >= lenb) return;
  int eh = 0;
  if (colmax[c] > 0.0) {
    int e;
    (void)frexp(colmax[c], &e);
    eh = 38 - e;
  }
  const int64_t groups = ldk / 4;
  double zsum = 0.0;
  for (int64_t g = 0; g < groups; ++g) {
    uint32_t w[5] = {0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = 4 * g + u;
      int64_t z = k < lenb ? static_cast<int64_t>(rint(ldexp(hcol[k * lenb + c], eh)))
                           : 0;
      zsum += static_cast<double>(z);
#pragma unroll
      for (int t = 4; t >= 0; --t) {
        const int64_t d = t ? ((z + 128) & 255) - 128 : z;
        z = (z - d) >> 8;
        w[t] |= (static_cast<uint32_t>(d) & 0xffu) << (8 * u);
      }
    }
#pragma unroll
    for (int t = 0; t < 5; ++t) hdig[(t * lenb + c) * groups + g] = w[t];
  }
  const double corr = kOffsetUnit * zsum;
  hscale[c] = ldexp(1.0, -eh);
  hcorr[c] = corr;
}

// GEMM tile: 128 M rows (64 alpha strings, real and imaginary rows) x 64
// columns per workgroup of 8 waves (4 x 2, 32 x 32 per wave: 2 x 2 MFMA
// tiles of 16 x 16, six accumulators each, 96 registers), two waves per
// SIMD, K in steps of 64.  With 64 x 32 per wave (192 accumulators, one
// wave per SIMD) the LDS operand reads per MFMA were fewer, but every wait
// for a K step's loads and fragments was exposed: 1.59 ms instead of 1.35
// at norb 14.
// The operand tiles are double-buffered in LDS (2 x 68 KB, one workgroup
// per CU) and filled by direct buffer-to-LDS loads (buffer_load_dwordx4
// ... lds: lane L of an instruction writes bytes 16 L .. 16 L + 15 past the
// instruction's LDS base), so no staging registers: one instruction moves
// 16 rows x 64 bytes of one digit plane.  Rows are 64 bytes (16 banks),
// unpadded; the four 16-byte chunks of row r sit XOR-swizzled by (r / 2) % 4,
// so each 8 lanes of a fragment read (rows r..r+7 at one chunk: even and odd
// rows on the two bank halves, the four rows of each parity on four chunks)
// cover the 32 banks once (keyed by (r / 4) % 4, rows r and r + 2 shared
// banks: 1.5-2% slower).
// Step ks + 1's loads go out after the barrier of step ks and land in the
// other buffer while step ks's MFMAs run; one wait and one barrier per
// step.
constexpr int kBdBuf = (6 * kBgTM + 5 * kBgTN) * kBgK;   // bytes per buffer
constexpr int kBdNoRow = 0x7ff00000;        // offset of a row past the matrix

__device__ __forceinline__ int bd_off(int row, int chunk) {
  return row * kBgK + 16 * (chunk ^ ((row >> 1) & 3));
}

typedef __attribute__((address_space(3))) void lds_void_t;

__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_beta_gemm_oz(int64_t lena, int64_t lenb, int64_t ldk,
                  const uint8_t* __restrict__ cdig,
                  const double* __restrict__ cscale,
                  const uint8_t* __restrict__ hdig,
                  const double* __restrict__ hscale,
                  const double* __restrict__ hcorr, double* __restrict__ out,
                  const int32_t* __restrict__ done, int tm, int tn,
                  int per_xcd) {
  // two buffers as two arrays: the compiler can tell the loads into one
  // from the fragment reads of the other
  __shared__ __align__(16) uint8_t buf0[kBdBuf];
  __shared__ __align__(16) uint8_t buf1[kBdBuf];
  if (done && *done) return;
  int im, in;
  if (!oz_tile(tm, tn, per_xcd, &im, &in)) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int64_t rows = 2 * lena;
  const int64_t m0 = static_cast<int64_t>(im) * kBgTM;
  const int64_t n0 = static_cast<int64_t>(in) * kBgTN;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  const __amdgpu_buffer_rsrc_t rcd = buffer_of(cdig, 6 * rows * ldk);
  const __amdgpu_buffer_rsrc_t rhd = buffer_of(hdig, 5 * lenb * ldk);
  // lane L of a load instruction: tile row r0 + L/4, LDS chunk L%4, which
  // holds global chunk (L%4) ^ ((L/8)%4) (r0 is a multiple of 16)
  const int lrow = lane >> 2;
  const int lq = (lane & 3) ^ ((lane >> 3) & 3);
  // C: 6 planes x 8 instructions of 16 rows; H: 5 planes x 4; wave w takes
  // instructions w, w + 8, ... (six C, and three or two H: waves 0..3 take
  // H instruction 16 + w)  Byte offsets at k0 = 0, once: a row past
  // the matrix gets kBdNoRow, which stays past the end of the digit planes
  // for every k0 (their size is below kBdNoRow), so it reads zero.
  int coff[6], hoff[3];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int ins = wave + 8 * j;
    const int i = ins >> 3, r0 = (ins & 7) * 16;
    const int64_t gr = m0 + r0 + lrow;
    coff[j] = gr < rows ? static_cast<int>((i * rows + gr) * ldk + 16 * lq) : kBdNoRow;
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int ins = wave + 8 * j;
    const int i = ins >> 2, r0 = (ins & 3) * 16;
    const int64_t gn = n0 + r0 + lrow;
    hoff[j] = (ins < 20 && gn < lenb) ? static_cast<int>((i * lenb + gn) * ldk + 16 * lq) : kBdNoRow;
  }
  auto issue = [&](uint8_t* base, int k0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int ins = wave + 8 * j;
      const int i = ins >> 3, r0 = (ins & 7) * 16;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rcd, (lds_void_t*)(base + (i * kBgTM + r0) * kBgK), 16, coff[j] + k0,
          0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int ins = wave + 8 * j;
      if (ins < 20) {
        const int i = ins >> 2, r0 = (ins & 3) * 16;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rhd, (lds_void_t*)(base + 6 * kBgTM * kBgK + (i * kBgTN + r0) * kBgK),
            16, hoff[j] + k0, 0, 0, 0);
      }
    }
  };
  int4v acc[2][2][6];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int m = 0; m < 6; ++m) acc[a][b][m] = int4v{0, 0, 0, 0};
  // both column tiles' Hbb fragments held, one row tile's C fragments at a
  // time (64 operand registers)
  auto compute = [&](const uint8_t* base) {
    const uint8_t* sc = base;
    const uint8_t* sh = base + 6 * kBgTM * kBgK;
    const int q = lane >> 4;
    int4v b[2][5];
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int i = 0; i < 5; ++i)
        b[y][i] = *reinterpret_cast<const int4v*>(
            sh + i * kBgTN * kBgK + bd_off(wn + 16 * y + (lane & 15), q));
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      int4v a[6];
#pragma unroll
      for (int i = 0; i < 6; ++i)
        a[i] = *reinterpret_cast<const int4v*>(
            sc + i * kBgTM * kBgK + bd_off(wm + 16 * x + (lane & 15), q));
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
          for (int t = 0; t < 5; ++t)
            if (t <= i)
              acc[x][y][i - t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                  a[i], b[y][t], acc[x][y][i - t], 0, 0, 0);
    }
  };
  const int nk = static_cast<int>(ldk / kBgK);
  issue(buf0, 0);
  for (int ks = 0; ks < nk; ++ks) {
    // step ks from one buffer; step ks + 1 loads into the other meanwhile
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    uint8_t* cur = (ks & 1) ? buf1 : buf0;
    uint8_t* nxt = (ks & 1) ? buf0 : buf1;
    if (ks + 1 < nk) issue(nxt, (ks + 1) * kBgK);
    compute(cur);
  }
  // output item e of lane l: M row 4 (l/16) + e, column l%16; rows 2 Ia
  // (real) and 2 Ia + 1 (imaginary) are items e and e + 1
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int64_t col = n0 + wn + 16 * y + (lane & 15);
      if (col >= lenb) continue;
      const double hs = hscale[col];
      const double hc = hcorr[col];
#pragma unroll
      for (int e = 0; e < 4; e += 2) {
        const int64_t row = m0 + wm + 16 * x + 4 * (lane >> 4) + e;
        if (row >= rows) continue;
        const int64_t ia = row >> 1;
        double v[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          double t = static_cast<double>(acc[x][y][5][e + c]);
#pragma unroll
          for (int m = 4; m >= 0; --m)
            t = fma(t, 256.0, static_cast<double>(acc[x][y][m][e + c]));
          v[c] = fma(t, 4294967296.0, hc) * hs;
        }
        const double cs = cscale[ia];
        *reinterpret_cast<double2*>(out + (ia * lenb + col) * 2) =
            make_double2(v[0] * cs, v[1] * cs);
      }
    }
}

}  // namespace

int64_t beta_oz_ldk(int64_t lenb) { return (lenb + kBgK - 1) / kBgK * kBgK; }

// C digit planes of the apply: 6 x 2 lena x ldk bytes, then lena doubles.
int64_t beta_oz_workspace_bytes(int64_t lena, int64_t lenb) {
  return 6 * 2 * lena * beta_oz_ldk(lenb) + 8 * lena;
}

bool beta_oz_shape_ok(int64_t lena, int64_t lenb) {
  // byte offsets into the digit planes stay below the out-of-range row
  // offset of k_beta_gemm_oz
  const int64_t lim = kBdNoRow - beta_oz_ldk(lenb);
  return lenb < 26214 && 6 * 2 * lena * beta_oz_ldk(lenb) < lim &&
         5 * lenb * beta_oz_ldk(lenb) < lim;
}

int launch_beta_gemm_oz(int64_t lena, int64_t lenb, const void* hoz,
                        const double* coeff, const double* rowmax,
                        void* cwork, double* out, hipStream_t st,
                        const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  const int64_t ldk = beta_oz_ldk(lenb);
  uint8_t* cdig = static_cast<uint8_t*>(cwork);
  double* cscale = reinterpret_cast<double*>(cdig + 6 * 2 * lena * ldk);
  const uint8_t* hdig = static_cast<const uint8_t*>(hoz);
  const double* hscale = reinterpret_cast<const double*>(hdig + 5 * lenb * ldk);
  const double* hcorr = hscale + lenb;
  const int64_t nthr = lena * (ldk / 4);
  k_c_digits<<<blocks_for(nthr, 256), 256, 0, st>>>(
      lena, lenb, ldk, coeff, rowmax, reinterpret_cast<uint32_t*>(cdig), cscale,
      done);
  FQEHIP_CHECK_LAUNCH("k_c_digits");
  const int tm = static_cast<int>((2 * lena + kBgTM - 1) / kBgTM);
  const int tn = static_cast<int>((lenb + kBgTN - 1) / kBgTN);
  const int per_xcd = (tm * tn + kBgXcds - 1) / kBgXcds;
  k_beta_gemm_oz<<<static_cast<unsigned>(per_xcd * kBgXcds), 512, 0, st>>>(
      lena, lenb, ldk, cdig, cscale, hdig, hscale, hcorr, out, done, tm, tn,
      per_xcd);
  FQEHIP_CHECK_LAUNCH("k_beta_gemm_oz");
  return FQEHIP_OK;
}

int launch_row_absmax(int64_t lena, int64_t lenb, const double* coeff,
                      double* rowmax, hipStream_t st, const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  const unsigned grid = static_cast<unsigned>(lena < 65536 ? lena : 65536);
  k_row_absmax<<<grid, 256, 0, st>>>(lena, lenb, coeff, rowmax, done);
  FQEHIP_CHECK_LAUNCH("k_row_absmax");
  return FQEHIP_OK;
}

}  // namespace fqehip

using namespace fqehip;

extern "C" int64_t fqehip_beta_block_digits_bytes(int64_t lenb) {
  if (lenb < 0) return -1;
  return 5 * lenb * beta_oz_ldk(lenb) + 16 * lenb;
}

extern "C" int fqehip_beta_block_digits(int64_t lenb, const double* d_hcol,
                                        void* d_hoz, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(lenb >= 0, "lenb=%lld", static_cast<long long>(lenb));
  if (lenb == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_hcol && d_hoz, "NULL pointer");
  FQEHIP_CHECK_ARG(beta_oz_shape_ok(1, lenb), "lenb=%lld too large",
                   static_cast<long long>(lenb));
  hipStream_t st = as_stream(stream);
  const int64_t ldk = beta_oz_ldk(lenb);
  uint32_t* hdig = static_cast<uint32_t*>(d_hoz);
  double* hscale = reinterpret_cast<double*>(static_cast<uint8_t*>(d_hoz) + 5 * lenb * ldk);
  double* hcorr = hscale + lenb;
  // the correction slots hold the column maxima until the digits pass
  k_hbb_colmax<<<blocks_for(lenb, 256), 256, 0, st>>>(lenb, d_hcol, hcorr);
  FQEHIP_CHECK_LAUNCH("k_hbb_colmax");
  k_hbb_digits<<<blocks_for(lenb, 256), 256, 0, st>>>(lenb, ldk, d_hcol, hcorr,
                                                      hdig, hscale, hcorr);
  FQEHIP_CHECK_LAUNCH("k_hbb_digits");
  return FQEHIP_OK;
}

extern "C" int fqehip_beta_gemm_digits(int64_t lena, int64_t lenb,
                                       const void* d_hoz, const void* d_coeff,
                                       void* d_work, void* d_out,
                                       const int32_t* d_done, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(lena >= 0 && lenb >= 0, "negative size");
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  FQEHIP_CHECK_ARG(d_hoz && d_coeff && d_work && d_out, "NULL pointer");
  FQEHIP_CHECK_ARG(beta_oz_shape_ok(lena, lenb), "sector too large");
  hipStream_t st = as_stream(stream);
  double* rowmax = static_cast<double*>(d_work);
  int rc = launch_row_absmax(lena, lenb, static_cast<const double*>(d_coeff),
                             rowmax, st, d_done);
  if (rc) return rc;
  return launch_beta_gemm_oz(lena, lenb, d_hoz,
                             static_cast<const double*>(d_coeff), rowmax,
                             rowmax + lena, static_cast<double*>(d_out), st,
                             d_done);
}
