// Fused apply for real operators: one workgroup per alpha string (row Ia
// of the CI matrix).
//
//   k_fused_*:  for each chunk of beta columns of row Ia
//       1. gather the dvec tile D[q][Ia, chunk] into LDS from C
//          (alpha part: rows C[src_a(q, Ia), chunk], coalesced; beta part:
//          C[Ia, src_b(q, Ib)] within row Ia);
//       2. Y tile = A D tile on the matrix cores (A = pair matrix + h1' row,
//          held in registers for the whole workgroup);
//       3. scatter the Y tile without atomics to global memory:
//          - one-body row, alpha-diagonal pairs and every beta excitation
//            land in row Ia itself: LDS accumulator row (LDS atomics);
//          - an off-diagonal alpha excitation E_p Ia = tgt lands in a
//            per-target "slot" row slots[tgt][k] (k = index of (Ia, p)
//            among tgt's off-diagonal de-excitations): a plain store.
//       The LDS row is written out as rowacc[Ia].
//   k_slot_reduce: out[I] = alpha (rowacc[I] + sum_k slots[Ia][k][Ib])
//       + beta cin[I], with the Taylor epilogue (FP64 sum, FP64 norm
//       partials).  A streaming pass.
//
// The dvec never reaches HBM: per apply the traffic is C (cached), the
// slot rows (n_alpha (norb - n_alpha) per determinant), rowacc and out.
#include "fqehip_common.h"
#include "rows_common.h"


namespace fqehip {
namespace {

using namespace rows;

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMPad = 128;  // output rows of the Y tile (npair + 1 <= 128)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float x) {
  return static_cast<__bf16>(x);  // round to nearest even (v_cvt_pk_bf16_f32)
}

__device__ __forceinline__ float from_bf16(__bf16 h) {
  return __uint_as_float(static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h))
                         << 16);
}


__device__ __forceinline__ float sgnf(int32_t v) {
  return v > 0 ? 1.0f : (v < 0 ? -1.0f : 0.0f);
}

// ---------------------------------------------------------------------------
// tables of the fused path
// ---------------------------------------------------------------------------

// fwd[p][src] = sign * (tgt + 1) where <tgt|E_p|src> = sign, from the gather
// table; fwd is zero-filled first.
__global__ void __launch_bounds__(kThreads)
k_forward_table(int64_t total, int64_t len, const int32_t* __restrict__ pt,
                int32_t* __restrict__ fwd) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= total) return;
  const int32_t e = pt[idx];
  if (!e) return;
  const int64_t p = idx / len;
  const int64_t tgt = idx - p * len;
  const int64_t src = iabs(e) - 1;
  fwd[p * len + src] = (e > 0 ? 1 : -1) * static_cast<int32_t>(tgt + 1);
}

// slot[p][src] = k for the off-diagonal de-excitation rows of every target,
// numbered in dexc order; -1 elsewhere (pre-filled).
__global__ void __launch_bounds__(kThreads)
k_slot_table(int norb, int nexc, int64_t len, const int32_t* __restrict__ dexc,
             int compressed, int32_t* __restrict__ slot) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t >= len) return;
  int k = 0;
  for (int e = 0; e < nexc; ++e) {
    const int32_t* row = dexc + (t * nexc + e) * 3;
    const int32_t src = row[0];
    if (src == t) continue;
    const int i = row[1] / norb;
    const int j = row[1] % norb;
    const int p = compressed ? comp_pair(i, j) : row[1];
    slot[static_cast<int64_t>(p) * len + src] = k++;
  }
}

__global__ void __launch_bounds__(kThreads)
k_fill(int64_t n, int32_t v, int32_t* out) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
       k < n; k += static_cast<int64_t>(gridDim.x) * kThreads)
    out[k] = v;
}

// ---------------------------------------------------------------------------
// Fused kernels.  The Y tile has 128 (padded) rows: one 16-row tile per
// wave (8 waves).  Columns of the D / Y tiles are split into a real block
// and an imaginary block, so the lane that holds Re Y[p][Ib] also holds
// Im Y[p][Ib] and the scatter works on complex values.  Per alpha row the
// workgroup caches in LDS: the alpha source offset and sign of every pair
// (gather), and the slot-row base and sign of every pair's alpha target
// (scatter); beta entries come from the pair / forward tables in L2.
// ---------------------------------------------------------------------------
constexpr int kFWaves = 8;
constexpr int kFThreads = kFWaves * 64;

// Scatter of one complex Y element (p, jb) of row ia (see file comment).
template <typename T, typename A>
__device__ __forceinline__ void scatter_complex(
    A yr, A yi, int p, int npair, int64_t jb, int32_t fbv, RowTables rt,
    A* accrow, T* __restrict__ slots) {
  if (p == npair) {
    atomicAdd(&accrow[2 * jb], yr);
    atomicAdd(&accrow[2 * jb + 1], yi);
    return;
  }
  if (fbv) {
    const int64_t tb = 2 * (static_cast<int64_t>(iabs(fbv)) - 1);
    atomicAdd(&accrow[tb], fbv > 0 ? yr : -yr);
    atomicAdd(&accrow[tb + 1], fbv > 0 ? yi : -yi);
  }
  const int64_t sb = rt.sbase[p];
  if (sb == -1) return;
  const A s = static_cast<A>(rt.ssgn[p]);
  if (sb == -2) {
    atomicAdd(&accrow[2 * jb], s * yr);
    atomicAdd(&accrow[2 * jb + 1], s * yi);
  } else {
    T* dst = slots + sb + 2 * jb;
    dst[0] = static_cast<T>(s * yr);
    dst[1] = static_cast<T>(s * yi);
  }
}

// ---------------- mixed precision: bf16 hi/lo x3, FP32 accumulate ---------
// NW waves, two or more workgroups (alpha rows) per CU, so one row's gather
// overlaps the other's MFMAs and scatter.
//
// kSplit (pair-symmetric operator, beta-beta block done by
// k_beta_gemm_bf16x3 into rowacc beforehand): the gathered operand is
// D'[q] = D^a[q] + 2 D^b[q] (q < npair) and D'[npair] = C[ia]; A' has the
// extra column A'[p][npair] = h1'[p], and its row npair is rebuilt per
// alpha row as the sum of the rows of the occupied diagonal pairs (ii).
// The scatter then stores off-diagonal alpha excitations into slot rows
// and adds row npair to rowacc[ia]: no atomics, no LDS accumulator row.
constexpr int kBfTbC = 32;                 // complex columns per chunk
constexpr int kBfKPad = 128;
constexpr int kBfKSteps = kBfKPad / 32;
constexpr int kBfStride = kBfKPad + 8;     // bf16 per LDS column (padding)
constexpr int kBfCols = 2 * kBfTbC;        // re block then im block

// Per-row body and A' fragment setup of k_fused_bf16x3, shared with the
// grouped launch over several sectors (k_fused_group).
template <int NW, bool kSplit>
__device__ __forceinline__ void fused_a_frags(
    int npair, const double* __restrict__ amat,
    bf16x8 (&ahi)[8 / NW][kBfKSteps], bf16x8 (&alo)[8 / NW][kBfKSteps]) {
  constexpr int kRT = 8 / NW;              // row tiles per wave
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int K = npair;
  const int M = npair + 1;
#pragma unroll
  for (int r = 0; r < kRT; ++r) {
    const int row = (wave * kRT + r) * 16 + (lane & 15);
#pragma unroll
    for (int kt = 0; kt < kBfKSteps; ++kt) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int kk = kt * 32 + 8 * (lane >> 4) + j;
        double v = 0.0;
        if (kSplit) {
          if (row < npair && kk < K) v = amat[static_cast<int64_t>(row) * K + kk];
          else if (row < npair && kk == K) v = amat[static_cast<int64_t>(npair) * K + row];
        } else if (row < M && kk < K) {
          v = amat[static_cast<int64_t>(row) * K + kk];
        }
        const __bf16 h = to_bf16(static_cast<float>(v));
        ahi[r][kt][j] = h;
        alo[r][kt][j] = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
      }
    }
  }
}

template <int NW, bool kSplit>
__device__ __forceinline__ void fused_row(
    int64_t ia, int norb, int compressed, int npair, int64_t lena, int64_t lenb,
    const double* __restrict__ amat, const int32_t* __restrict__ pta,
    const int32_t* __restrict__ ptb, const int32_t* __restrict__ fwa,
    const int32_t* __restrict__ fwb, const int32_t* __restrict__ sla, int nslot,
    const float* __restrict__ coeff, float* __restrict__ slots,
    float* __restrict__ rowacc, char* smem,
    bf16x8 (&ahi)[8 / NW][kBfKSteps], bf16x8 (&alo)[8 / NW][kBfKSteps]) {
  constexpr int kT = NW * 64;              // threads
  constexpr int kRT = 8 / NW;              // row tiles per wave
  constexpr int kGroups = kT / kBfTbC;     // gather pair groups of 8
  constexpr int kGIter = kBfKPad / (8 * kGroups);
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  __bf16* dhi = reinterpret_cast<__bf16*>(smem);
  __bf16* dlo = dhi + kBfCols * kBfStride;
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(dlo + kBfCols * kBfStride);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  float* accrow = reinterpret_cast<float*>(rt.ssgn + kRowTabs);
  double* drow = reinterpret_cast<double*>(rt.ssgn + kRowTabs);  // kSplit
  const int Kg = kSplit ? K + 1 : K;       // gathered rows of the D tile
  // k rows the MFMAs read (32 per step): small sectors skip the padding
  const int ksteps = (Kg + 31) / 32;
  const int kpad = 32 * ksteps;

  const bool wave_live = wave * kRT * 16 <= npair;
  const int t = tid % kBfTbC;            // gather column
  const int g = tid / kBfTbC;            // gather pair group
  const int nchunks = static_cast<int>((lenb + kBfTbC - 1) / kBfTbC);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(float));
  const __amdgpu_buffer_rsrc_t rpb = buffer_of(ptb, static_cast<int64_t>(K) * lenb * 4);
  const __amdgpu_buffer_rsrc_t rfb = buffer_of(fwb, static_cast<int64_t>(K) * lenb * 4);


  load_row_tables<float>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                         kT, kSplit ? K : -1);
  if (kSplit) {
    __syncthreads();
    // row npair of A' for this alpha string: sum over occupied orbitals i
    // (E^a_ii ia = ia, marked -2 in the row tables) of A'[ii][:]
    for (int k = tid; k < Kg; k += kT) {
      double v = 0.0;
      for (int i = 0; i < norb; ++i) {
        const int dp = compressed ? i + i * (i + 1) / 2 : i * (norb + 1);
        if (rt.sbase[dp] == -2)
          v += k < K ? amat[static_cast<int64_t>(dp) * K + k]
                     : amat[static_cast<int64_t>(npair) * K + dp];
      }
      drow[k] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRT; ++r) {
      if ((wave * kRT + r) * 16 + (lane & 15) == npair) {
#pragma unroll
        for (int kt = 0; kt < kBfKSteps; ++kt)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int kk = kt * 32 + 8 * (lane >> 4) + j;
            const double v = kk < Kg ? drow[kk] : 0.0;
            const __bf16 h = to_bf16(static_cast<float>(v));
            ahi[r][kt][j] = h;
            alo[r][kt][j] = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
          }
      }
    }
  } else {
    for (int64_t x = tid; x < 2 * lenb; x += kT) accrow[x] = 0.0f;
  }
  const int crow_b = static_cast<int>(ia * lenb * 2 * sizeof(float));
  int32_t bt[kGIter][8];                 // gather-table entries, this chunk
  auto load_bt = [&](int64_t ib0) {
    const int ib = static_cast<int>(ib0) + t;
#pragma unroll
    for (int it = 0; it < kGIter; ++it)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int q = 8 * (g + it * kGroups) + j;
        bt[it][j] = q < kpad ? bload_i32(rpb, (ib < lenb && q < K)
                                                  ? (q * static_cast<int>(lenb) + ib) * 4 : -1)
                             : 0;
      }
  };
  load_bt(0);
  __syncthreads();
  for (int chunk = 0; chunk < nchunks; ++chunk) {
    const int64_t ib0 = static_cast<int64_t>(chunk) * kBfTbC;
    {
      const int ib = static_cast<int>(ib0) + t;
      const bool colok = ib < lenb;
#pragma unroll
      for (int it = 0; it < kGIter; ++it) {
        const int q0 = 8 * (g + it * kGroups);
        if (q0 >= kpad) continue;        // rows no MFMA reads
        bf16x8 hr, lr, hi, li;
        float2 ca[8], cb[8];
        float sa[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int q = q0 + j;
          const int32_t off = (colok && q < Kg) ? rt.aoff[q] : -1;
          sa[j] = off >= 0 ? static_cast<float>(rt.asgn[q]) : 0.0f;
          ca[j] = bload_f2(rc, off >= 0 ? (off + 2 * ib) * 4 : -1);
          const int32_t b = bt[it][j];
          cb[j] = bload_f2(rc, b ? crow_b + 8 * (iabs(b) - 1) : -1);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float sb = (kSplit ? 2.0f : 1.0f) * sgnf(bt[it][j]);
          const float re = sa[j] * ca[j].x + sb * cb[j].x;
          const float im = sa[j] * ca[j].y + sb * cb[j].y;
          const __bf16 h0 = to_bf16(re), h1 = to_bf16(im);
          hr[j] = h0;
          lr[j] = to_bf16(re - from_bf16(h0));
          hi[j] = h1;
          li[j] = to_bf16(im - from_bf16(h1));
        }
        *reinterpret_cast<bf16x8*>(dhi + t * kBfStride + q0) = hr;
        *reinterpret_cast<bf16x8*>(dlo + t * kBfStride + q0) = lr;
        *reinterpret_cast<bf16x8*>(dhi + (kBfTbC + t) * kBfStride + q0) = hi;
        *reinterpret_cast<bf16x8*>(dlo + (kBfTbC + t) * kBfStride + q0) = li;
      }
    }
    __syncthreads();
    // the next chunk's gather table is in flight during the MFMAs
    if (chunk + 1 < nchunks) load_bt(ib0 + kBfTbC);
    if (wave_live) {
      // two passes over 16-column halves h: tile h holds Re, tile h+2 Im
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int32_t fbv[kRT][4];
#pragma unroll
        for (int r = 0; r < kRT; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int p = (wave * kRT + r) * 16 + 4 * (lane >> 4) + i;
            const int jb = static_cast<int>(ib0) + h * 16 + (lane & 15);
            fbv[r][i] = kSplit ? 0 : bload_i32(rfb, (p < npair && jb < lenb)
                                           ? (p * static_cast<int>(lenb) + jb) * 4 : -1);
          }
        floatx4 are[kRT], aim[kRT];
#pragma unroll
        for (int r = 0; r < kRT; ++r) {
          are[r] = floatx4{0, 0, 0, 0};
          aim[r] = floatx4{0, 0, 0, 0};
        }
#pragma unroll
        for (int kt = 0; kt < kBfKSteps; ++kt) {
          if (kt >= ksteps) break;
          const int kq = kt * 32 + 8 * (lane >> 4);
          const int cre = h * 16 + (lane & 15);
          const int cim = kBfTbC + cre;
          const bf16x8 bhr = *reinterpret_cast<const bf16x8*>(dhi + cre * kBfStride + kq);
          const bf16x8 blr = *reinterpret_cast<const bf16x8*>(dlo + cre * kBfStride + kq);
          const bf16x8 bhi = *reinterpret_cast<const bf16x8*>(dhi + cim * kBfStride + kq);
          const bf16x8 bli = *reinterpret_cast<const bf16x8*>(dlo + cim * kBfStride + kq);
#pragma unroll
          for (int r = 0; r < kRT; ++r) {
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bhr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bhi, aim[r], 0, 0, 0);
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], blr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bli, aim[r], 0, 0, 0);
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r][kt], bhr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r][kt], bhi, aim[r], 0, 0, 0);
          }
        }
        // f32 MFMA layout: item i of lane l is row 4*(l/16)+i, col l%16
#pragma unroll
        for (int r = 0; r < kRT; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int p = (wave * kRT + r) * 16 + 4 * (lane >> 4) + i;
            const int64_t jb = ib0 + h * 16 + (lane & 15);
            if (kSplit) {
              if (jb >= lenb || p > npair) continue;
              if (p == npair) {
                float2* d = reinterpret_cast<float2*>(rowacc + (ia * lenb + jb) * 2);
                float2 o = *d;
                o.x += are[r][i];
                o.y += aim[r][i];
                *d = o;
              } else {
                const int64_t sb = rt.sbase[p];
                if (sb >= 0) {
                  const float sg = static_cast<float>(rt.ssgn[p]);
                  stream_store2(slots + sb + 2 * jb, sg * are[r][i], sg * aim[r][i]);
                }
              }
            } else if (p <= npair && jb < lenb) {
              scatter_complex<float, float>(are[r][i], aim[r][i], p, npair,
                                            jb, fbv[r][i], rt, accrow, slots);
            }
          }
      }
    }
    __syncthreads();
  }
  if (!kSplit) {
    float* dst = rowacc + ia * 2 * lenb;
    for (int64_t x = tid; x < 2 * lenb; x += kT) dst[x] = accrow[x];
  }
  __syncthreads();
}

template <int NW, int WPE, bool kSplit>
__global__ void __launch_bounds__(NW * 64, WPE)
k_fused_bf16x3(int norb, int compressed, int npair, int64_t lena, int64_t lenb,
               const double* __restrict__ amat,
               const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
               const int32_t* __restrict__ fwa, const int32_t* __restrict__ fwb,
               const int32_t* __restrict__ sla, int nslot,
               const float* __restrict__ coeff, float* __restrict__ slots,
               float* __restrict__ rowacc, const int32_t* __restrict__ done) {
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  bf16x8 ahi[8 / NW][kBfKSteps], alo[8 / NW][kBfKSteps];
  fused_a_frags<NW, kSplit>(npair, amat, ahi, alo);
  for (int64_t ia = blockIdx.x; ia < lena; ia += gridDim.x)
    fused_row<NW, kSplit>(ia, norb, compressed, npair, lena, lenb, amat, pta,
                          ptb, fwa, fwb, sla, nslot, coeff, slots, rowacc, smem,
                          ahi, alo);
}

// ---------------- split kernel over the live pair rows ------------------------
// Of the npair + 1 rows of Y = A' D only those whose alpha excitation of
// the row's string ia is a different string (sbase >= 0, nalpha * (norb -
// nalpha) of them) and the folded row npair are scattered.  This kernel
// compacts them per alpha row into at most kLiveRows rows: the A'
// fragments of the live rows are rebuilt in registers per alpha row, so
// the chunk loop runs half the MFMAs of the all-rows kernel, and each
// lane's scatter targets (slot base and sign of its 4 output rows) are
// computed once per alpha row and kept in registers.  TBC beta columns per
// chunk; ROWLDS stages row ia of C in LDS for the beta half of the gather
// (random 8-byte reads within the row: from LDS instead of one cache line
// per lane through L1/L2).

template <int NW, int WPE, int TBC, bool ROWLDS>
__global__ void __launch_bounds__(NW * 64, WPE)
k_split_live(int norb, int compressed, int npair, int64_t lena, int64_t lenb,
             const double* __restrict__ amat,
             const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
             const int32_t* __restrict__ fwa, const int32_t* __restrict__ sla,
             int nslot, const float* __restrict__ coeff,
             float* __restrict__ slots, float* __restrict__ rowacc,
             const int32_t* __restrict__ done) {
  constexpr int kT = NW * 64;
  constexpr int kRT = kLiveRows / 16 / NW;     // live row tiles per wave
  constexpr int kGroups = kT / TBC;            // gather pair groups of 8
  constexpr int kCols = 2 * TBC;
  constexpr int kHalves = TBC / 16;
  constexpr int kGIter = kBfKPad / (8 * kGroups);
  static_assert(kRT >= 1 && kRT * NW * 16 == kLiveRows, "row tiles");
  static_assert(kGIter >= 1 && kGIter * 8 * kGroups == kBfKPad, "gather");
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  const int Kg = K + 1;                        // gathered rows: pairs + C
  __bf16* dhi = reinterpret_cast<__bf16*>(smem);
  __bf16* dlo = dhi + kCols * kBfStride;
  double* drow = reinterpret_cast<double*>(dlo + kCols * kBfStride);
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(drow + kBfKPad);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  int32_t* lpair = rt.ssgn + kRowTabs;         // [kLiveRows] pair of live row
  float2* crow = reinterpret_cast<float2*>(lpair + kLiveRows);  // ROWLDS

  const int t = tid % TBC;                     // gather column
  const int g = tid / TBC;                     // gather pair group
  const int nchunks = static_cast<int>((lenb + TBC - 1) / TBC);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(float));
  const __amdgpu_buffer_rsrc_t rpb = buffer_of(ptb, static_cast<int64_t>(K) * lenb * 4);

  for (int64_t ia = blockIdx.x; ia < lena; ia += gridDim.x) {
    load_row_tables<float>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                           kT, K);
    __syncthreads();
    // live rows: the pairs with sbase >= 0 in increasing order (wave 0,
    // ballot compaction), then row npair, then padding (-1)
    if (wave == 0) {
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
    }
    // row npair of A' for this alpha string: sum over occupied orbitals i
    // (E^a_ii ia = ia, marked -2 in the row tables) of A'[ii][:]
    for (int k = tid; k < Kg; k += kT) {
      double v = 0.0;
      for (int i = 0; i < norb; ++i) {
        const int dp = compressed ? i + i * (i + 1) / 2 : i * (norb + 1);
        if (rt.sbase[dp] == -2)
          v += k < K ? amat[static_cast<int64_t>(dp) * K + k]
                     : amat[static_cast<int64_t>(npair) * K + dp];
      }
      drow[k] = v;
    }
    if (ROWLDS) {
      const float2* src = reinterpret_cast<const float2*>(coeff) + ia * lenb;
      for (int64_t x = tid; x < lenb; x += kT) crow[x] = src[x];
    }
    __syncthreads();
    // A' fragments of the live rows (bf16 hi / lo split), the scatter
    // targets of this lane's output rows and the alpha gather entries of
    // this thread's pair rows
    bf16x8 ahi[kRT][kBfKSteps], alo[kRT][kBfKSteps];
    int64_t obase[kRT][4];                     // -1 none, -3 rowacc
    float osgn[kRT][4];
#pragma unroll
    for (int r = 0; r < kRT; ++r) {
      const int p = lpair[(wave * kRT + r) * 16 + (lane & 15)];
#pragma unroll
      for (int kt = 0; kt < kBfKSteps; ++kt) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int kk = kt * 32 + 8 * (lane >> 4) + j;
          double v = 0.0;
          if (p == npair) {
            if (kk < Kg) v = drow[kk];
          } else if (p >= 0) {
            if (kk < K) v = amat[static_cast<int64_t>(p) * K + kk];
            else if (kk == K) v = amat[static_cast<int64_t>(npair) * K + p];
          }
          const __bf16 h = to_bf16(static_cast<float>(v));
          ahi[r][kt][j] = h;
          alo[r][kt][j] = to_bf16(static_cast<float>(v - static_cast<double>(from_bf16(h))));
        }
      }
      // f32 MFMA layout: item i of lane l is row 4*(l/16)+i, col l%16
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int po = lpair[(wave * kRT + r) * 16 + 4 * (lane >> 4) + i];
        obase[r][i] = po == npair ? -3 : (po < 0 ? -1 : rt.sbase[po]);
        osgn[r][i] = po >= 0 && po < npair ? static_cast<float>(rt.ssgn[po]) : 1.0f;
      }
    }
    int32_t aof[kGIter][8];
    float asg[kGIter][8];
#pragma unroll
    for (int it = 0; it < kGIter; ++it)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int q = 8 * (g + it * kGroups) + j;
        const int32_t off = q < Kg ? rt.aoff[q] : -1;
        aof[it][j] = off;
        asg[it][j] = off >= 0 ? static_cast<float>(rt.asgn[q]) : 0.0f;
      }
    const int crow_b = static_cast<int>(ia * lenb * 2 * sizeof(float));
    float* racc = rowacc + ia * lenb * 2;
    for (int chunk = 0; chunk < nchunks; ++chunk) {
      const int64_t ib0 = static_cast<int64_t>(chunk) * TBC;
      {
        const int ib = static_cast<int>(ib0) + t;
        const bool colok = ib < lenb;
#pragma unroll
        for (int it = 0; it < kGIter; ++it) {
          const int q0 = 8 * (g + it * kGroups);
          bf16x8 hr, lr, hi, li;
          float2 ca[8], cb[8];
          int32_t bt[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int q = q0 + j;
            bt[j] = bload_i32(rpb, (colok && q < K)
                                       ? (q * static_cast<int>(lenb) + ib) * 4 : -1);
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int32_t off = aof[it][j];
            ca[j] = bload_f2(rc, (colok && off >= 0) ? (off + 2 * ib) * 4 : -1);
            if (ROWLDS) cb[j] = crow[bt[j] ? iabs(bt[j]) - 1 : 0];
            else cb[j] = bload_f2(rc, bt[j] ? crow_b + 8 * (iabs(bt[j]) - 1) : -1);
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const float sb = 2.0f * sgnf(bt[j]);
            const float re = asg[it][j] * ca[j].x + sb * cb[j].x;
            const float im = asg[it][j] * ca[j].y + sb * cb[j].y;
            const __bf16 h0 = to_bf16(re), h1 = to_bf16(im);
            hr[j] = h0;
            lr[j] = to_bf16(re - from_bf16(h0));
            hi[j] = h1;
            li[j] = to_bf16(im - from_bf16(h1));
          }
          *reinterpret_cast<bf16x8*>(dhi + t * kBfStride + q0) = hr;
          *reinterpret_cast<bf16x8*>(dlo + t * kBfStride + q0) = lr;
          *reinterpret_cast<bf16x8*>(dhi + (TBC + t) * kBfStride + q0) = hi;
          *reinterpret_cast<bf16x8*>(dlo + (TBC + t) * kBfStride + q0) = li;
        }
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < kHalves; ++h) {
        floatx4 are[kRT], aim[kRT];
#pragma unroll
        for (int r = 0; r < kRT; ++r) {
          are[r] = floatx4{0, 0, 0, 0};
          aim[r] = floatx4{0, 0, 0, 0};
        }
#pragma unroll
        for (int kt = 0; kt < kBfKSteps; ++kt) {
          const int kq = kt * 32 + 8 * (lane >> 4);
          const int cre = h * 16 + (lane & 15);
          const int cim = TBC + cre;
          const bf16x8 bhr = *reinterpret_cast<const bf16x8*>(dhi + cre * kBfStride + kq);
          const bf16x8 blr = *reinterpret_cast<const bf16x8*>(dlo + cre * kBfStride + kq);
          const bf16x8 bhi = *reinterpret_cast<const bf16x8*>(dhi + cim * kBfStride + kq);
          const bf16x8 bli = *reinterpret_cast<const bf16x8*>(dlo + cim * kBfStride + kq);
#pragma unroll
          for (int r = 0; r < kRT; ++r) {
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bhr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bhi, aim[r], 0, 0, 0);
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], blr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r][kt], bli, aim[r], 0, 0, 0);
            are[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r][kt], bhr, are[r], 0, 0, 0);
            aim[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r][kt], bhi, aim[r], 0, 0, 0);
          }
        }
        const int64_t jb = ib0 + h * 16 + (lane & 15);
        if (jb < lenb) {
#pragma unroll
          for (int r = 0; r < kRT; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int64_t sb = obase[r][i];
              if (sb >= 0) {
                stream_store2(slots + sb + 2 * jb, osgn[r][i] * are[r][i],
                              osgn[r][i] * aim[r][i]);
              } else if (sb == -3) {
                float2* d = reinterpret_cast<float2*>(racc + 2 * jb);
                float2 o = *d;
                o.x += are[r][i];
                o.y += aim[r][i];
                *d = o;
              }
            }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------- FP64: v_mfma_f64_16x16x4f64 -----------------------------
// One workgroup per CU (the FP64 row accumulator fills the LDS); the D tile
// is double-buffered so the gather of chunk c+1 (coefficients loaded one
// chunk ahead, table entries two chunks ahead) overlaps the FP64 MFMAs of
// chunk c, with one barrier per chunk.
constexpr int kF64TbC = 16;                // complex columns per chunk
constexpr int kF64Cols = 2 * kF64TbC;      // re block then im block
constexpr int kF64KPad = 128;              // gather rows (MFMA uses ceil(K/4))
constexpr int kF64KMax = 112;              // npair limit of the FP64 kernel
constexpr int kF64KSteps = kF64KMax / 4;
constexpr int kF64Stride = kF64Cols + 1;   // doubles per LDS row (padding)
constexpr int kF64Groups = kFThreads / kF64TbC;
constexpr int kF64GIter = kF64KPad / kF64Groups;
constexpr int kF64Tile = kF64KPad * kF64Stride;

// A' fragments of the wave's 16 rows in MFMA A-operand registers (lane l:
// row 16 wave + l%16, k = 4 kt + l/16), kept for every alpha row.
__device__ __forceinline__ void f64_a_regs(int npair,
                                           const double* __restrict__ amat,
                                           double (&areg)[kF64KSteps]) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int K = npair;
  const int M = npair + 1;
  const int row = wave * 16 + (lane & 15);
#pragma unroll
  for (int kt = 0; kt < kF64KSteps; ++kt) {
    const int kk = kt * 4 + (lane >> 4);
    areg[kt] = (row < M && kk < K) ? amat[static_cast<int64_t>(row) * K + kk] : 0.0;
  }
}

// Alpha row ia of k_fused_f64 (shared with the grouped launch over the
// sectors of a batch, k_fused_group_f64).
__device__ __forceinline__ void f64_row(
    int64_t ia, int npair, int64_t lena, int64_t lenb,
    const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
    const int32_t* __restrict__ fwa, const int32_t* __restrict__ fwb,
    const int32_t* __restrict__ sla, int nslot,
    const double* __restrict__ coeff, double* __restrict__ slots,
    double* __restrict__ rowacc, char* smem,
    const double (&areg)[kF64KSteps]) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  const int ksteps = (K + 3) / 4;
  double* dt = reinterpret_cast<double*>(smem);     // 2 x [kF64KPad][kF64Stride]
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(dt + 2 * kF64Tile);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  double* accrow = reinterpret_cast<double*>(rt.ssgn + kRowTabs);

  const bool wave_live = wave * 16 <= npair;
  const int t = tid % kF64TbC;
  const int g = tid / kF64TbC;
  const int nchunks = static_cast<int>((lenb + kF64TbC - 1) / kF64TbC);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(double));
  const __amdgpu_buffer_rsrc_t rpb = buffer_of(ptb, static_cast<int64_t>(K) * lenb * 4);
  const __amdgpu_buffer_rsrc_t rfb = buffer_of(fwb, static_cast<int64_t>(K) * lenb * 4);

  {
    load_row_tables<double>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                            kFThreads);
    for (int64_t x = tid; x < 2 * lenb; x += kFThreads) accrow[x] = 0.0;
    const int crow_b = static_cast<int>(ia * lenb * 2 * sizeof(double));
    __syncthreads();

    int32_t bt[kF64GIter];                 // gather-table entries (ahead)
    double2 ca[kF64GIter], cb[kF64GIter];  // coefficients (one chunk ahead)
    int32_t as[kF64GIter], bs[kF64GIter];
    auto load_bt = [&](int64_t ib0) {
      const int ib = static_cast<int>(ib0) + t;
#pragma unroll
      for (int it = 0; it < kF64GIter; ++it) {
        const int q = g + it * kF64Groups;
        bt[it] = bload_i32(rpb, (ib < lenb && q < K)
                                    ? (q * static_cast<int>(lenb) + ib) * 4 : -1);
      }
    };
    auto load_coeff = [&](int64_t ib0) {
      const int ib = static_cast<int>(ib0) + t;
      const bool colok = ib < lenb;
#pragma unroll
      for (int it = 0; it < kF64GIter; ++it) {
        const int q = g + it * kF64Groups;
        const int32_t off = (colok && q < K) ? rt.aoff[q] : -1;
        as[it] = off >= 0 ? rt.asgn[q] : 0;
        ca[it] = bload_d2(rc, off >= 0 ? (off + 2 * ib) * 8 : -1);
        bs[it] = bt[it];
        cb[it] = bload_d2(rc, bt[it] ? crow_b + 16 * (iabs(bt[it]) - 1) : -1);
      }
    };
    auto store_tile = [&](double* tile) {
#pragma unroll
      for (int it = 0; it < kF64GIter; ++it) {
        const int q = g + it * kF64Groups;
        const double sa = static_cast<double>(as[it]);
        const double sb = bs[it] > 0 ? 1.0 : (bs[it] < 0 ? -1.0 : 0.0);
        tile[q * kF64Stride + t] = sa * ca[it].x + sb * cb[it].x;
        tile[q * kF64Stride + kF64TbC + t] = sa * ca[it].y + sb * cb[it].y;
      }
    };
    // prologue: tile 0 in buffer 0, tables of chunk 1 in registers
    load_bt(0);
    load_coeff(0);
    if (nchunks > 1) load_bt(kF64TbC);
    store_tile(dt);
    __syncthreads();

    for (int chunk = 0; chunk < nchunks; ++chunk) {
      const int64_t ib0 = static_cast<int64_t>(chunk) * kF64TbC;
      const double* cur = dt + (chunk & 1) * kF64Tile;
      double* nxt = dt + ((chunk + 1) & 1) * kF64Tile;
      const bool more = chunk + 1 < nchunks;
      if (more) load_coeff(ib0 + kF64TbC);
      if (chunk + 2 < nchunks) load_bt(ib0 + 2 * kF64TbC);
      int32_t fbv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = wave * 16 + (lane >> 4) + 4 * i;
        const int jb = static_cast<int>(ib0) + (lane & 15);
        fbv[i] = bload_i32(rfb, (wave_live && p < npair && jb < lenb)
                                    ? (p * static_cast<int>(lenb) + jb) * 4 : -1);
      }
      double4_t acc_re = {0, 0, 0, 0}, acc_im = {0, 0, 0, 0};
      if (wave_live) {
#pragma unroll
        for (int kt = 0; kt < kF64KSteps; ++kt) {
          if (kt < ksteps) {
            const double* r = cur + (kt * 4 + (lane >> 4)) * kF64Stride + (lane & 15);
            acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[kt], r[0], acc_re, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[kt], r[kF64TbC], acc_im, 0, 0, 0);
          }
        }
      }
      if (more) store_tile(nxt);
      if (wave_live) {
        // f64 MFMA layout: item i of lane l is row l/16 + 4 i, col l%16
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int p = wave * 16 + (lane >> 4) + 4 * i;
          const int64_t jb = ib0 + (lane & 15);
          if (p <= npair && jb < lenb)
            scatter_complex<double, double>(acc_re[i], acc_im[i], p, npair, jb,
                                            fbv[i], rt, accrow, slots);
        }
      }
      __syncthreads();
    }
    double* dst = rowacc + ia * 2 * lenb;
    for (int64_t x = tid; x < 2 * lenb; x += kFThreads) dst[x] = accrow[x];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kFThreads, 2)
k_fused_f64(int npair, int64_t lena, int64_t lenb,
            const double* __restrict__ amat, const int32_t* __restrict__ pta,
            const int32_t* __restrict__ ptb, const int32_t* __restrict__ fwa,
            const int32_t* __restrict__ fwb, const int32_t* __restrict__ sla,
            int nslot, const double* __restrict__ coeff,
            double* __restrict__ slots, double* __restrict__ rowacc,
            const int32_t* __restrict__ done) {
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  double areg[kF64KSteps];
  f64_a_regs(npair, amat, areg);
  for (int64_t ia = blockIdx.x; ia < lena; ia += gridDim.x)
    f64_row(ia, npair, lena, lenb, pta, ptb, fwa, fwb, sla, nslot, coeff,
            slots, rowacc, smem, areg);
}

// ---------------- FP64 split kernel over the live pair rows ------------------
// The complex128 counterpart of k_split_live: the beta-beta block has
// already been written to the row accumulators by k_beta_gemm_f64, and this
// kernel multiplies the live rows of A' (at most kLiveRows: four waves of
// 16 rows) by D' = D^a + 2 D^b plus the C column on v_mfma_f64_16x16x4f64.
// The D tile is double-buffered as in k_fused_f64: the gather of chunk c+1
// (coefficients one chunk ahead, beta table entries two chunks ahead)
// overlaps the MFMAs of chunk c, with one barrier per chunk.
constexpr int kSF64Threads = 256;
constexpr int kSF64Groups = kSF64Threads / kF64TbC;
constexpr int kSF64GIter = kF64KPad / kSF64Groups;

__global__ void __launch_bounds__(kSF64Threads, 2)
k_split_live_f64(int norb, int compressed, int npair, int64_t lena,
                 int64_t lenb, const double* __restrict__ amat,
                 const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
                 const int32_t* __restrict__ fwa, const int32_t* __restrict__ sla,
                 int nslot, const double* __restrict__ coeff,
                 double* __restrict__ slots, double* __restrict__ rowacc,
                 const int32_t* __restrict__ done) {
  static_assert(kLiveRows == 4 * 16, "one 16-row tile per wave");
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  const int Kg = K + 1;                     // gathered rows: pairs + C
  const int ksteps = (Kg + 3) / 4;
  double* dt = reinterpret_cast<double*>(smem);     // 2 x [kF64KPad][kF64Stride]
  double* drow = dt + 2 * kF64Tile;
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(drow + kF64KPad);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  int32_t* lpair = rt.ssgn + kRowTabs;      // [kLiveRows] pair of live row

  const int t = tid % kF64TbC;              // gather column
  const int g = tid / kF64TbC;              // gather pair offset
  const int nchunks = static_cast<int>((lenb + kF64TbC - 1) / kF64TbC);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(double));
  const __amdgpu_buffer_rsrc_t rpb = buffer_of(ptb, static_cast<int64_t>(K) * lenb * 4);

  for (int64_t ia = blockIdx.x; ia < lena; ia += gridDim.x) {
    load_row_tables<double>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                            kSF64Threads, K);
    __syncthreads();
    // live rows: the pairs with sbase >= 0 in increasing order (wave 0,
    // ballot compaction), then row npair, then padding (-1)
    if (wave == 0) {
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
    }
    // row npair of A' for this alpha string (occupied diagonal pairs)
    for (int k = tid; k < Kg; k += kSF64Threads) {
      double v = 0.0;
      for (int i = 0; i < norb; ++i) {
        const int dp = compressed ? i + i * (i + 1) / 2 : i * (norb + 1);
        if (rt.sbase[dp] == -2)
          v += k < K ? amat[static_cast<int64_t>(dp) * K + k]
                     : amat[static_cast<int64_t>(npair) * K + dp];
      }
      drow[k] = v;
    }
    __syncthreads();
    // A' fragment of this wave's 16 live rows (A[r = l%16][k = l/16]), the
    // slot targets of this lane's output rows (l/16 + 4 i) and the alpha
    // gather entries of this thread's pair rows
    double areg[kF64KSteps];
    {
      const int p = lpair[wave * 16 + (lane & 15)];
#pragma unroll
      for (int kt = 0; kt < kF64KSteps; ++kt) {
        const int kk = kt * 4 + (lane >> 4);
        double v = 0.0;
        if (p == npair) {
          if (kk < Kg) v = drow[kk];
        } else if (p >= 0) {
          if (kk < K) v = amat[static_cast<int64_t>(p) * K + kk];
          else if (kk == K) v = amat[static_cast<int64_t>(npair) * K + p];
        }
        areg[kt] = v;
      }
    }
    // slot bases are even element offsets: bit 0 carries the sign
    int64_t obase[4];                       // -1 none, -3 rowacc
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int po = lpair[wave * 16 + (lane >> 4) + 4 * i];
      obase[i] = po == npair ? -3
                             : (po < 0 ? -1 : rt.sbase[po] | (rt.ssgn[po] < 0 ? 1 : 0));
    }
    const int crow_b = static_cast<int>(ia * lenb * 2 * sizeof(double));
    double* racc = rowacc + ia * lenb * 2;

    int32_t bt[kSF64GIter];                 // beta table entries (ahead)
    auto load_bt = [&](int64_t ib0) {
      const int ib = static_cast<int>(ib0) + t;
#pragma unroll
      for (int it = 0; it < kSF64GIter; ++it) {
        const int q = g + it * kSF64Groups;
        bt[it] = bload_i32(rpb, (ib < lenb && q < K)
                                    ? (q * static_cast<int>(lenb) + ib) * 4 : -1);
      }
    };
    // gather of one chunk into an LDS tile: coefficients loaded and stored
    // in one pass (the MFMAs of the previous chunk are already issued)
    auto gather = [&](int64_t ib0, double* tile) {
      const int ib = static_cast<int>(ib0) + t;
      const bool colok = ib < lenb;
      constexpr int kHalf = kSF64GIter / 2;   // two passes bound the VGPRs
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double2 ca[kHalf], cb[kHalf];
        int32_t as[kHalf];
#pragma unroll
        for (int j = 0; j < kHalf; ++j) {
          const int it = h * kHalf + j;
          const int q = g + it * kSF64Groups;
          const int32_t off = q < Kg ? rt.aoff[q] : -1;
          as[j] = off >= 0 ? rt.asgn[q] : 0;
          ca[j] = bload_d2(rc, (colok && off >= 0) ? (off + 2 * ib) * 8 : -1);
          cb[j] = bload_d2(rc, bt[it] ? crow_b + 16 * (iabs(bt[it]) - 1) : -1);
        }
#pragma unroll
        for (int j = 0; j < kHalf; ++j) {
          const int it = h * kHalf + j;
          const int q = g + it * kSF64Groups;
          const double sa = static_cast<double>(as[j]);
          const double sb = bt[it] > 0 ? 2.0 : (bt[it] < 0 ? -2.0 : 0.0);
          tile[q * kF64Stride + t] = sa * ca[j].x + sb * cb[j].x;
          tile[q * kF64Stride + kF64TbC + t] = sa * ca[j].y + sb * cb[j].y;
        }
      }
    };
    // prologue: tile 0 in buffer 0, table entries of chunk 1 in registers
    load_bt(0);
    gather(0, dt);
    if (nchunks > 1) load_bt(kF64TbC);
    __syncthreads();

    for (int chunk = 0; chunk < nchunks; ++chunk) {
      const int64_t ib0 = static_cast<int64_t>(chunk) * kF64TbC;
      const double* cur = dt + (chunk & 1) * kF64Tile;
      double* nxt = dt + ((chunk + 1) & 1) * kF64Tile;
      const bool more = chunk + 1 < nchunks;
      double4_t acc_re = {0, 0, 0, 0}, acc_im = {0, 0, 0, 0};
#pragma unroll
      for (int kt = 0; kt < kF64KSteps; ++kt) {
        if (kt < ksteps) {
          const double* r = cur + (kt * 4 + (lane >> 4)) * kF64Stride + (lane & 15);
          acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[kt], r[0], acc_re, 0, 0, 0);
          acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[kt], r[kF64TbC], acc_im, 0, 0, 0);
        }
      }
      if (more) {
        gather(ib0 + kF64TbC, nxt);
        if (chunk + 2 < nchunks) load_bt(ib0 + 2 * kF64TbC);
      }
      // f64 MFMA layout: item i of lane l is row l/16 + 4 i, col l%16
      const int64_t jb = ib0 + (lane & 15);
      if (jb < lenb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t sb = obase[i];
          if (sb >= 0) {
            const double sg = (sb & 1) ? -1.0 : 1.0;
            stream_store2(slots + (sb & ~int64_t(1)) + 2 * jb, sg * acc_re[i],
                          sg * acc_im[i]);
          } else if (sb == -3) {
            double2* d = reinterpret_cast<double2*>(racc + 2 * jb);
            double2 o = *d;
            o.x += acc_re[i];
            o.y += acc_im[i];
            *d = o;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------- FP64 split kernel, column-owned waves ---------------------
// Same product as k_split_live_f64, organised without a shared D tile: the
// live rows' A' fragments sit in LDS (written once per alpha string), and
// each wave owns 16 columns of a 64-column chunk.  A wave gathers the D'
// values of its columns straight into its MFMA B-operand registers
// (lane l: k = 4 kt + l/16, column l%16), then runs all four 16-row tiles
// of live rows over them.  The chunk loop has no barrier, so one wave's
// gather overlaps another wave's MFMAs.
constexpr int kCF64Threads = 256;
constexpr int kCF64Cols = 64;               // columns per chunk (4 waves x 16)
constexpr int kCF64KMax = 128;              // npair + 1 limit (norb 15)
constexpr int kCF64AS = kLiveRows + 4;      // doubles per LDS A' k-row (padding)
constexpr int kCF64Batch = 7;               // k steps gathered per load batch
constexpr int kCF64Parts = 4;               // k-step parts held in registers

// KMAX: the gather depth the kernel is built for, kF64KMax (112: the
// sectors up to norb 14) or kCF64KMax (128: norb 15); the narrower build
// keeps fewer k steps in registers.
template <int KMAX>
__global__ void __launch_bounds__(kCF64Threads, 2)
k_split_cols_f64(int norb, int compressed, int npair, int64_t lena,
                 int64_t lenb, const double* __restrict__ amat,
                 const int32_t* __restrict__ pta, const int32_t* __restrict__ ptb,
                 const int32_t* __restrict__ fwa, const int32_t* __restrict__ sla,
                 int nslot, const double* __restrict__ coeff,
                 double* __restrict__ slots, double* __restrict__ rowacc,
                 const int32_t* __restrict__ done) {
  static_assert((KMAX / 4) % kCF64Parts == 0, "parts of the k steps");
  extern __shared__ __align__(16) char smem[];
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int K = npair;
  const int Kg = K + 1;                     // gathered rows: pairs + C
  const int ksteps = (Kg + 3) / 4;
  double* alds = reinterpret_cast<double*>(smem);     // [KMAX][kCF64AS]
  double* drow = alds + KMAX * kCF64AS;
  RowTables rt;
  rt.sbase = reinterpret_cast<int64_t*>(drow + kF64KPad);
  rt.aoff = reinterpret_cast<int32_t*>(rt.sbase + kRowTabs);
  rt.asgn = rt.aoff + kRowTabs;
  rt.ssgn = rt.asgn + kRowTabs;
  int32_t* lpair = rt.ssgn + kRowTabs;      // [kLiveRows] pair of live row

  const int nchunks = static_cast<int>((lenb + kCF64Cols - 1) / kCF64Cols);
  const __amdgpu_buffer_rsrc_t rc = buffer_of(coeff, lena * lenb * 2 * sizeof(double));
  const __amdgpu_buffer_rsrc_t rpb = buffer_of(ptb, static_cast<int64_t>(K) * lenb * 4);
  const int kq = lane >> 4;                 // k within a k step
  const int cl = lane & 15;                 // column within the wave's 16

  for (int64_t ia = blockIdx.x; ia < lena; ia += gridDim.x) {
    load_row_tables<double>(ia, K, lena, lenb, nslot, pta, fwa, sla, rt, tid,
                            kCF64Threads, K);
    __syncthreads();
    if (wave == 0) {
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
    }
    for (int k = tid; k < Kg; k += kCF64Threads) {
      double v = 0.0;
      for (int i = 0; i < norb; ++i) {
        const int dp = compressed ? i + i * (i + 1) / 2 : i * (norb + 1);
        if (rt.sbase[dp] == -2)
          v += k < K ? amat[static_cast<int64_t>(dp) * K + k]
                     : amat[static_cast<int64_t>(npair) * K + dp];
      }
      drow[k] = v;
    }
    __syncthreads();
    // A'[live row m][k] into LDS as [k][m]
    for (int x = tid; x < KMAX * kLiveRows; x += kCF64Threads) {
      const int k = x / kLiveRows;
      const int m = x - k * kLiveRows;
      const int p = lpair[m];
      double v = 0.0;
      if (p == npair) {
        if (k < Kg) v = drow[k];
      } else if (p >= 0) {
        if (k < K) v = amat[static_cast<int64_t>(p) * K + k];
        else if (k == K) v = amat[static_cast<int64_t>(npair) * K + p];
      }
      alds[k * kCF64AS + m] = v;
    }
    __syncthreads();
    const int crow_b = static_cast<int>(ia * lenb * 2 * sizeof(double));
    double* racc = rowacc + ia * lenb * 2;

    for (int chunk = 0; chunk < nchunks; ++chunk) {
      const int64_t jb = static_cast<int64_t>(chunk) * kCF64Cols + wave * 16 + cl;
      const bool colok = jb < lenb;
      const int ib = static_cast<int>(colok ? jb : 0);
      // the k steps in kCF64Parts parts: B fragments D'[4 kt + kq][column]
      // of one part in registers, accumulators of all four row tiles kept
      double4_t accr[kLiveRows / 16], acci[kLiveRows / 16];
#pragma unroll
      for (int r = 0; r < kLiveRows / 16; ++r) {
        accr[r] = double4_t{0, 0, 0, 0};
        acci[r] = double4_t{0, 0, 0, 0};
      }
#pragma unroll 1
      for (int hk = 0; hk < kCF64Parts; ++hk) {
        constexpr int kH = (KMAX / 4) / kCF64Parts;
        if (hk * kH >= ksteps) break;      // parts no MFMA reads
        double bre[kH], bim[kH];
#pragma unroll
        for (int b0 = 0; b0 < kH; b0 += kCF64Batch) {
          int32_t bt[kCF64Batch], as[kCF64Batch];
          double2 ca[kCF64Batch], cb[kCF64Batch];
#pragma unroll
          for (int j = 0; j < kCF64Batch; ++j) {
            const int q = 4 * (hk * kH + b0 + j) + kq;
            bt[j] = (b0 + j < kH)
                        ? bload_i32(rpb, (colok && q < K)
                                             ? (q * static_cast<int>(lenb) + ib) * 4 : -1)
                        : 0;
          }
#pragma unroll
          for (int j = 0; j < kCF64Batch; ++j) {
            const int q = 4 * (hk * kH + b0 + j) + kq;
            const int32_t off = (colok && b0 + j < kH && q < Kg) ? rt.aoff[q] : -1;
            as[j] = off >= 0 ? rt.asgn[q] : 0;
            ca[j] = bload_d2(rc, off >= 0 ? (off + 2 * ib) * 8 : -1);
            cb[j] = bload_d2(rc, bt[j] ? crow_b + 16 * (iabs(bt[j]) - 1) : -1);
          }
#pragma unroll
          for (int j = 0; j < kCF64Batch; ++j) {
            if (b0 + j < kH) {
              const double sa = static_cast<double>(as[j]);
              const double sb = bt[j] > 0 ? 2.0 : (bt[j] < 0 ? -2.0 : 0.0);
              bre[b0 + j] = sa * ca[j].x + sb * cb[j].x;
              bim[b0 + j] = sa * ca[j].y + sb * cb[j].y;
            }
          }
        }
        // A from LDS: A[r = l%16][k = l/16]
#pragma unroll
        for (int r = 0; r < kLiveRows / 16; ++r) {
          const double* ab = alds + (hk * kH * 4 + kq) * kCF64AS + r * 16 + cl;
          // all of the part's A values first: the LDS reads are in flight
          // together instead of one wait per MFMA pair
          double av[kH];
#pragma unroll
          for (int kt = 0; kt < kH; ++kt) av[kt] = ab[kt * 4 * kCF64AS];
#pragma unroll
          for (int kt = 0; kt < kH; ++kt) {
            if (hk * kH + kt < ksteps) {
              accr[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kt], bre[kt], accr[r], 0, 0, 0);
              acci[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kt], bim[kt], acci[r], 0, 0, 0);
            }
          }
        }
      }
      // f64 MFMA layout: item i of lane l is row l/16 + 4 i, col l%16
      if (colok) {
#pragma unroll
        for (int r = 0; r < kLiveRows / 16; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int po = lpair[r * 16 + kq + 4 * i];
            if (po == npair) {
              double2* d = reinterpret_cast<double2*>(racc + 2 * jb);
              double2 o = *d;
              o.x += accr[r][i];
              o.y += acci[r][i];
              *d = o;
            } else if (po >= 0) {
              const int64_t sb = rt.sbase[po];
              const double sg = static_cast<double>(rt.ssgn[po]);
              stream_store2(slots + sb + 2 * jb, sg * accr[r][i], sg * acci[r][i]);
            }
          }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// slot reduction + epilogue
// ---------------------------------------------------------------------------
// Element idx = (ia, ib) of a slot reduction: rowacc + the nslot slots in
// FP64, out = alpha * that + beta * cin, acc += the FP64 value; returns
// |out|^2 (shared by k_slot_reduce and k_slot_reduce_group).  P48: the
// slots are complex128 values packed in 48 bits per double
// (stream_store_p48).
template <typename T, bool P48 = false>
__device__ __forceinline__ double slot_element(
    int64_t idx, int64_t ia, int64_t ib, int nslot, int64_t lenb,
    const void* __restrict__ slotv, const T* __restrict__ rowacc,
    double alpha_re, double alpha_im, double beta_re, double beta_im,
    const T* cin, T* out, double* __restrict__ acc) {
  double re = static_cast<double>(rowacc[2 * idx]);
  double im = static_cast<double>(rowacc[2 * idx + 1]);
  if constexpr (P48) {
    const uint32_t* s = static_cast<const uint32_t*>(slotv) + (ia * nslot * lenb + ib) * 3;
    for (int k = 0; k < nslot; ++k) {
      const u3v v = stream_load_p48(s);
      re += p48_re(v);
      im += p48_im(v);
      s += 3 * lenb;
    }
  } else {
    const T* s = static_cast<const T*>(slotv) + (ia * nslot * lenb + ib) * 2;
    // complex64 slots: eight loads in flight per thread (complex128 slots
    // keep one, which measured faster); the FP64 sum keeps slot order
    constexpr int kU = sizeof(T) == 4 ? 8 : 1;
    typedef T t2v __attribute__((ext_vector_type(2)));
    int k = 0;
    for (; k + kU <= nslot; k += kU) {
      T vr[kU], vi[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const t2v v = __builtin_nontemporal_load(
            reinterpret_cast<const t2v*>(s + 2 * lenb * u));
        vr[u] = v[0];
        vi[u] = v[1];
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        re += static_cast<double>(vr[u]);
        im += static_cast<double>(vi[u]);
      }
      s += 2 * lenb * kU;
    }
    for (; k < nslot; ++k) {
      const t2v v = __builtin_nontemporal_load(reinterpret_cast<const t2v*>(s));
      re += static_cast<double>(v[0]);
      im += static_cast<double>(v[1]);
      s += 2 * lenb;
    }
  }
  double vr = alpha_re * re - alpha_im * im;
  double vi = alpha_re * im + alpha_im * re;
  if (beta_re != 0.0 || beta_im != 0.0) {
    const double cr = static_cast<double>(cin[2 * idx]);
    const double ci = static_cast<double>(cin[2 * idx + 1]);
    vr += beta_re * cr - beta_im * ci;
    vi += beta_re * ci + beta_im * cr;
  }
  const T orr = static_cast<T>(vr);
  const T oi = static_cast<T>(vi);
  out[2 * idx] = orr;
  out[2 * idx + 1] = oi;
  if (acc) {
    acc[2 * idx] += vr;
    acc[2 * idx + 1] += vi;
  }
  return static_cast<double>(orr) * orr + static_cast<double>(oi) * oi;
}

template <typename T, bool P48>
__global__ void __launch_bounds__(kThreads)
k_slot_reduce(int nslot, int64_t lena, int64_t lenb, double inv_lenb,
              const void* __restrict__ slots, const T* __restrict__ rowacc,
              double alpha_re, double alpha_im, double beta_re, double beta_im,
              const T* cin, T* out, double* __restrict__ acc,
              double* __restrict__ partials, const int32_t* __restrict__ done) {
  __shared__ double red[kWaves];
  if (done && *done) return;
  const int64_t n = lena * lenb;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  double sq = 0.0;
  if (idx < n) {
    int64_t ia = static_cast<int64_t>(static_cast<double>(idx) * inv_lenb);
    int64_t ib = idx - ia * lenb;
    if (ib < 0) { --ia; ib += lenb; } else if (ib >= lenb) { ++ia; ib -= lenb; }
    sq = slot_element<T, P48>(idx, ia, ib, nslot, lenb, slots, rowacc,
                              alpha_re, alpha_im, beta_re, beta_im, cin, out,
                              acc);
  }
  if (partials) {
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < kWaves; ++w) t += red[w];
      partials[blockIdx.x] = t;
    }
  }
}

// ---------------- grouped launches over the sectors of a batch --------------
// One launch covers one Taylor step of every sector of the group: global
// alpha row gr belongs to the last sector with row0 <= gr (a uniform
// binary search), which then runs the row body of k_fused_bf16x3.  The
// rows of all sectors together fill the GPU where each small sector alone
// would not.
template <int NW, int WPE>
__global__ void __launch_bounds__(NW * 64, WPE)
k_fused_group(const GroupSector* __restrict__ gs, int nsec, int64_t rows,
              int step) {
  extern __shared__ __align__(16) char smem[];
  bf16x8 ahi[8 / NW][kBfKSteps], alo[8 / NW][kBfKSteps];
  // a contiguous run of rows per workgroup: the A' fragments are rebuilt
  // only where the run crosses into another sector
  const int64_t per = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * per;
  const int64_t r1 = r0 + per < rows ? r0 + per : rows;
  int lo = 0;
  if (r0 < rows) {
    int hi = nsec - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (gs[mid].row0 <= r0) lo = mid;
      else hi = mid - 1;
    }
  }
  int cur = -1;
  for (int64_t gr = r0; gr < r1; ++gr) {
    while (lo + 1 < nsec && gs[lo + 1].row0 <= gr) ++lo;
    const GroupSector& sc = gs[lo];
    if (step == 0 && sc.apply_out == nullptr) continue;
    if (lo != cur) {
      fused_a_frags<NW, false>(sc.npair, sc.amat, ahi, alo);
      cur = lo;
    }
    fused_row<NW, false>(gr - sc.row0, sc.norb, sc.compressed, sc.npair,
                         sc.lena, sc.lenb, sc.amat, sc.pta, sc.ptb, sc.fwa,
                         sc.fwb, sc.sla, sc.nslot,
                         static_cast<const float*>(step == 0 ? sc.coeff0 : sc.term),
                         static_cast<float*>(sc.slots),
                         static_cast<float*>(sc.rowacc), smem, ahi, alo);
  }
}

// The complex128 counterpart: the row body of k_fused_f64 over the rows of
// every sector of the group, A' registers reloaded where a workgroup's run
// of rows crosses into another sector.
__global__ void __launch_bounds__(kFThreads, 2)
k_fused_group_f64(const GroupSector* __restrict__ gs, int nsec, int64_t rows,
                  int step) {
  extern __shared__ __align__(16) char smem[];
  double areg[kF64KSteps];
  const int64_t per = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * per;
  const int64_t r1 = r0 + per < rows ? r0 + per : rows;
  int lo = 0;
  if (r0 < rows) {
    int hi = nsec - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (gs[mid].row0 <= r0) lo = mid;
      else hi = mid - 1;
    }
  }
  int cur = -1;
  for (int64_t gr = r0; gr < r1; ++gr) {
    while (lo + 1 < nsec && gs[lo + 1].row0 <= gr) ++lo;
    const GroupSector& sc = gs[lo];
    if (step == 0 && sc.apply_out == nullptr) continue;
    if (lo != cur) {
      f64_a_regs(sc.npair, sc.amat, areg);
      cur = lo;
    }
    f64_row(gr - sc.row0, sc.npair, sc.lena, sc.lenb, sc.pta, sc.ptb, sc.fwa,
            sc.fwb, sc.sla, sc.nslot,
            static_cast<const double*>(step == 0 ? sc.coeff0 : sc.term),
            static_cast<double*>(sc.slots), static_cast<double*>(sc.rowacc),
            smem, areg);
  }
}

// Step 0: apply_out = rowacc + slots.  Step k >= 1: term = (ar + i ai)
// (rowacc + slots) in place and result += term in FP64.  Block gb belongs
// to the last sector with blk0 <= gb.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_slot_reduce_group(const GroupSector* __restrict__ gs, int nsec, int step,
                    double ar, double ai) {
  const int64_t gb = blockIdx.x;
  int lo = 0, hi = nsec - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gs[mid].blk0 <= gb) lo = mid;
    else hi = mid - 1;
  }
  const GroupSector& sc = gs[lo];
  T* out = static_cast<T*>(step == 0 ? sc.apply_out : sc.term);
  if (out == nullptr) return;
  const int64_t idx = (gb - sc.blk0) * kThreads + threadIdx.x;
  if (idx >= sc.lena * sc.lenb) return;
  const int64_t ia = idx / sc.lenb;
  const int64_t ib = idx - ia * sc.lenb;
  (void)slot_element<T>(idx, ia, ib, sc.nslot, sc.lenb,
                        static_cast<const T*>(sc.slots),
                        static_cast<const T*>(sc.rowacc), step == 0 ? 1.0 : ar,
                        step == 0 ? 0.0 : ai, 0.0, 0.0, nullptr, out,
                        step == 0 ? nullptr : sc.result);
}

size_t row_tables_bytes() {
  return kRowTabs * (sizeof(int64_t) + 3 * sizeof(int32_t));
}

size_t bf16_lds_bytes(int64_t lenb, bool split = false) {
  return 2 * kBfCols * kBfStride * sizeof(__bf16) + row_tables_bytes() +
         (split ? kBfKPad * sizeof(double) : 2 * lenb * sizeof(float));
}

size_t split_live_lds_bytes(int tbc, bool rowlds, int64_t lenb) {
  return 2 * (2 * tbc) * kBfStride * sizeof(__bf16) + kBfKPad * sizeof(double) +
         row_tables_bytes() + kLiveRows * sizeof(int32_t) +
         (rowlds ? lenb * 8 : 0);
}

size_t split_live_f64_lds_bytes() {
  return (2 * kF64Tile + kF64KPad) * sizeof(double) + row_tables_bytes() +
         kLiveRows * sizeof(int32_t);
}

size_t split_cols_f64_lds_bytes(int kmax = kCF64KMax) {
  return (kmax * kCF64AS + kF64KPad) * sizeof(double) + row_tables_bytes() +
         kLiveRows * sizeof(int32_t);
}

size_t f64_lds_bytes(int64_t lenb) {
  return 2 * kF64Tile * sizeof(double) + row_tables_bytes() +
         2 * lenb * sizeof(double);
}

constexpr size_t kLdsLimit = 160 * 1024;

unsigned fused_grid(int64_t lena) {
  return static_cast<unsigned>(lena < 65536 ? lena : 65536);
}

}  // namespace

// Whether the fused path can run this sector / operator.
bool fused_supported(int precision, int npair, int64_t lena, int64_t lenb) {
  if (npair + 1 > kMPad) return false;
  // byte offsets of C and of the pair tables are 32-bit (buffer loads)
  const int64_t lim = int64_t(1) << 31;
  if (static_cast<int64_t>(npair) * lenb * 4 >= lim) return false;
  if (precision == FQEHIP_C128)
    return npair <= kF64KMax && f64_lds_bytes(lenb) <= kLdsLimit &&
           lena * lenb * 2 * 8 < lim;
  return npair <= kBfKPad && bf16_lds_bytes(lenb) <= kLdsLimit &&
         lena * lenb * 2 * 4 < lim;
}

// FP64 split: the live-rows kernel's LDS does not grow with lenb (the row
// accumulator is the global rowacc), and its gather depth is kF64KMax.
// The FP64 split kernels: the column-owned k_split_cols_f64 (default tile)
// gathers up to kCF64KMax pair rows, the D-tile k_split_live_f64 kF64KMax.
bool split_f64_supported(int npair, int64_t lena, int64_t lenb, int tile) {
  const int64_t lim = int64_t(1) << 31;
  const bool dtile = tile == FQEHIP_TILE_SPLIT_F64_DTILE;
  const int kmax = dtile ? kF64KMax : kCF64KMax;
  const size_t lds = dtile ? split_live_f64_lds_bytes() : split_cols_f64_lds_bytes();
  return npair + 1 <= kmax && static_cast<int64_t>(npair) * lenb * 4 < lim &&
         lena * lenb * 2 * 8 < lim && lds <= kLdsLimit;
}

bool split_supported(int npair, int64_t lena, int64_t lenb) {
  return npair + 1 <= kBfKPad && fused_supported(FQEHIP_C64, npair, lena, lenb) &&
         bf16_lds_bytes(lenb, true) <= kLdsLimit;
}

int launch_fused(int precision, int norb, int npair, int64_t lena,
                 int64_t lenb, const double* a, const int32_t* pta,
                 const int32_t* ptb, const int32_t* fwa, const int32_t* fwb,
                 const int32_t* sla, int nslot, const void* coeff, void* slots,
                 void* rowacc, bool split, int tile, hipStream_t st,
                 const int32_t* done) {
  if (lena == 0 || lenb == 0) return FQEHIP_OK;
  const int compressed = npair == norb * (norb + 1) / 2 ? 1 : 0;
  if (split && precision == FQEHIP_C128) {
    if (!split_f64_supported(npair, lena, lenb, tile) ||
        norb * norb / 4 + 1 > kLiveRows) {
      set_error("the FP64 split apply needs npair + 1 <= %d (%d with the "
                "shared D tile) and norb^2/4 + 1 <= %d",
                kCF64KMax, kF64KMax, kLiveRows);
      return FQEHIP_ERR_INVALID;
    }
    if (tile == FQEHIP_TILE_SPLIT_F64_DTILE) {
      k_split_live_f64<<<fused_grid(lena), kSF64Threads, split_live_f64_lds_bytes(), st>>>(
          norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, sla, nslot,
          static_cast<const double*>(coeff), static_cast<double*>(slots),
          static_cast<double*>(rowacc), done);
      FQEHIP_CHECK_LAUNCH("k_split_live_f64");
      return FQEHIP_OK;
    }
    if (npair + 1 <= kF64KMax)
      k_split_cols_f64<kF64KMax><<<fused_grid(lena), kCF64Threads,
                                   split_cols_f64_lds_bytes(kF64KMax), st>>>(
          norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, sla, nslot,
          static_cast<const double*>(coeff), static_cast<double*>(slots),
          static_cast<double*>(rowacc), done);
    else
      k_split_cols_f64<kCF64KMax><<<fused_grid(lena), kCF64Threads,
                                    split_cols_f64_lds_bytes(kCF64KMax), st>>>(
          norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, sla, nslot,
          static_cast<const double*>(coeff), static_cast<double*>(slots),
          static_cast<double*>(rowacc), done);
    FQEHIP_CHECK_LAUNCH("k_split_cols_f64");
    return FQEHIP_OK;
  }
  if (split) {
    if (precision != FQEHIP_C64 || !split_supported(npair, lena, lenb)) {
      set_error("the split apply needs complex64 and npair + 1 <= %d", kBfKPad);
      return FQEHIP_ERR_INVALID;
    }
    if (tile != FQEHIP_TILE_SPLIT_ALLROWS) {
      // live rows: at most norb^2/4 excitations to other strings plus row
      // npair (57 for the largest sector here, norb 15)
      if (norb * norb / 4 + 1 > kLiveRows) {
        set_error("the live-rows kernel needs norb^2/4 + 1 <= %d", kLiveRows);
        return FQEHIP_ERR_INVALID;
      }
      // 16 columns per chunk with row ia in LDS while three workgroups
      // still fit a CU's LDS, else 32 columns and the row from L2
      const size_t lds16 = split_live_lds_bytes(16, true, lenb);
      if (lds16 <= kLdsLimit / 3) {
        k_split_live<4, 3, 16, true><<<fused_grid(lena), 256, lds16, st>>>(
            norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, sla, nslot,
            static_cast<const float*>(coeff), static_cast<float*>(slots),
            static_cast<float*>(rowacc), done);
      } else {
        k_split_live<4, 3, 32, false>
            <<<fused_grid(lena), 256, split_live_lds_bytes(32, false, lenb), st>>>(
                norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, sla, nslot,
                static_cast<const float*>(coeff), static_cast<float*>(slots),
                static_cast<float*>(rowacc), done);
      }
      FQEHIP_CHECK_LAUNCH("k_split_live");
      return FQEHIP_OK;
    }
    // all npair + 1 rows (the live-rows kernel's cross-check); three
    // workgroups per CU: no LDS accumulator row, and the higher occupancy
    // outweighs a few spilled registers
    k_fused_bf16x3<4, 3, true><<<fused_grid(lena), 256, bf16_lds_bytes(lenb, true), st>>>(
        norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, fwb, sla, nslot,
        static_cast<const float*>(coeff), static_cast<float*>(slots),
        static_cast<float*>(rowacc), done);
    FQEHIP_CHECK_LAUNCH("k_fused_bf16x3_split");
    return FQEHIP_OK;
  }
  if (precision == FQEHIP_C128) {
    const size_t lds = f64_lds_bytes(lenb);
    k_fused_f64<<<fused_grid(lena), kFThreads, lds, st>>>(
        npair, lena, lenb, a, pta, ptb, fwa, fwb, sla, nslot,
        static_cast<const double*>(coeff), static_cast<double*>(slots),
        static_cast<double*>(rowacc), done);
    FQEHIP_CHECK_LAUNCH("k_fused_f64");
  } else {
    const size_t lds = bf16_lds_bytes(lenb);
    // variant 4: 4 waves x 2 row tiles, two workgroups per CU;
    // variant 8: 8 waves x 1 row tile, one workgroup per CU
    if (tile == FQEHIP_TILE_ROWS8)
      k_fused_bf16x3<8, 2, false><<<fused_grid(lena), 512, lds, st>>>(
          norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, fwb, sla, nslot,
          static_cast<const float*>(coeff), static_cast<float*>(slots),
          static_cast<float*>(rowacc), done);
    else
      k_fused_bf16x3<4, 2, false><<<fused_grid(lena), 256, lds, st>>>(
          norb, compressed, npair, lena, lenb, a, pta, ptb, fwa, fwb, sla, nslot,
          static_cast<const float*>(coeff), static_cast<float*>(slots),
          static_cast<float*>(rowacc), done);
    FQEHIP_CHECK_LAUNCH("k_fused_bf16x3");
  }
  return FQEHIP_OK;
}

int launch_slot_reduce(int precision, bool p48, int nslot, int64_t lena,
                       int64_t lenb, const void* slots, const void* rowacc,
                       double ar, double ai, double br, double bi,
                       const void* cin, void* out, double* acc,
                       double* partials, hipStream_t st, const int32_t* done) {
  const int64_t n = lena * lenb;
  if (n == 0) return FQEHIP_OK;
  const unsigned nb = blocks_for(n, kThreads);
  if (precision == FQEHIP_C128 && p48) {
    k_slot_reduce<double, true><<<nb, kThreads, 0, st>>>(
        nslot, lena, lenb, 1.0 / lenb, slots,
        static_cast<const double*>(rowacc), ar, ai, br, bi,
        static_cast<const double*>(cin), static_cast<double*>(out), acc,
        partials, done);
  } else if (precision == FQEHIP_C128) {
    k_slot_reduce<double, false><<<nb, kThreads, 0, st>>>(
        nslot, lena, lenb, 1.0 / lenb, slots,
        static_cast<const double*>(rowacc), ar, ai, br, bi,
        static_cast<const double*>(cin), static_cast<double*>(out), acc,
        partials, done);
  } else {
    k_slot_reduce<float, false><<<nb, kThreads, 0, st>>>(
        nslot, lena, lenb, 1.0 / lenb, slots,
        static_cast<const float*>(rowacc), ar, ai, br, bi,
        static_cast<const float*>(cin), static_cast<float*>(out), acc,
        partials, done);
  }
  FQEHIP_CHECK_LAUNCH("k_slot_reduce");
  return FQEHIP_OK;
}

// Grouped launches (batch code): rows of every sector of the group, then
// their reductions; complex64: rows4 row body (two 4-wave workgroups per CU
// measured faster here than one 8-wave one); complex128: the k_fused_f64
// row body; LDS sized for the longest beta row.
int launch_group_rows(int precision, const GroupSector* d_secs, int nsec,
                      int64_t rows, int64_t max_lenb, int step,
                      hipStream_t st) {
  if (rows == 0 || nsec == 0) return FQEHIP_OK;
  const unsigned grid = static_cast<unsigned>(rows < 4096 ? rows : 4096);
  if (precision == FQEHIP_C128) {
    const size_t lds = f64_lds_bytes(max_lenb);
    if (lds > kLdsLimit) {
      set_error("grouped rows: lenb=%lld exceeds the LDS row accumulator",
                static_cast<long long>(max_lenb));
      return FQEHIP_ERR_INVALID;
    }
    k_fused_group_f64<<<grid, kFThreads, lds, st>>>(d_secs, nsec, rows, step);
    FQEHIP_CHECK_LAUNCH("k_fused_group_f64");
    return FQEHIP_OK;
  }
  const size_t lds = bf16_lds_bytes(max_lenb);
  if (lds > kLdsLimit) {
    set_error("grouped rows: lenb=%lld exceeds the LDS row accumulator",
              static_cast<long long>(max_lenb));
    return FQEHIP_ERR_INVALID;
  }
  // about four workgroups per CU at a time, several rounds: runs of
  // consecutive rows while the load stays balanced
  k_fused_group<4, 2><<<grid, 256, lds, st>>>(d_secs, nsec, rows, step);
  FQEHIP_CHECK_LAUNCH("k_fused_group");
  return FQEHIP_OK;
}

int launch_group_reduce(int precision, const GroupSector* d_secs, int nsec,
                        int64_t blocks, int step, double ar, double ai,
                        hipStream_t st) {
  if (blocks == 0 || nsec == 0) return FQEHIP_OK;
  if (blocks > 0x7fffffffLL) {
    set_error("grouped reduction: %lld blocks", static_cast<long long>(blocks));
    return FQEHIP_ERR_INVALID;
  }
  if (precision == FQEHIP_C128)
    k_slot_reduce_group<double><<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(
        d_secs, nsec, step, ar, ai);
  else
    k_slot_reduce_group<float><<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(
        d_secs, nsec, step, ar, ai);
  FQEHIP_CHECK_LAUNCH("k_slot_reduce_group");
  return FQEHIP_OK;
}

int group_reduce_threads() { return kThreads; }

}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_fused_tables(int norb, int nele, int64_t nstrings,
                                   const int32_t* d_dexc, int compressed,
                                   const int32_t* d_ptab, int32_t* d_fwd,
                                   int32_t* d_slot, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(norb >= 0 && norb <= kMaxNorb && nele >= 0 && nele <= norb,
                   "bad sector norb=%d nele=%d", norb, nele);
  FQEHIP_CHECK_ARG(compressed == 0 || compressed == 1, "compressed=%d",
                   compressed);
  FQEHIP_CHECK_ARG(d_ptab && d_fwd, "NULL table pointer");
  const int npair = compressed ? norb * (norb + 1) / 2 : norb * norb;
  const int64_t total = static_cast<int64_t>(npair) * nstrings;
  hipStream_t st = as_stream(stream);
  if (total == 0) return FQEHIP_OK;
  const unsigned nb = blocks_for(total, kThreads) < 4096u ? blocks_for(total, kThreads) : 4096u;
  k_fill<<<nb, kThreads, 0, st>>>(total, 0, d_fwd);
  FQEHIP_CHECK_LAUNCH("k_fill");
  k_forward_table<<<blocks_for(total, kThreads), kThreads, 0, st>>>(
      total, nstrings, d_ptab, d_fwd);
  FQEHIP_CHECK_LAUNCH("k_forward_table");
  if (d_slot) {
    k_fill<<<nb, kThreads, 0, st>>>(total, -1, d_slot);
    FQEHIP_CHECK_LAUNCH("k_fill");
    const int nexc = nele * (norb - nele + 1);
    if (nexc > 0) {
      FQEHIP_CHECK_ARG(d_dexc != nullptr, "d_dexc is NULL");
      k_slot_table<<<blocks_for(nstrings, kThreads), kThreads, 0, st>>>(
          norb, nexc, nstrings, d_dexc, compressed, d_slot);
      FQEHIP_CHECK_LAUNCH("k_slot_table");
    }
  }
  return FQEHIP_OK;
}
