// (c) FP64 pair GEMM of the unfused apply, Y = A B (+ Y), for an A matrix
// small enough to stay resident in LDS.
//
// A (m x k, leading dimension lda) is the (npair+1) x npair pair matrix of
// the operator; B (k x ncol) is the dvec, one pair plane per row; Y (m x
// ncol) gets one plane per output pair (reference src/fqe/fqe_data.py:647-
// 657, the einsum of V with the dvec).  With m, k ~ 100 and ncol ~ 2e7 the
// product is bound by the FP64 matrix cores (m/4 flops per byte of B), so
// the kernel is built to keep v_mfma_f64_16x16x4f64 busy:
//   - each workgroup copies A once into LDS, already laid out as MFMA
//     fragments (fragment (t, s) = the 64 lane values A[16t + l%16][4s +
//     l/16], 512 contiguous bytes, read with one conflict-free ds_read_b64);
//   - workgroups are persistent: each wave walks 32-column panels of B;
//   - a lane loads B[4s + l/16][c..c+1] (16 bytes, c = panel + 2(l%16)) and
//     feeds two MFMA tiles, one for the even and one for the odd columns, so
//     a k step costs one 16-byte load per lane for 2 RT MFMAs;
//   - B is read through a ring of 8 (4 for m > 112) k steps in registers.
#include "fqehip_common.h"

namespace fqehip {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

constexpr int kPgWaves = 8;                    // waves per workgroup
constexpr int kPgThreads = 64 * kPgWaves;
constexpr int kPgCols = 32;                    // columns per wave panel
// k steps of B in flight: 8, or 4 where 8 or more row tiles of
// accumulators leave too few registers
template <int RT>
constexpr int pg_depth() { return RT >= 8 ? 4 : 8; }
constexpr int kPgMaxRowTiles = 9;              // m <= 144
constexpr int64_t kPgLdsLimit = 160 * 1024;    // LDS of one CU

template <int RT>
__global__ void __launch_bounds__(kPgThreads)
k_pair_gemm_lds(int m, int k, const double* __restrict__ a, int lda,
                int64_t ncol, const double* __restrict__ b,
                double* __restrict__ y, int accumulate,
                const int32_t* __restrict__ done) {
  constexpr int kPgDepth = pg_depth<RT>();
  extern __shared__ double afrag[];           // [RT][ks][64]
  if (done && *done) return;
  const int ks = (k + 3) >> 2;
  const int nfrag = RT * ks * 64;
  for (int e = threadIdx.x; e < nfrag; e += kPgThreads) {
    const int l = e & 63;
    const int ts = e >> 6;
    const int t = ts / ks;
    const int s = ts - t * ks;
    const int r = 16 * t + (l & 15);
    const int kk = 4 * s + (l >> 4);
    afrag[e] = (r < m && kk < k) ? a[static_cast<int64_t>(r) * lda + kk] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int kl = lane >> 4;
  const int64_t npanel = (ncol + kPgCols - 1) / kPgCols;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kPgWaves;
  for (int64_t pan = static_cast<int64_t>(blockIdx.x) * kPgWaves + wave;
       pan < npanel; pan += stride) {
    const int64_t c = pan * kPgCols + 2 * (lane & 15);
    const bool cok = c < ncol;                 // ncol is even: c + 1 too
    // The B value of k step s for this lane.  No value is masked after its
    // load (that would make the wave wait for the load at once): rows past
    // k read row k-1 and meet zero A fragments, and a lane past ncol reads
    // column 0 into output columns it never stores.
    const double* bc = b + (cok ? c : 0);
    auto load = [&](int s) -> double2_t {
      const int kk = 4 * s + kl < k ? 4 * s + kl : k - 1;
      return *reinterpret_cast<const double2_t*>(bc + static_cast<int64_t>(kk) * ncol);
    };
    double2_t ring[kPgDepth];
#pragma unroll
    for (int j = 0; j < kPgDepth; ++j) ring[j] = load(j);
    double4_t acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      acc[t][0] = double4_t{0, 0, 0, 0};
      acc[t][1] = double4_t{0, 0, 0, 0};
    }
    auto step = [&](int s, const double2_t bv) {
      const double* af = afrag + static_cast<int64_t>(s) * 64 + lane;
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        const double av = af[static_cast<int64_t>(t) * ks * 64];
        acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[0], acc[t][0], 0, 0, 0);
        acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[1], acc[t][1], 0, 0, 0);
      }
      // keep the A reads of later steps below this one (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    };
    // full groups of kPgDepth steps without a branch, so that each use of
    // a ring entry waits only for its own load; then the tail
    int s0 = 0;
    for (; s0 + kPgDepth <= ks; s0 += kPgDepth) {
#pragma unroll
      for (int j = 0; j < kPgDepth; ++j) {
        // the refill is issued after the MFMAs that read ring[j], so it
        // lands in the same registers and nothing waits for it here
        step(s0 + j, ring[j]);
        ring[j] = load(s0 + j + kPgDepth);
      }
    }
#pragma unroll
    for (int j = 0; j < kPgDepth; ++j)
      if (s0 + j < ks) step(s0 + j, ring[j]);
    if (!cok) continue;
    // item i of lane l: row 16t + l/16 + 4i, columns c (tile 0), c+1 (tile 1)
#pragma unroll
    for (int t = 0; t < RT; ++t) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * t + kl + 4 * i;
        if (r < m) {
          double2_t* dst = reinterpret_cast<double2_t*>(y + static_cast<int64_t>(r) * ncol + c);
          double2_t v = double2_t{acc[t][0][i], acc[t][1][i]};
          if (accumulate) v += *dst;
          __builtin_nontemporal_store(v, dst);
        }
      }
    }
  }
}

// Larger A (complex operators in their real 2 x 2 block form, spin-orbital
// operators): m up to 8 RT row tiles, RT <= 4.  A no longer fits in LDS, so it is
// streamed through it in chunks of kPcSteps k steps (16 k), already in
// fragment layout, double-buffered: the global loads of chunk c+1 are
// issued before the MFMAs of chunk c and written to LDS after them, one
// barrier per chunk.  The 8 waves of a workgroup are 8 row groups of RT
// row tiles over the same 32-column panel (its B rows come once from HBM,
// then from the cache); wave g owns tiles g, g + 8, g + 16, ..., so the
// tiles past m (skipped) are spread over the waves and their SIMDs; B goes through the same register ring as above,
// and k is padded to a multiple of the ring depth (zero A fragments, rows
// of B clamped).
constexpr int kPcRowGroups = kPgWaves;
constexpr int kPcSteps = 4;                    // k steps per A chunk
constexpr int kPcDepth = 8;                    // ring depth, 2 chunks
constexpr int kPcMaxRowTiles = 4;              // m <= 512
// A fragment takes 68 doubles of LDS: with the 4-double pad, the 64 lanes
// of one staging write (4 rows x 16 k of the chunk) fall on 32 distinct
// bank pairs, each twice, instead of 4 (16-way conflicts)
constexpr int kPcFrag = 68;

template <int RT>
__global__ void __launch_bounds__(kPgThreads)
k_pair_gemm_chunked(int m, int k, const double* __restrict__ a, int lda,
                    int64_t ncol, const double* __restrict__ b,
                    double* __restrict__ y, int accumulate,
                    const int32_t* __restrict__ done) {
  constexpr int T = kPcRowGroups * RT;         // row tiles of the workgroup
  constexpr int kChunk = T * kPcSteps * kPcFrag;   // doubles per LDS buffer
  constexpr int kPer = T * kPcSteps * 64 / kPgThreads;   // staged per thread
  constexpr int kChunksPerGroup = kPcDepth / kPcSteps;
  static_assert(T * kPcSteps * 64 % kPgThreads == 0, "chunk split");
  extern __shared__ double afr[];              // [2][T][kPcSteps][kPcFrag]
  if (done && *done) return;
  const int tid = threadIdx.x;
  const int rg = tid >> 6;
  const int lane = tid & 63;
  const int ntile = (m + 15) >> 4;
  const int kl = lane >> 4;
  const int ks = (((k + 3) >> 2) + kPcDepth - 1) / kPcDepth * kPcDepth;
  // staging: element e = tid + kPgThreads i of chunk c is A[r][16c + q],
  // r = r0 + 32 i, q = tid % 16 (128 contiguous bytes per row); its LDS
  // slot is fragment (r/16, q/4), lane r%16 + 16(q%4), i.e. a fixed slot of
  // the thread plus 2 kPcSteps kPcFrag i doubles.  A is read through a buffer resource
  // ending at its last element; an element outside A (row >= m or column
  // >= k) gets the offset -1, out of range, and reads zero.
  const int q = tid & 15;
  const int r0 = tid >> 4;
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(a), 0,
      static_cast<int>((static_cast<int64_t>(m - 1) * lda + k) * 8), 0x00020000);
  const int slot = ((r0 >> 4) * kPcSteps + (q >> 2)) * kPcFrag + (r0 & 15) + 16 * (q & 3);
  double stage[kPer];
  auto fetch = [&](int c) {
    const int kk = 16 * c + q;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int r = r0 + (kPgThreads / 16) * i;
      const int off = (r < m && kk < k) ? (r * lda + kk) * 8 : -1;
      const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(ra, off, 0, 0);
      stage[i] = __builtin_bit_cast(double, v);
    }
  };
  auto put = [&](int buf) {
    double* dst = afr + buf * kChunk + slot;
#pragma unroll
    for (int i = 0; i < kPer; ++i) dst[2 * kPcSteps * kPcFrag * i] = stage[i];
  };
  const int64_t npanel = (ncol + kPgCols - 1) / kPgCols;
  for (int64_t pan = blockIdx.x; pan < npanel; pan += gridDim.x) {
    const int64_t c = pan * kPgCols + 2 * (lane & 15);
    const bool cok = c < ncol;
    const double* bc = b + (cok ? c : 0);
    auto load = [&](int s) -> double2_t {
      const int kk = 4 * s + kl < k ? 4 * s + kl : k - 1;
      // streamed once: non-temporal, so that it does not push A out of L2
      return __builtin_nontemporal_load(
          reinterpret_cast<const double2_t*>(bc + static_cast<int64_t>(kk) * ncol));
    };
    double2_t ring[kPcDepth];
#pragma unroll
    for (int j = 0; j < kPcDepth; ++j) ring[j] = load(j);
    double4_t acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      acc[t][0] = double4_t{0, 0, 0, 0};
      acc[t][1] = double4_t{0, 0, 0, 0};
    }
    __syncthreads();                           // previous panel's reads done
    fetch(0);
    put(0);
    __syncthreads();
    for (int s0 = 0; s0 < ks; s0 += kPcDepth) {
#pragma unroll
      for (int h = 0; h < kChunksPerGroup; ++h) {
        // chunk ch + 1 is fetched and stored unconditionally: after the last
        // chunk every column is past k (offset -1, zero, no memory access)
        // and lands in the buffer the next panel re-stages behind a
        // barrier; a branch here would make the wait for each ring entry
        // count only the loads of the path without the fetch
        const int ch = s0 / kPcSteps + h;
        fetch(ch + 1);
        const double* buf = afr + (ch & 1) * kChunk + rg * kPcSteps * kPcFrag + lane;
#pragma unroll
        for (int j = 0; j < kPcSteps; ++j) {
          const int jr = h * kPcSteps + j;
#pragma unroll
          for (int t = 0; t < RT; ++t) {
            if (t * kPcRowGroups + rg < ntile) {   // uniform over the wave
              const double av = buf[((t * kPcRowGroups) * kPcSteps + j) * kPcFrag];
              acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, ring[jr][0], acc[t][0], 0, 0, 0);
              acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, ring[jr][1], acc[t][1], 0, 0, 0);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          ring[jr] = load(s0 + jr + kPcDepth);
        }
        put((ch + 1) & 1);
        __syncthreads();
      }
    }
    if (!cok) continue;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * (t * kPcRowGroups + rg) + kl + 4 * i;
        if (r < m) {
          double2_t* dst = reinterpret_cast<double2_t*>(y + static_cast<int64_t>(r) * ncol + c);
          double2_t v = double2_t{acc[t][0][i], acc[t][1][i]};
          if (accumulate) v += *dst;
          __builtin_nontemporal_store(v, dst);
        }
      }
    }
  }
}

template <int RT>
int launch_chunked(int m, int k, const double* a, int lda, int64_t ncol,
                   const double* b, double* y, int accumulate, hipStream_t st,
                   const int32_t* done) {
  const size_t lds = 2 * static_cast<size_t>(kPcRowGroups) * RT * kPcSteps *
                     kPcFrag * sizeof(double);
  const int64_t npanel = (ncol + kPgCols - 1) / kPgCols;
  const unsigned blocks = static_cast<unsigned>(npanel < 4096 ? npanel : 4096);
  k_pair_gemm_chunked<RT><<<blocks, kPgThreads, lds, st>>>(
      m, k, a, lda, ncol, b, y, accumulate, done);
  FQEHIP_CHECK_LAUNCH("k_pair_gemm_chunked");
  return FQEHIP_OK;
}

template <int RT>
int launch_rt(int m, int k, const double* a, int lda, int64_t ncol,
              const double* b, double* y, int accumulate, hipStream_t st,
              const int32_t* done) {
  const int ks = (k + 3) >> 2;
  const size_t lds = static_cast<size_t>(RT) * ks * 64 * sizeof(double);
  const int64_t npanel = (ncol + kPgCols - 1) / kPgCols;
  int64_t blocks = (npanel + kPgWaves - 1) / kPgWaves;
  // persistent: a few workgroups per CU at most (1 per CU at norb 14)
  if (blocks > 1024) blocks = 1024;
  k_pair_gemm_lds<RT><<<static_cast<unsigned>(blocks), kPgThreads, lds, st>>>(
      m, k, a, lda, ncol, b, y, accumulate, done);
  FQEHIP_CHECK_LAUNCH("k_pair_gemm_lds");
  return FQEHIP_OK;
}

}  // namespace

// The shapes the two kernels take: an even column count, 16-byte aligned
// B and Y rows, and m at most 8 x 4 row tiles (A resident in LDS when its
// fragments fit; else streamed in chunks, m <= 512).
bool pair_gemm_lds_ok(int m, int k, int64_t ncol, const void* b,
                      const void* y) {
  if (m <= 0 || k <= 0 || ncol <= 0) return false;
  return (m + 15) / 16 <= kPcRowGroups * kPcMaxRowTiles && ncol % 2 == 0 &&
         (reinterpret_cast<uintptr_t>(b) & 15) == 0 &&
         (reinterpret_cast<uintptr_t>(y) & 15) == 0;
}

int launch_pair_gemm_lds(int m, int k, const double* a, int lda, int64_t ncol,
                         const double* b, double* y, int accumulate,
                         hipStream_t st, const int32_t* done) {
  const int rt = (m + 15) / 16;
  const int64_t resident = static_cast<int64_t>(rt) * ((k + 3) / 4) * 64 * 8;
  if (rt > kPgMaxRowTiles || resident > kPgLdsLimit) {
    switch ((rt + kPcRowGroups - 1) / kPcRowGroups) {
      case 1: return launch_chunked<1>(m, k, a, lda, ncol, b, y, accumulate, st, done);
      case 2: return launch_chunked<2>(m, k, a, lda, ncol, b, y, accumulate, st, done);
      case 3: return launch_chunked<3>(m, k, a, lda, ncol, b, y, accumulate, st, done);
      default: return launch_chunked<4>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    }
  }
  switch (rt) {
    case 1: return launch_rt<1>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 2: return launch_rt<2>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 3: return launch_rt<3>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 4: return launch_rt<4>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 5: return launch_rt<5>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 6: return launch_rt<6>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 7: return launch_rt<7>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 8: return launch_rt<8>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    case 9: return launch_rt<9>(m, k, a, lda, ncol, b, y, accumulate, st, done);
    default: break;
  }
  set_error("k_pair_gemm_lds: m=%d above %d rows", m, 16 * kPgMaxRowTiles);
  return FQEHIP_ERR_UNSUPPORTED;
}

}  // namespace fqehip
