// Spin-orbital apply on a full particle-number set of sectors (FqeDataSet).
//
// The set holds every (nalpha, nbeta) sector with nalpha + nbeta = nele and
// both counts <= norb, in ascending nbeta, concatenated into one complex128
// vector of length L (each sector row-major lena x lenb).  Spin-orbital
// index p < norb is alpha orbital p, p >= norb beta orbital p - norb, and
// M = 2 norb.  For a target determinant I and an ordered pair (p, q) there
// is at most one source determinant J with <I|a+_p a_q|J> != 0; every
// kernel here is a gather over that J, so no kernel writes to a location
// another thread writes and no atomics are used.
//
// Signs follow the reference: within one spin (-1)^(occupied orbitals
// strictly between i and j) (src/fqe/fci_graph.py:204-244); across spins
// (-1)^nalpha(J) times the parities of the sector-to-sector maps, which
// count the occupied orbitals above the orbital
// (src/fqe/fqe_data_set.py:737-796, src/fqe/fci_graph_set.py:137-150).
//
// Kernels:
//   k_set_dvec  D[pq][I] = sum_J <I|E_pq|J> C[J]   (one thread per I)
//   k_set_expand_v  the real block form [[Vr, -Vi], [Vi, Vr]] of V, with
//               the block form of the row h1' under it for the apply
//   k_set_gemm  Y = V D as one real FP64 matrix-core GEMM, k = 2 M^2 and
//               m = 2 M^2 (+ 2 rows that hold sum_pq h1'[pq] D[pq])
//   k_set_out   out[I]   = Y[h1' row][I] + sum_pq <I|E_pq|J> Y[pq][J]
//               for the 1+2-body apply, or
//               out[I]   = sum_pq h1[pq] <I|E_pq|J> C[J] for the 1-body one.  D and Y are planar per pair:
// row 2 pq holds the real parts, row 2 pq + 1 the imaginary parts, L
// doubles each, so the GEMM sees them as real 2 M^2 x L matrices.
#include "fqehip_common.h"

namespace fqehip {
namespace {

constexpr int kSetMaxNorb = 16;
constexpr int kSetMaxSec = kSetMaxNorb + 1;
constexpr int kSetBlock = 256;
constexpr int kSetPairChunk = 64;   // pairs per k_set_dvec block row
// k_set_gemm tiles: a wave owns 16 x 16 tiles of kGemmRows row tiles by
// kGemmCols column tiles, so each A value it loads feeds kGemmCols MFMAs and
// each B value kGemmRows; a block of 4 waves covers 64 rows x 256 columns.
constexpr int kGemmRows = 4;
constexpr int kGemmCols = 4;
constexpr int kGemmKc = 32;   // k values of a staged in LDS per chunk
static_assert(kGemmRows * 16 * kGemmKc % 256 == 0, "even LDS fill");

typedef double set_double4 __attribute__((ext_vector_type(4)));

struct SetDesc {
  int64_t total;                 // L
  int norb, nele, nsec, nbmin;   // sector s has nbeta = nbmin + s
  int64_t off[kSetMaxSec + 1];   // sector offsets, off[nsec] = L
  int32_t lenb[kSetMaxSec];
  int32_t binom[kSetMaxNorb + 1][kSetMaxNorb + 1];
};

__device__ __forceinline__ int popc(uint32_t x) { return __popc(x); }

// Address of an nele-electron string in the reference's order (ascending
// occupation tuples in lexicographic order):
// C(n, k) - 1 - sum_t C(n - 1 - c_t, k + 1 - t), occupied c_1 < ... < c_k.
__device__ __forceinline__ int32_t lex_rank(uint32_t s, int n, int k,
                                            const SetDesc& d) {
  int32_t r = d.binom[n][k] - 1;
  int t = 1;
  while (s) {
    int c = __ffs(s) - 1;
    s &= s - 1;
    r -= d.binom[n - 1 - c][k + 1 - t];
    ++t;
  }
  return r;
}

// Inverse of lex_rank: the mirrored set has colex rank C(n,k) - 1 - r.
__device__ __forceinline__ uint32_t lex_unrank(int32_t r, int n, int k,
                                               const SetDesc& d) {
  int32_t rem = d.binom[n][k] - 1 - r;
  uint32_t s = 0;
  int x = n - 1;
  for (int i = k; i >= 1; --i) {
    while (d.binom[x][i] > rem) --x;   // stops at x = i - 1 at the latest
    s |= 1u << (n - 1 - x);
    rem -= d.binom[x][i];
    --x;
  }
  return s;
}

__device__ __forceinline__ int bits_between32(uint32_t s, int i, int j) {
  int lo = i < j ? i : j, hi = i < j ? j : i;
  uint32_t mask = ((1u << hi) - 1u) & ~((2u << lo) - 1u);
  return popc(s & mask);
}

// Target determinant of sector `sec` with strings (a, b).  For pair (p, q)
// returns false when <I|E_pq|J> = 0 for every J; otherwise the flat index
// of J in the set vector and its sign.
__device__ __forceinline__ bool set_source(int p, int q, int sec, int na,
                                           int nb, uint32_t a, uint32_t b,
                                           const SetDesc& d, int64_t* jflat,
                                           double* sign) {
  const int n = d.norb;
  if (p < n && q < n) {            // alpha a+_i a_j, same sector
    const int i = p, j = q;
    if (!((a >> i) & 1u)) return false;
    if (i != j && ((a >> j) & 1u)) return false;
    const uint32_t src = (a ^ (1u << i)) | (1u << j);
    *sign = (bits_between32(a, i, j) & 1) ? -1.0 : 1.0;
    const int32_t ia = lex_rank(src, n, na, d);
    const int64_t lenb = d.lenb[sec];
    *jflat = d.off[sec] + ia * lenb + lex_rank(b, n, nb, d);
    return true;
  }
  if (p >= n && q >= n) {          // beta a+_i a_j, same sector
    const int i = p - n, j = q - n;
    if (!((b >> i) & 1u)) return false;
    if (i != j && ((b >> j) & 1u)) return false;
    const uint32_t src = (b ^ (1u << i)) | (1u << j);
    *sign = (bits_between32(b, i, j) & 1) ? -1.0 : 1.0;
    const int64_t lenb = d.lenb[sec];
    *jflat = d.off[sec] + lex_rank(a, n, na, d) * lenb + lex_rank(src, n, nb, d);
    return true;
  }
  if (p < n) {                     // a+_{i alpha} a_{j beta}: J in (na-1, nb+1)
    const int i = p, j = q - n;
    if (!((a >> i) & 1u) || ((b >> j) & 1u)) return false;
    const int ssec = sec + 1;
    if (ssec >= d.nsec) return false;
    const uint32_t sa = a ^ (1u << i), sb = b | (1u << j);
    const int par = (na - 1) + popc(a >> (i + 1)) + popc(b >> (j + 1));
    *sign = (par & 1) ? -1.0 : 1.0;
    const int64_t lenb = d.lenb[ssec];
    *jflat = d.off[ssec] + lex_rank(sa, n, na - 1, d) * lenb +
             lex_rank(sb, n, nb + 1, d);
    return true;
  }
  {                                // a+_{i beta} a_{j alpha}: J in (na+1, nb-1)
    const int i = p - n, j = q;
    if (!((b >> i) & 1u) || ((a >> j) & 1u)) return false;
    const int ssec = sec - 1;
    if (ssec < 0) return false;
    const uint32_t sa = a | (1u << j), sb = b ^ (1u << i);
    const int par = na + popc(a >> (j + 1)) + popc(b >> (i + 1));
    *sign = (par & 1) ? -1.0 : 1.0;
    const int64_t lenb = d.lenb[ssec];
    *jflat = d.off[ssec] + lex_rank(sa, n, na + 1, d) * lenb +
             lex_rank(sb, n, nb - 1, d);
    return true;
  }
}

// Sector, counts and strings of flat index g (g < L).
__device__ __forceinline__ void locate(int64_t g, const SetDesc& d, int* sec,
                                       int* na, int* nb, uint32_t* a,
                                       uint32_t* b) {
  int s = 0;
  while (s + 1 < d.nsec && g >= d.off[s + 1]) ++s;
  *sec = s;
  *nb = d.nbmin + s;
  *na = d.nele - *nb;
  const int64_t r = g - d.off[s];
  const int64_t lenb = d.lenb[s];
  *a = lex_unrank(static_cast<int32_t>(r / lenb), d.norb, *na, d);
  *b = lex_unrank(static_cast<int32_t>(r % lenb), d.norb, *nb, d);
}

// One thread per determinant and blockIdx.y-th chunk of kSetPairChunk
// pairs, so that the grid has enough blocks to fill the chip.
__global__ void __launch_bounds__(kSetBlock)
k_set_dvec(SetDesc d, int m2, const double2* __restrict__ c,
           double* __restrict__ dvec) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kSetBlock + threadIdx.x;
  if (g >= d.total) return;
  int sec, na, nb;
  uint32_t a, b;
  locate(g, d, &sec, &na, &nb, &a, &b);
  const int m = 2 * d.norb;
  const int pq0 = blockIdx.y * kSetPairChunk;
  const int pq1 = pq0 + kSetPairChunk < m2 ? pq0 + kSetPairChunk : m2;
  for (int pq = pq0; pq < pq1; ++pq) {
    int64_t j;
    double sg;
    double re = 0.0, im = 0.0;
    if (set_source(pq / m, pq % m, sec, na, nb, a, b, d, &j, &sg)) {
      const double2 x = c[j];
      re = sg * x.x;
      im = sg * x.y;
    }
    dvec[static_cast<int64_t>(2 * pq) * d.total + g] = re;
    dvec[static_cast<int64_t>(2 * pq + 1) * d.total + g] = im;
  }
}

// a[(2 pq + u) * 2 m2 + 2 rs + w] = [[Vr, -Vi], [Vi, Vr]][u][w] of
// V[pq][rs] (complex128 [m2][m2]) for rows < 2 m2; rows 2 m2 and 2 m2 + 1
// (when rows = 2 m2 + 2) hold the same block form of the row h1[rs].  One
// thread per element of a.
__global__ void __launch_bounds__(kSetBlock)
k_set_expand_v(int m2, int rows, const double2* __restrict__ v,
               const double2* __restrict__ h1, double* __restrict__ a) {
  const int64_t n2 = 2LL * m2;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kSetBlock + threadIdx.x;
  if (e >= rows * n2) return;
  const int64_t row = e / n2, col = e % n2;
  const double2 h = row < n2 ? v[(row >> 1) * m2 + (col >> 1)] : h1[col >> 1];
  const int u = static_cast<int>(row & 1), w = static_cast<int>(col & 1);
  a[e] = (u == w) ? h.x : (u == 0 ? -h.y : h.y);
}

// y[m][ncol] = a[m][k] b[k][ncol], all real row-major FP64, on
// v_mfma_f64_16x16x4f64 (A operand: row lane%16, k lane/16; B operand:
// k lane/16, column lane%16; result item i: row lane/16 + 4 i, column
// lane%16, the layout k_gemm_f64 stores).  Out-of-range rows, columns and
// k are read as zero and not written.
__global__ void __launch_bounds__(kSetBlock, 2)
k_set_gemm(int m, int k, const double* __restrict__ a, int64_t ncol,
           const double* __restrict__ b, double* __restrict__ y) {
  // The block's 4 waves share its kGemmRows*16 rows of a: each chunk of
  // kGemmKc k values of those rows is staged in LDS once, with every
  // thread of the block taking part.  The b operands of k step k0 + 4 are
  // loaded while the MFMAs of step k0 run.  No wave leaves early, so every
  // wave reaches every barrier.
  __shared__ double sa[kGemmRows * 16][kGemmKc + 1];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t col0 =
      (static_cast<int64_t>(blockIdx.x) * 4 + wave) * (16 * kGemmCols);
  const int row0 = blockIdx.y * kGemmRows * 16;
  const int kl = lane >> 4;
  const int li = lane & 15;
  set_double4 acc[kGemmRows][kGemmCols];
#pragma unroll
  for (int t = 0; t < kGemmRows; ++t)
#pragma unroll
    for (int u = 0; u < kGemmCols; ++u) acc[t][u] = set_double4{0, 0, 0, 0};
  auto load_b = [&](int k0, double* bn) {
    const int kk = k0 + kl;
#pragma unroll
    for (int u = 0; u < kGemmCols; ++u) {
      const int64_t col = col0 + u * 16 + li;
      bn[u] = (kk < k && col < ncol) ? b[static_cast<int64_t>(kk) * ncol + col]
                                     : 0.0;
    }
  };
  double bv[kGemmCols];
  load_b(0, bv);
  for (int kc0 = 0; kc0 < k; kc0 += kGemmKc) {
    __syncthreads();   // the previous chunk's reads of sa are done
#pragma unroll
    for (int j = 0; j < kGemmRows * 16 * kGemmKc / kSetBlock; ++j) {
      const int e = threadIdx.x + j * kSetBlock;
      const int r = e / kGemmKc, c = e % kGemmKc;
      const int gr = row0 + r, gc = kc0 + c;
      sa[r][c] = (gr < m && gc < k) ? a[static_cast<int64_t>(gr) * k + gc]
                                    : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGemmKc; ks += 4) {
      double bn[kGemmCols];
      load_b(kc0 + ks + 4, bn);   // zeros past the end of k
#pragma unroll
      for (int t = 0; t < kGemmRows; ++t) {
        const double av = sa[t * 16 + li][ks + kl];
#pragma unroll
        for (int u = 0; u < kGemmCols; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[u],
                                                           acc[t][u], 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < kGemmCols; ++u) bv[u] = bn[u];
    }
  }
#pragma unroll
  for (int u = 0; u < kGemmCols; ++u) {
    const int64_t col = col0 + u * 16 + li;
    if (col >= ncol) continue;
#pragma unroll
    for (int t = 0; t < kGemmRows; ++t) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + t * 16 + kl + 4 * i;
        if (r < m) y[static_cast<int64_t>(r) * ncol + col] = acc[t][u][i];
      }
    }
  }
}

// kPlanarG: out[I] = Y[2 m2][I] + i Y[2 m2 + 1][I]
//                    + sum_pq <I|E_pq|J> (Y[2 pq][J] + i Y[2 pq + 1][J]);
// otherwise: out[I] = sum_pq hg[pq] <I|E_pq|J> C[J], C interleaved.
template <bool kPlanarG>
__global__ void __launch_bounds__(kSetBlock)
k_set_out(SetDesc d, int m2, const double2* __restrict__ hg,
          const double* __restrict__ gsrc, double2* __restrict__ out) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kSetBlock + threadIdx.x;
  if (g >= d.total) return;
  int sec, na, nb;
  uint32_t a, b;
  locate(g, d, &sec, &na, &nb, &a, &b);
  const int m = 2 * d.norb;
  double re = 0.0, im = 0.0;
  if (kPlanarG) {
    re = gsrc[static_cast<int64_t>(2 * m2) * d.total + g];
    im = gsrc[static_cast<int64_t>(2 * m2 + 1) * d.total + g];
  }
  for (int pq = 0; pq < m2; ++pq) {
    int64_t j;
    double sg;
    if (set_source(pq / m, pq % m, sec, na, nb, a, b, d, &j, &sg)) {
      if (kPlanarG) {
        re += sg * gsrc[static_cast<int64_t>(2 * pq) * d.total + j];
        im += sg * gsrc[static_cast<int64_t>(2 * pq + 1) * d.total + j];
      } else {
        const double xr = gsrc[2 * j], xi = gsrc[2 * j + 1];
        const double hr = sg * hg[pq].x, hi = sg * hg[pq].y;
        re += hr * xr - hi * xi;
        im += hr * xi + hi * xr;
      }
    }
  }
  out[g] = make_double2(re, im);
}

int64_t host_binom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

int make_desc(int norb, int nele, SetDesc* d) {
  FQEHIP_CHECK_ARG(norb >= 1 && norb <= kSetMaxNorb,
                   "set apply: norb %d outside 1..%d", norb, kSetMaxNorb);
  FQEHIP_CHECK_ARG(nele >= 0 && nele <= 2 * norb,
                   "set apply: nele %d outside 0..%d", nele, 2 * norb);
  *d = SetDesc{};
  d->norb = norb;
  d->nele = nele;
  d->nbmin = nele > norb ? nele - norb : 0;
  const int nbmax = nele < norb ? nele : norb;
  d->nsec = nbmax - d->nbmin + 1;
  for (int n = 0; n <= kSetMaxNorb; ++n)
    for (int k = 0; k <= kSetMaxNorb; ++k)
      d->binom[n][k] = static_cast<int32_t>(host_binom(n, k));
  int64_t off = 0;
  for (int s = 0; s < d->nsec; ++s) {
    const int nb = d->nbmin + s, na = nele - nb;
    d->off[s] = off;
    d->lenb[s] = static_cast<int32_t>(host_binom(norb, nb));
    off += host_binom(norb, na) * host_binom(norb, nb);
  }
  d->off[d->nsec] = off;
  d->total = off;
  return FQEHIP_OK;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int64_t fqehip_set_length(int norb, int nele) {
  SetDesc d;
  if (make_desc(norb, nele, &d) != FQEHIP_OK) return -1;
  return d.total;
}

// Workspace: D (2 m2 L doubles), Y ((2 m2 + 2) L doubles), then the real
// block form of V with the h1' row ((2 m2 + 2) 2 m2 doubles).
extern "C" int64_t fqehip_set_workspace_bytes(int norb, int nele) {
  SetDesc d;
  if (make_desc(norb, nele, &d) != FQEHIP_OK) return -1;
  const int64_t m2 = 4LL * norb * norb;
  return ((4 * m2 + 2) * d.total + (2 * m2 + 2) * 2 * m2) *
         static_cast<int64_t>(sizeof(double));
}

extern "C" int fqehip_set_dvec(int norb, int nele, const double* d_coeff,
                               double* d_dvec, void* stream) {
  clear_error();
  SetDesc d;
  if (int rc = make_desc(norb, nele, &d)) return rc;
  FQEHIP_CHECK_ARG(d_coeff != nullptr && d_dvec != nullptr,
                   "fqehip_set_dvec: null pointer");
  if (d.total == 0) return FQEHIP_OK;
  const int m2 = 4 * norb * norb;
  const dim3 grid(blocks_for(d.total, kSetBlock),
                  (m2 + kSetPairChunk - 1) / kSetPairChunk);
  k_set_dvec<<<grid, kSetBlock, 0, as_stream(stream)>>>(d, m2,
                                    reinterpret_cast<const double2*>(d_coeff),
                                    d_dvec);
  FQEHIP_CHECK_LAUNCH("k_set_dvec");
  return FQEHIP_OK;
}

namespace fqehip {
namespace {

// Y = [V; h1'] D on the workspace (h1 may be null: V rows only).
int set_contract(const SetDesc& d, int m2, const double* v, const double* h1,
                 double* ws, hipStream_t st) {
  const int n2 = 2 * m2;
  const int rows = n2 + (h1 != nullptr ? 2 : 0);
  double* dv = ws;
  double* y = dv + static_cast<int64_t>(n2) * d.total;
  double* a = y + static_cast<int64_t>(n2 + 2) * d.total;
  const int64_t na = static_cast<int64_t>(rows) * n2;
  k_set_expand_v<<<blocks_for(na, kSetBlock), kSetBlock, 0, st>>>(
      m2, rows, reinterpret_cast<const double2*>(v),
      reinterpret_cast<const double2*>(h1), a);
  FQEHIP_CHECK_LAUNCH("k_set_expand_v");
  dim3 grid(blocks_for(d.total, 4 * 16 * kGemmCols),
            (rows + kGemmRows * 16 - 1) / (kGemmRows * 16));
  k_set_gemm<<<grid, kSetBlock, 0, st>>>(rows, n2, a, d.total, dv, y);
  FQEHIP_CHECK_LAUNCH("k_set_gemm");
  return FQEHIP_OK;
}

}  // namespace
}  // namespace fqehip

extern "C" int fqehip_set_contract(int norb, int nele, const double* d_v,
                                   double* d_workspace, void* stream) {
  clear_error();
  SetDesc d;
  if (int rc = make_desc(norb, nele, &d)) return rc;
  FQEHIP_CHECK_ARG(d_v != nullptr && d_workspace != nullptr,
                   "fqehip_set_contract: null pointer");
  if (d.total == 0) return FQEHIP_OK;
  return set_contract(d, 4 * norb * norb, d_v, nullptr, d_workspace,
                      as_stream(stream));
}

extern "C" int fqehip_set_apply(int norb, int nele, const double* d_h1,
                                const double* d_v, const double* d_coeff,
                                double* d_out, double* d_workspace,
                                void* stream) {
  clear_error();
  SetDesc d;
  if (int rc = make_desc(norb, nele, &d)) return rc;
  FQEHIP_CHECK_ARG(d_h1 != nullptr && d_coeff != nullptr && d_out != nullptr,
                   "fqehip_set_apply: null pointer");
  FQEHIP_CHECK_ARG(d_v == nullptr || d_workspace != nullptr,
                   "fqehip_set_apply: the two-body apply needs a workspace");
  FQEHIP_CHECK_ARG(d_coeff != d_out,
                   "fqehip_set_apply: output must not alias the input");
  if (d.total == 0) return FQEHIP_OK;
  const int m2 = 4 * norb * norb;
  const unsigned nblk = blocks_for(d.total, kSetBlock);
  hipStream_t s = as_stream(stream);
  const double2* h1 = reinterpret_cast<const double2*>(d_h1);
  const double2* c = reinterpret_cast<const double2*>(d_coeff);
  double2* out = reinterpret_cast<double2*>(d_out);
  if (d_v == nullptr) {
    k_set_out<false><<<nblk, kSetBlock, 0, s>>>(d, m2, h1, d_coeff, out);
    FQEHIP_CHECK_LAUNCH("k_set_out");
    return FQEHIP_OK;
  }
  double* dv = d_workspace;
  double* y = dv + 2LL * m2 * d.total;
  const dim3 dgrid(nblk, (m2 + kSetPairChunk - 1) / kSetPairChunk);
  k_set_dvec<<<dgrid, kSetBlock, 0, s>>>(d, m2, c, dv);
  FQEHIP_CHECK_LAUNCH("k_set_dvec");
  if (int rc = set_contract(d, m2, d_v, d_h1, d_workspace, s)) return rc;
  k_set_out<true><<<nblk, kSetBlock, 0, s>>>(d, m2, nullptr, y, out);
  FQEHIP_CHECK_LAUNCH("k_set_out");
  return FQEHIP_OK;
}
