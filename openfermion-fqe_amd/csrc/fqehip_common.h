// Shared helpers of the fqehip library: status codes, the per-thread error
// message behind fqehip_last_error(), launch checking and small device
// utilities (binomials, Knowles-Handy Z matrix in LDS, string ranks).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>

#include "../../include/fqehip.h"

namespace fqehip {

void set_error(const char* fmt, ...);
void clear_error();

#define FQEHIP_CHECK_ARG(cond, ...)                 \
  do {                                              \
    if (!(cond)) {                                  \
      ::fqehip::set_error(__VA_ARGS__);             \
      return FQEHIP_ERR_INVALID;                    \
    }                                               \
  } while (0)

#define FQEHIP_CHECK_LAUNCH(name)                                          \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) {                                                \
      ::fqehip::set_error("%s: launch failed: %s", name,                    \
                          hipGetErrorString(e_));                          \
      return FQEHIP_ERR_HIP;                                               \
    }                                                                      \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// XCD-aware order of n work items (alpha strings, element blocks):
// workgroup i runs on XCD i % 8 of an MI355X, so XCD x takes the
// contiguous range of items starting at x n/8 (+1 per earlier XCD with a
// remainder item); items that read the same rows then share one L2.  A
// bijection of [0, n).
__device__ __forceinline__ int64_t xcd_order(int64_t i, int64_t n) {
  const int64_t per = n / 8, rem = n % 8;
  const int64_t x = i % 8, j = i / 8;
  return x * per + (x < rem ? x : rem) + j;
}

// Grid-size helper: number of blocks of `block` threads covering n items.
inline unsigned blocks_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// Largest orbital count the string kernels accept (strings are uint64 and
// all table sizes are int32-addressed per spin).
constexpr int kMaxNorb = 32;

__device__ __forceinline__ int64_t dbinom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Knowles-Handy Z matrix: Z[k-1][l-1] (k electron, l orbital, 1-based) is
// sum_{m=norb-l+1}^{norb-k} [C(m, nele-k) - C(m-1, nele-k-1)] for k < nele
// and l - nele for k = nele; the address of a string is sum_k Z[k][occ_k].
// Every thread of the block fills part of `z` (nele*norb entries); the
// caller synchronises.
__device__ __forceinline__ void fill_z_matrix(int32_t* z, int norb, int nele) {
  for (int idx = threadIdx.x; idx < nele * norb; idx += blockDim.x) {
    const int k = idx / norb + 1;
    const int l = idx % norb + 1;
    int64_t v = 0;
    if (k < nele) {
      if (l >= k && l <= norb - nele + k) {
        for (int m = norb - l + 1; m < norb - k + 1; ++m)
          v += dbinom(m, nele - k) - dbinom(m - 1, nele - k - 1);
      }
    } else {
      if (l >= nele) v = l - nele;
    }
    z[idx] = static_cast<int32_t>(v);
  }
}

__device__ __forceinline__ int32_t string_rank(uint64_t s, const int32_t* z,
                                               int norb) {
  int32_t addr = 0;
  int k = 0;
  while (s) {
    const int orb = __ffsll(static_cast<unsigned long long>(s)) - 1;
    addr += z[k * norb + orb];
    ++k;
    s &= s - 1;
  }
  return addr;
}

// Occupied orbitals strictly between i and j.
__device__ __forceinline__ int bits_between(uint64_t s, int i, int j) {
  const int lo = i < j ? i : j;
  const int hi = i < j ? j : i;
  if (hi - lo < 2) return 0;
  const uint64_t mask = ((1ull << hi) - 1ull) & ~((2ull << lo) - 1ull);
  return __popcll(s & mask);
}

// Index of unordered pair {i, j} in the compressed (triangular) basis.
__device__ __forceinline__ int comp_pair(int i, int j) {
  const int hi = i > j ? i : j;
  const int lo = i > j ? j : i;
  return lo + hi * (hi + 1) / 2;
}

// One sector of a grouped batch launch (device array, built by the batch
// code, read by k_fused_group / k_slot_reduce_group): the sector's fused
// tables and operator, its buffers, and its first global row and first
// reduction block (prefix sums over the group).
struct GroupSector {
  int norb, compressed, npair, nslot;
  int64_t lena, lenb;
  int64_t row0;                 // first global alpha row of this sector
  int64_t blk0;                 // first block of the grouped reduction
  const double* amat;
  const int32_t* pta;
  const int32_t* ptb;
  const int32_t* fwa;
  const int32_t* fwb;
  const int32_t* sla;
  // complex64 (float) or complex128 (double) arrays, by the group's precision
  const void* coeff0;           // step 0 input
  void* apply_out;              // step 0 output; NULL: the sector has no step 0
  void* term;                   // steps >= 1: input and output (in place)
  double* result;               // steps >= 1: FP64 Taylor sum
  void* slots;                  // workspace: slot rows
  void* rowacc;                 // workspace: row accumulators
};

// fqehip_apply that, with reuse_tables, takes the Ozaki-digit split apply's
// operator tables from the workspace as an earlier apply of the same sector
// and operator with the same workspace built them (apply.hip).
int apply_reusing_tables(int precision, const fqehip_sector* s,
                         const fqehip_operator* op, const void* d_coeff,
                         double alpha_re, double alpha_im, double beta_re,
                         double beta_im, const void* d_cin, void* d_out,
                         double* d_acc, double* d_norm_partials,
                         const int32_t* d_done, void* d_workspace,
                         void* stream, bool reuse_tables);

}  // namespace fqehip
