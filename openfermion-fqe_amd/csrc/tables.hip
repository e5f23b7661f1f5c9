// (a) String tables on the GPU: strings in address order, de-excitation
// rows per target (dexc), excitation maps per orbital pair, and the dense
// pair tables the apply kernels gather through.
//
// Every table is a pure function of (norb, nele): strings come from
// unranking the address in the lexicographic combination order, source
// addresses from the Knowles-Handy Z matrix (rebuilt in LDS by every
// workgroup), signs from the parity of the occupied orbitals strictly
// between the two orbitals of the excitation.  Bit-for-bit the tables of
// the reference's FciGraph (src/fqe/fci_graph.py:36-154,
// src/fqe/lib/fci_graph.c:27-147).
#include "fqehip_common.h"

namespace fqehip {
namespace {

constexpr int kBlock = 256;

// address -> string: lexicographic unranking of the occupation tuple.
__global__ void __launch_bounds__(kBlock)
k_strings(int norb, int nele, int64_t nstrings, uint64_t* strings) {
  const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= nstrings) return;
  int64_t rem = a;
  int c = 0;
  uint64_t s = 0;
  for (int m = 0; m < nele; ++m) {
    for (;;) {
      const int64_t cnt = dbinom(norb - 1 - c, nele - 1 - m);
      if (rem < cnt) {
        s |= 1ull << c;
        ++c;
        break;
      }
      rem -= cnt;
      ++c;
    }
  }
  strings[a] = s;
}

// One thread per target string: its nexc rows (source, i*norb+j, sign) in
// ascending i*norb+j, i occupied in the target, j == i or j empty.
__global__ void __launch_bounds__(kBlock)
k_dexc(int norb, int nele, int64_t nstrings, const uint64_t* strings,
       int32_t* dexc) {
  __shared__ int32_t z[kMaxNorb * kMaxNorb];
  fill_z_matrix(z, norb, nele);
  __syncthreads();
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nstrings) return;
  const uint64_t s = strings[t];
  const int nexc = nele * (norb - nele + 1);
  int32_t* out = dexc + t * nexc * 3;
  int e = 0;
  for (int i = 0; i < norb; ++i) {
    if (!((s >> i) & 1ull)) continue;
    for (int j = 0; j < norb; ++j) {
      int32_t src_addr;
      int sign;
      if (j == i) {
        src_addr = static_cast<int32_t>(t);
        sign = 1;
      } else if (!((s >> j) & 1ull)) {
        const uint64_t src = (s & ~(1ull << i)) | (1ull << j);
        src_addr = string_rank(src, z, norb);
        sign = (bits_between(src, i, j) & 1) ? -1 : 1;
      } else {
        continue;
      }
      out[3 * e + 0] = src_addr;
      out[3 * e + 1] = i * norb + j;
      out[3 * e + 2] = sign;
      ++e;
    }
  }
}

// One workgroup per ordered pair (i, j): stream-compacts the sources in
// address order into the map rows (source, target, sign).
__global__ void __launch_bounds__(kBlock)
k_maps(int norb, int nele, int64_t nstrings, const uint64_t* strings,
       const int64_t* offsets, int32_t* maps) {
  __shared__ int32_t z[kMaxNorb * kMaxNorb];
  __shared__ int wave_count[kBlock / 64];
  fill_z_matrix(z, norb, nele);
  __syncthreads();
  const int ij = blockIdx.x;
  const int i = ij / norb;
  const int j = ij % norb;
  const int64_t base = offsets[ij];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int64_t written = 0;
  for (int64_t c0 = 0; c0 < nstrings; c0 += kBlock) {
    const int64_t a = c0 + threadIdx.x;
    bool pred = false;
    uint64_t s = 0, t = 0;
    int sign = 1;
    if (a < nstrings) {
      s = strings[a];
      if (((s >> j) & 1ull) && !((s >> i) & 1ull)) {
        pred = true;
        t = (s | (1ull << i)) & ~(1ull << j);
        sign = (bits_between(s, i, j) & 1) ? -1 : 1;
      } else if (i == j && ((s >> i) & 1ull)) {
        pred = true;
        t = s;
      }
    }
    const uint64_t ballot = __ballot(pred);
    const int before = __popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) wave_count[wave] = __popcll(ballot);
    __syncthreads();
    int woff = 0, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wave) woff += wave_count[w];
      total += wave_count[w];
    }
    if (pred) {
      const int64_t row = base + written + woff + before;
      maps[3 * row + 0] = static_cast<int32_t>(a);
      maps[3 * row + 1] = (t == s) ? static_cast<int32_t>(a)
                                   : string_rank(t, z, norb);
      maps[3 * row + 2] = sign;
    }
    written += total;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock)
k_fill_i32(int64_t n, int32_t value, int32_t* out) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       k < n; k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[k] = value;
}

// Pair table [npair][nstrings] from dexc; one thread per dexc row.
__global__ void __launch_bounds__(kBlock)
k_pair_table(int norb, int nexc, int64_t nstrings, const int32_t* dexc,
             int compressed, int32_t* ptab) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nstrings * nexc) return;
  const int64_t t = r / nexc;
  const int32_t src = dexc[3 * r + 0];
  const int32_t ij = dexc[3 * r + 1];
  const int32_t sign = dexc[3 * r + 2];
  const int i = ij / norb;
  const int j = ij % norb;
  const int p = compressed ? comp_pair(i, j) : ij;
  ptab[static_cast<int64_t>(p) * nstrings + t] = sign * (src + 1);
}

int check_sector(int norb, int nele, int64_t nstrings) {
  FQEHIP_CHECK_ARG(norb >= 0 && norb <= kMaxNorb,
                   "norb=%d outside [0, %d]", norb, kMaxNorb);
  FQEHIP_CHECK_ARG(nele >= 0 && nele <= norb, "nele=%d outside [0, norb=%d]",
                   nele, norb);
  // C(norb, nele) on the host, exact for norb <= 32
  int64_t c = 1;
  for (int k = 1; k <= nele; ++k) c = c * (norb - nele + k) / k;
  FQEHIP_CHECK_ARG(nstrings == c, "nstrings=%lld != C(%d,%d)=%lld",
                   (long long)nstrings, norb, nele, (long long)c);
  FQEHIP_CHECK_ARG(c * (nele * (norb - nele + 1)) < (int64_t(1) << 31),
                   "tables of C(%d,%d) strings exceed int32 addressing", norb,
                   nele);
  return FQEHIP_OK;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_string_tables(int norb, int nele, int64_t nstrings,
                                    uint64_t* d_strings, int32_t* d_dexc,
                                    void* stream) {
  clear_error();
  if (int rc = check_sector(norb, nele, nstrings)) return rc;
  FQEHIP_CHECK_ARG(d_strings != nullptr, "d_strings is NULL");
  hipStream_t st = as_stream(stream);
  k_strings<<<blocks_for(nstrings, kBlock), kBlock, 0, st>>>(norb, nele, nstrings,
                                                           d_strings);
  FQEHIP_CHECK_LAUNCH("k_strings");
  if (d_dexc != nullptr && nele > 0) {
    k_dexc<<<blocks_for(nstrings, kBlock), kBlock, 0, st>>>(norb, nele, nstrings,
                                                          d_strings, d_dexc);
    FQEHIP_CHECK_LAUNCH("k_dexc");
  }
  return FQEHIP_OK;
}

extern "C" int fqehip_string_maps(int norb, int nele, int64_t nstrings,
                                  const uint64_t* d_strings,
                                  const int64_t* d_offsets, int32_t* d_maps,
                                  void* stream) {
  clear_error();
  if (int rc = check_sector(norb, nele, nstrings)) return rc;
  FQEHIP_CHECK_ARG(d_strings && d_offsets && d_maps, "NULL table pointer");
  if (norb == 0 || nele == 0) return FQEHIP_OK;
  k_maps<<<norb * norb, kBlock, 0, as_stream(stream)>>>(
      norb, nele, nstrings, d_strings, d_offsets, d_maps);
  FQEHIP_CHECK_LAUNCH("k_maps");
  return FQEHIP_OK;
}

extern "C" int fqehip_pair_table(int norb, int nele, int64_t nstrings,
                                 const int32_t* d_dexc, int compressed,
                                 int32_t* d_ptab, void* stream) {
  clear_error();
  if (int rc = check_sector(norb, nele, nstrings)) return rc;
  FQEHIP_CHECK_ARG(d_ptab != nullptr, "d_ptab is NULL");
  FQEHIP_CHECK_ARG(compressed == 0 || compressed == 1, "compressed=%d",
                   compressed);
  const int npair = compressed ? norb * (norb + 1) / 2 : norb * norb;
  const int64_t total = static_cast<int64_t>(npair) * nstrings;
  hipStream_t st = as_stream(stream);
  if (total > 0) {
    const unsigned nb = blocks_for(total, kBlock) < 4096u
                            ? blocks_for(total, kBlock) : 4096u;
    k_fill_i32<<<nb, kBlock, 0, st>>>(total, 0, d_ptab);
    FQEHIP_CHECK_LAUNCH("k_fill_i32");
  }
  const int nexc = nele * (norb - nele + 1);
  const int64_t rows = nstrings * nexc;
  if (rows > 0) {
    FQEHIP_CHECK_ARG(d_dexc != nullptr, "d_dexc is NULL");
    k_pair_table<<<blocks_for(rows, kBlock), kBlock, 0, st>>>(
        norb, nexc, nstrings, d_dexc, compressed, d_ptab);
    FQEHIP_CHECK_LAUNCH("k_pair_table");
  }
  return FQEHIP_OK;
}
