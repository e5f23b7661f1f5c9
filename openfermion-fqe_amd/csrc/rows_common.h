// Device helpers shared by the alpha-row kernels of the fused apply
// (fused.hip) and of the Ozaki-digit apply (ozaki.hip): raw buffer loads
// and the per-alpha-row LDS tables (gather offsets and signs, slot-row
// bases and signs of every pair).
#pragma once

#include "fqehip_common.h"

namespace fqehip {
namespace rows {

__device__ __forceinline__ int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

// Raw buffer loads.  A buffer resource carries the byte size of the array;
// the hardware returns 0 for an offset outside it, so absent table entries
// are loaded with offset -1 instead of a branch around the load (all of a
// thread's loads can then be in flight together).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p,
                                                            int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0,
                                           static_cast<int>(bytes), 0x00020000);
}

__device__ __forceinline__ int32_t bload_i32(__amdgpu_buffer_rsrc_t r, int off) {
  return static_cast<int32_t>(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

__device__ __forceinline__ float2 bload_f2(__amdgpu_buffer_rsrc_t r, int off) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}

__device__ __forceinline__ double2 bload_d2(__amdgpu_buffer_rsrc_t r, int off) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  const uint64_t lo = static_cast<uint64_t>(v[0]) | (static_cast<uint64_t>(v[1]) << 32);
  const uint64_t hi = static_cast<uint64_t>(v[2]) | (static_cast<uint64_t>(v[3]) << 32);
  return make_double2(__longlong_as_double(static_cast<long long>(lo)),
                      __longlong_as_double(static_cast<long long>(hi)));
}

// Streaming (non-temporal) stores and loads of one complex value: the slot
// rows are written once by the row kernels and read once by the reduction,
// and should not evict the gathered C rows from L2 and the last-level cache.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stream_store2(double* p, double x, double y) {
  __builtin_nontemporal_store(d2v{x, y}, reinterpret_cast<d2v*>(p));
}
__device__ __forceinline__ void stream_store2(float* p, float x, float y) {
  __builtin_nontemporal_store(f2v{x, y}, reinterpret_cast<f2v*>(p));
}

// Slot rows of the Ozaki-digit engine in 48 bits per double: sign,
// exponent and the top 36 mantissa bits, rounded to nearest (ties away
// from zero) at bit 16, so the relative rounding error is at most 2^-37.
// One complex value is three words {hi(re), hi(im), lo(re) | lo(im) << 16}
// (hi: bits 63..32, lo: bits 31..16), 12 bytes at a 4-byte alignment.
typedef uint32_t u3v __attribute__((ext_vector_type(3)));
typedef u3v u3v_a4 __attribute__((aligned(4)));
__device__ __forceinline__ void stream_store_p48(uint32_t* p, double x,
                                                 double y) {
  const uint64_t bx = static_cast<uint64_t>(__double_as_longlong(x)) + 0x8000u;
  const uint64_t by = static_cast<uint64_t>(__double_as_longlong(y)) + 0x8000u;
  const u3v v{static_cast<uint32_t>(bx >> 32), static_cast<uint32_t>(by >> 32),
              static_cast<uint32_t>((bx >> 16) & 0xffffu) |
                  (static_cast<uint32_t>(by >> 16) << 16)};
  __builtin_nontemporal_store(v, reinterpret_cast<u3v_a4*>(p));
}
__device__ __forceinline__ u3v stream_load_p48(const uint32_t* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const u3v_a4*>(p));
}
__device__ __forceinline__ double p48_re(u3v v) {
  return __longlong_as_double(static_cast<long long>(
      (static_cast<uint64_t>(v[0]) << 32) | (static_cast<uint64_t>(v[2] & 0xffffu) << 16)));
}
__device__ __forceinline__ double p48_im(u3v v) {
  return __longlong_as_double(static_cast<long long>(
      (static_cast<uint64_t>(v[1]) << 32) | (static_cast<uint64_t>(v[2] >> 16) << 16)));
}

constexpr int kRowTabs = 128;              // per-row LDS tables (>= npair + 1)

struct RowTables {
  int32_t* aoff;    // [q] offset of C row src_a(q) (elements), -1 if none
  int32_t* asgn;    // [q] sign of the alpha gather entry
  int64_t* sbase;   // [p] slot row base of the alpha target, -1 none, -2 diag
  int32_t* ssgn;    // [p] sign of the alpha target
};

// own_q >= 0 (split kernel): gather row own_q is C[ia] itself, sign +1.
template <typename T>
__device__ __forceinline__ void load_row_tables(
    int64_t ia, int K, int64_t lena, int64_t lenb, int nslot,
    const int32_t* __restrict__ pta, const int32_t* __restrict__ fwa,
    const int32_t* __restrict__ sla, RowTables rt, int tid, int nthreads,
    int own_q = -1) {
  for (int q = tid; q < kRowTabs; q += nthreads) {
    int32_t off = -1, sg = 0;
    int64_t sb = -1;
    int32_t ss = 0;
    if (q == own_q) {
      off = static_cast<int32_t>(ia * lenb * 2);
      sg = 1;
    } else if (q < K) {
      const int32_t a = pta[static_cast<int64_t>(q) * lena + ia];
      if (a) {
        off = static_cast<int32_t>((iabs(a) - 1) * lenb * 2);
        sg = a > 0 ? 1 : -1;
      }
      const int32_t f = fwa[static_cast<int64_t>(q) * lena + ia];
      if (f) {
        const int64_t tgt = iabs(f) - 1;
        ss = f > 0 ? 1 : -1;
        sb = (tgt == ia) ? -2
                         : (tgt * nslot + sla[static_cast<int64_t>(q) * lena + ia]) * lenb * 2;
      }
    }
    rt.aoff[q] = off;
    rt.asgn[q] = sg;
    rt.sbase[q] = sb;
    rt.ssgn[q] = ss;
  }
}

// Live pair rows of an alpha string in the split row kernels (excitations
// to another string plus the folded diagonal row): 4 MFMA tiles of 16.
constexpr int kLiveRows = 64;;

}  // namespace rows
}  // namespace fqehip
