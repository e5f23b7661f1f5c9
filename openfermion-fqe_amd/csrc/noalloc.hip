// Apply entry point that allocates nothing: the workspace of fqehip_apply
// comes from a library-owned device array instead of the caller.
//
// The array `scratch` is a file-scope device variable, so every GPU the
// library is loaded on holds its own copy (FQEHIP_NOALLOC_SCRATCH_BYTES)
// until the process ends.  Host code never uses the name `scratch` as an
// address: on the first call on a device, with that device current, the
// device address of that GPU's copy is looked up once with
// hipGetSymbolAddress and kept per device; kernels receive that pointer.
//
// Every call on one GPU shares that GPU's array.  The kernels write every
// workspace element they read (slot rows, row accumulators, dvec and Y
// planes), so nothing needs clearing between calls; calls that overlap on
// one GPU (two streams, two host threads) race on the array.
#include <mutex>

#include "fqehip_common.h"

namespace {

static __device__ __align__(256) unsigned char
    scratch[FQEHIP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void* g_scratch[kMaxDevices];
hipError_t g_status[kMaxDevices];

// Workspace bytes of the path fqehip_apply takes for this sector and
// operator (the one-body apply needs none).
int64_t needed_bytes(int precision, const fqehip_sector* s,
                     const fqehip_operator* op) {
  if (op->h1) return 0;
  const int64_t n = s->lena * s->lenb;
  const int64_t elem = precision == FQEHIP_C128 ? 8 : 4;
  if (fqehip_fused_supported(precision, s, op)) {
    int64_t b = static_cast<int64_t>(s->nslot + 1) * 2 * n * elem;
    if (precision == FQEHIP_C128 && s->hbb_f64 && s->tile == FQEHIP_TILE_OZAKI) {
      b += s->lenb * 128 * 4 + 128 * (5 * 128 + 12);  // beta table, pair digits
      b += s->lena * 8;  // row maxima of |C|
      if (s->hbb_oz)     // C digit planes and row scales
        b += 6 * 2 * s->lena * ((s->lenb + 63) / 64 * 64) + 8 * s->lena;
    }
    return b;
  }
  return static_cast<int64_t>(2 * s->npair + 1) * 2 * n * elem;
}

}  // namespace

extern "C" int fqehip_apply_noalloc(int precision, const fqehip_sector* s,
                                    const fqehip_operator* op,
                                    const void* d_coeff, double alpha_re,
                                    double alpha_im, double beta_re,
                                    double beta_im, const void* d_cin,
                                    void* d_out, double* d_acc,
                                    double* d_norm_partials,
                                    const int32_t* d_done, void* stream) {
  fqehip::clear_error();
  FQEHIP_CHECK_ARG(s && op, "NULL sector or operator");
  FQEHIP_CHECK_ARG(precision == FQEHIP_C128 || precision == FQEHIP_C64,
                   "precision=%d", precision);
  FQEHIP_CHECK_ARG(s->lena >= 0 && s->lenb >= 0 && s->npair >= 0 &&
                       s->nslot >= 0, "negative sector size");
  const int64_t need = needed_bytes(precision, s, op);
  if (need > static_cast<int64_t>(FQEHIP_NOALLOC_SCRATCH_BYTES)) {
    fqehip::set_error(
        "fqehip_apply_noalloc: the sector needs %lld workspace bytes, above "
        "the library's fixed %lld (use fqehip_apply with a caller workspace)",
        static_cast<long long>(need),
        static_cast<long long>(FQEHIP_NOALLOC_SCRATCH_BYTES));
    return FQEHIP_ERR_UNSUPPORTED;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fqehip::set_error("fqehip_apply_noalloc: no usable current device");
    return FQEHIP_ERR_HIP;
  }
  std::call_once(g_once[dev], [dev] {
    void* p = nullptr;
    g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch[dev] = p;
  });
  if (g_status[dev] != hipSuccess || g_scratch[dev] == nullptr) {
    fqehip::set_error("fqehip_apply_noalloc: hipGetSymbolAddress: %s",
                      hipGetErrorString(g_status[dev]));
    return FQEHIP_ERR_HIP;
  }
  return fqehip_apply(precision, s, op, d_coeff, alpha_re, alpha_im, beta_re,
                      beta_im, d_cin, d_out, d_acc, d_norm_partials, d_done,
                      g_scratch[dev], stream);
}
