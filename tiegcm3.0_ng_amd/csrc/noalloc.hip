// Allocation-free form of tgx_trsolv_dev_f64 for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, hosts that
// forbid hipMalloc).  The Thomas ratios live in a file-scope device array of
// fixed size instead of a library-allocated buffer; see include/tgx.h.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tgx.h"
#include "tgx_common.h"

extern "C" int tgx_launch_thomas_f64(const double *, const double *, const double *, const double *,
                                     double *, double *, int, int, int, int, int, hipStream_t);
extern "C" int tgx_launch_pcr(const double *, const double *, const double *, const double *,
                              double *, int, int, int, int, int, int, int, hipStream_t);

// One copy per GPU, held until the process ends.  Only device code may name
// it: host code uses the address hipGetSymbolAddress returns for the current
// device (g_scratch_dev below).
static __device__ __align__(256) double scratch[TGX_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_ptr(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tgx_set_error(TGX_ERR_NODEV, "no HIP device: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return tgx_set_error(TGX_ERR_NODEV, "tgx_trsolv_dev_f64_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return tgx_set_error(TGX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

extern "C" int tgx_trsolv_dev_f64_noalloc(const double *a, const double *b, const double *c,
                                          const double *f, double *x, int ldk, int k1, int k2,
                                          int col0, int ncol, int scheme, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (ldk <= 0 || k1 < 0 || k2 >= ldk || col0 < 0 || ncol < 0)
    return tgx_set_error(TGX_ERR_ARG,
                         "tgx_trsolv_dev_f64_noalloc: bad shape ldk=%d k1=%d k2=%d col0=%d ncol=%d", ldk,
                         k1, k2, col0, ncol);
  if (scheme != TGX_SCHEME_AUTO && scheme != TGX_SCHEME_THOMAS && scheme != TGX_SCHEME_PCR)
    return tgx_set_error(TGX_ERR_ARG, "tgx_trsolv_dev_f64_noalloc: unknown scheme %d", scheme);
  if (ncol == 0 || k2 < k1) return 0;
  const int nrow = k2 - k1 + 1;
  if (scheme == TGX_SCHEME_AUTO) {  // the autotuner's cached choice; never timed here
    scheme = tgx_autotune_choice(ldk, k1, k2, ncol);
    if (scheme == TGX_SCHEME_AUTO) scheme = TGX_SCHEME_THOMAS;
  }
  if (scheme == TGX_SCHEME_PCR && nrow > 256) scheme = TGX_SCHEME_THOMAS;
  if (scheme == TGX_SCHEME_PCR) return tgx_launch_pcr(a, b, c, f, x, ldk, k1, k2, col0, ncol, 0, 0, s);
  const unsigned long long need = (unsigned long long)ldk * (unsigned long long)ncol;
  if (need > (unsigned long long)TGX_NOALLOC_SCRATCH_DOUBLES)
    return tgx_set_error(TGX_ERR_ARG,
                         "tgx_trsolv_dev_f64_noalloc: ldk*ncol = %llu doubles of Thomas ratios exceed the "
                         "fixed scratch of %llu doubles (%d MiB); split the batch or use tgx_trsolv_dev_f64",
                         need, (unsigned long long)TGX_NOALLOC_SCRATCH_DOUBLES,
                         (int)(TGX_NOALLOC_SCRATCH_DOUBLES * sizeof(double) >> 20));
  double *w2 = nullptr;
  if (int rc = scratch_ptr(&w2)) return rc;
  return tgx_launch_thomas_f64(a, b, c, f, x, w2, ldk, k1, k2, col0, ncol, s);
}
