// C-ABI of libtgx: drop-in host entry points with the reference's argument
// layout, stream-ordered device entry points, error reporting and the solver
// autotuner.  See include/tgx.h for the contract of every symbol.
#include <hip/hip_runtime.h>

#include <sys/syscall.h>
#include <unistd.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/tgx.h"
#include "tgx_common.h"

// ---- error slot ------------------------------------------------------------
static thread_local char g_err[512] = "";
static thread_local int g_status = 0;

int tgx_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  g_status = code;
  return code;
}
extern "C" const char *tgx_last_error(void) { return g_err; }
extern "C" int tgx_last_status(void) { return g_status; }
extern "C" void tgx_clear_error(void) {
  g_err[0] = 0;
  g_status = 0;
}
extern "C" int tgx_version(void) { return TGX_VERSION; }

bool tgx_on_main_thread() { return (pid_t)syscall(SYS_gettid) == getpid(); }

// ---- launchers (solvers.hip) -------------------------------------------------
extern "C" int tgx_launch_thomas_f64(const double *, const double *, const double *, const double *,
                                     double *, double *, int, int, int, int, int, hipStream_t);
extern "C" int tgx_launch_pcr(const double *, const double *, const double *, const double *,
                              double *, int, int, int, int, int, int, int, hipStream_t);
extern "C" int tgx_launch_blktri2_f64(const double *, const double *, const double *, const double *,
                                      double *, double *, int, int, int, int, hipStream_t);
extern "C" int tgx_launch_ringfilter(double *, int, long, int, const long *, const int *, int, int, int,
                                     double, hipStream_t);
extern "C" int tgx_launch_blktri3_f64(const double *, const double *, const double *, const double *,
                                      double *, double *, int, int, int, int, hipStream_t);

// ---- per-thread workspace ------------------------------------------------------
// Scratch for the device entry points (Thomas ratios, block gammas) is kept
// per calling host thread, device and stream: work on one stream runs in
// order, so a buffer reused by the next call on the same stream is never
// read by a kernel still in flight, and calls from other threads or on other
// streams get buffers of their own.  Staging buffers and the copy stream of
// the host drop-ins (which return only when their copies are done) are kept
// per host thread and device.  Buffers are grown on demand.  A worker
// thread's buffers and stream are released when the thread exits, after the
// device has finished the work queued on them; the main thread's are held
// until the process ends.
namespace {
struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
};
struct Stage {
  Buf buf;
  hipStream_t stream = nullptr;
};
std::mutex g_mu;  // guards the autotuner's table

// The scratch buffers are keyed by the caller's stream, which the caller may
// already have destroyed when the thread exits: wait for the whole device.
struct ScratchSet {
  std::map<std::pair<int, hipStream_t>, Buf> m;
  ~ScratchSet() {
    if (tgx_on_main_thread()) return;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    for (auto &kv : m) {
      if (!kv.second.p) continue;
      if (hipSetDevice(kv.first.first) != hipSuccess) continue;
      (void)hipDeviceSynchronize();
      (void)hipFree(kv.second.p);
    }
    (void)hipSetDevice(cur);
  }
};
struct StageSet {
  std::map<int, Stage> m;
  ~StageSet() {
    if (tgx_on_main_thread()) return;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    for (auto &kv : m) {
      if (hipSetDevice(kv.first) != hipSuccess) continue;
      if (kv.second.stream) (void)hipStreamSynchronize(kv.second.stream);
      if (kv.second.buf.p) (void)hipFree(kv.second.buf.p);
      if (kv.second.stream) (void)hipStreamDestroy(kv.second.stream);
    }
    (void)hipSetDevice(cur);
  }
};
thread_local ScratchSet t_scratch;
thread_local StageSet t_stage;

int current_device(int *dev) {
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return tgx_set_error(3, "no HIP device: %s", hipGetErrorString(e));
  return 0;
}

// Grows b to at least need bytes.  The old buffer may still be read by work
// queued on stream s, so s is drained before it is freed.
int grow(Buf *b, size_t need, hipStream_t s) {
  if (b->bytes >= need) return 0;
  if (b->p) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return tgx_set_error(2, "hipStreamSynchronize: %s", hipGetErrorString(e));
    e = hipFree(b->p);
    if (e != hipSuccess) return tgx_set_error(2, "hipFree: %s", hipGetErrorString(e));
    b->p = nullptr;
    b->bytes = 0;
  }
  size_t sz = need + need / 4 + 4096;
  hipError_t e = hipMalloc(&b->p, sz);
  if (e != hipSuccess) return tgx_set_error(2, "hipMalloc(%zu): %s", sz, hipGetErrorString(e));
  b->bytes = sz;
  return 0;
}

int get_scratch(size_t bytes, hipStream_t s, void **out) {
  int dev;
  if (int rc = current_device(&dev)) return rc;
  Buf &b = t_scratch.m[{dev, s}];
  if (int rc = grow(&b, bytes, s)) return rc;
  *out = b.p;
  return 0;
}

int get_stage(size_t bytes, void **out, hipStream_t *s) {
  int dev;
  if (int rc = current_device(&dev)) return rc;
  Stage &d = t_stage.m[dev];
  if (!d.stream) {
    hipError_t e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return tgx_set_error(2, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  if (int rc = grow(&d.buf, bytes, d.stream)) return rc;
  *out = d.buf.p;
  *s = d.stream;
  return 0;
}

// ---- autotuner ------------------------------------------------------------------
// Picks Thomas or PCR for a scalar batch shape the first time it is seen on
// a device, by timing both on the caller's data with hipEvents; the choice
// is cached per (device, ldk, rows, column bucket).
std::map<std::tuple<int, int, int, int>, int> g_tune;

int col_bucket(int ncol) {
  int b = 0;
  while ((1 << b) < ncol) ++b;
  return b;
}

int run_scalar(int scheme, const double *a, const double *b, const double *c, const double *f,
               double *x, int ldk, int k1, int k2, int col0, int ncol, hipStream_t s) {
  if (scheme == TGX_SCHEME_PCR) return tgx_launch_pcr(a, b, c, f, x, ldk, k1, k2, col0, ncol, 0, 0, s);
  void *w2 = nullptr;
  size_t need = sizeof(double) * (size_t)ldk * (size_t)ncol;
  if (int rc = get_scratch(need, s, &w2)) return rc;
  return tgx_launch_thomas_f64(a, b, c, f, x, (double *)w2, ldk, k1, k2, col0, ncol, s);
}

// hipEvent pair destroyed on every exit path.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

int time_scheme(int scheme, const double *a, const double *b, const double *c, const double *f,
                double *x, int ldk, int k1, int k2, int col0, int ncol, hipStream_t s, float *ms) {
  EventPair ev;
  TGX_HIP(hipEventCreate(&ev.e0));
  TGX_HIP(hipEventCreate(&ev.e1));
  if (int rc = run_scalar(scheme, a, b, c, f, x, ldk, k1, k2, col0, ncol, s)) return rc;  // warm
  TGX_HIP(hipEventRecord(ev.e0, s));
  for (int i = 0; i < 3; ++i)
    if (int rc = run_scalar(scheme, a, b, c, f, x, ldk, k1, k2, col0, ncol, s)) return rc;
  TGX_HIP(hipEventRecord(ev.e1, s));
  TGX_HIP(hipEventSynchronize(ev.e1));
  TGX_HIP(hipEventElapsedTime(ms, ev.e0, ev.e1));
  return 0;
}
}  // namespace

extern "C" int tgx_autotune_choice(int ldk, int k1, int k2, int ncol) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TGX_SCHEME_AUTO;  // no device: nothing tuned
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tune.find({dev, ldk, k2 - k1 + 1, col_bucket(ncol)});
  return it == g_tune.end() ? TGX_SCHEME_AUTO : it->second;
}

// ---- device entry points ----------------------------------------------------
extern "C" int tgx_trsolv_dev_f64(const double *a, const double *b, const double *c,
                                  const double *f, double *x, int ldk, int k1, int k2, int col0,
                                  int ncol, int scheme, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (ldk <= 0 || k1 < 0 || k2 >= ldk || col0 < 0 || ncol < 0)
    return tgx_set_error(1, "tgx_trsolv_dev_f64: bad shape ldk=%d k1=%d k2=%d col0=%d ncol=%d",
                         ldk, k1, k2, col0, ncol);
  if (ncol == 0 || k2 < k1) return 0;
  if (scheme == TGX_SCHEME_PCR && k2 - k1 + 1 > 256) scheme = TGX_SCHEME_THOMAS;
  if (scheme == TGX_SCHEME_AUTO) {
    int dev = 0;
    if (int rc = current_device(&dev)) return rc;
    auto key = std::make_tuple(dev, ldk, k2 - k1 + 1, col_bucket(ncol));
    int choice = TGX_SCHEME_AUTO;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      auto it = g_tune.find(key);
      if (it != g_tune.end()) choice = it->second;
    }
    if (choice == TGX_SCHEME_AUTO) {
      float tt = 0, tp = 1e30f;
      if (int rc = time_scheme(TGX_SCHEME_THOMAS, a, b, c, f, x, ldk, k1, k2, col0, ncol, s, &tt)) return rc;
      if (k2 - k1 + 1 <= 256)
        if (int rc = time_scheme(TGX_SCHEME_PCR, a, b, c, f, x, ldk, k1, k2, col0, ncol, s, &tp)) return rc;
      choice = tp < tt ? TGX_SCHEME_PCR : TGX_SCHEME_THOMAS;
      std::lock_guard<std::mutex> lk(g_mu);
      g_tune[key] = choice;
    }
    scheme = choice;
  }
  return run_scalar(scheme, a, b, c, f, x, ldk, k1, k2, col0, ncol, s);
}

extern "C" int tgx_trsolv_dev_mixed(const double *a, const double *b, const double *c,
                                    const double *f, double *x, int ldk, int k1, int k2,
                                    int col0, int ncol, int nsweep, void *stream) {
  if (ldk <= 0 || k1 < 0 || k2 >= ldk || col0 < 0 || ncol < 0 || nsweep < 0 || nsweep > 8)
    return tgx_set_error(1, "tgx_trsolv_dev_mixed: bad arguments");
  if (k2 - k1 + 1 > 256) return tgx_set_error(1, "tgx_trsolv_dev_mixed: at most 256 rows");
  return tgx_launch_pcr(a, b, c, f, x, ldk, k1, k2, col0, ncol, nsweep, 1, (hipStream_t)stream);
}

extern "C" int tgx_blktri2_dev_f64(const double *a, const double *b, const double *c,
                                   const double *f, double *x, int ldk, int nrow, int col0,
                                   int ncol, void *stream) {
  if (ldk <= 0 || nrow < 1 || nrow > ldk || col0 < 0 || ncol < 0)
    return tgx_set_error(1, "tgx_blktri2_dev_f64: bad shape");
  void *gs = nullptr;
  if (int rc = get_scratch(sizeof(double) * 4 * (size_t)ldk * (size_t)(col0 + ncol), (hipStream_t)stream, &gs)) return rc;
  return tgx_launch_blktri2_f64(a, b, c, f, x, (double *)gs, ldk, nrow, col0, ncol, (hipStream_t)stream);
}

extern "C" int tgx_blktri3_dev_f64(const double *p, const double *q, const double *r,
                                   const double *f, double *x, int ldk, int nrow, int col0,
                                   int ncol, void *stream) {
  if (ldk <= 0 || nrow < 1 || nrow > ldk || col0 < 0 || ncol < 0)
    return tgx_set_error(1, "tgx_blktri3_dev_f64: bad shape");
  void *gs = nullptr;
  if (int rc = get_scratch(sizeof(double) * 9 * (size_t)ldk * (size_t)(col0 + ncol), (hipStream_t)stream, &gs)) return rc;
  return tgx_launch_blktri3_f64(p, q, r, f, x, (double *)gs, ldk, nrow, col0, ncol, (hipStream_t)stream);
}

// ---- host drop-ins ------------------------------------------------------------
// Host arrays in, host array out, synchronous.  x is staged in as well so the
// entries the reference leaves untouched keep their values.
namespace {
int host_scalar(const double *a, const double *b, const double *c, const double *f, double *x,
                int ldk, int ncols_total, int k1, int k2, int col0, int ncol) {
  size_t n = (size_t)ldk * (size_t)ncols_total, bytes = sizeof(double) * n;
  void *st = nullptr;
  hipStream_t s;
  if (int rc = get_stage(5 * bytes, &st, &s)) return rc;
  double *d = (double *)st;
  double *da = d, *db = d + n, *dc = d + 2 * n, *df = d + 3 * n, *dx = d + 4 * n;
  TGX_HIP(hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s));
  TGX_HIP(hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s));
  TGX_HIP(hipMemcpyAsync(dc, c, bytes, hipMemcpyHostToDevice, s));
  TGX_HIP(hipMemcpyAsync(df, f, bytes, hipMemcpyHostToDevice, s));
  TGX_HIP(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
  if (int rc = tgx_trsolv_dev_f64(da, db, dc, df, dx, ldk, k1, k2, col0, ncol, TGX_SCHEME_THOMAS, s))
    return rc;
  TGX_HIP(hipMemcpyAsync(x, dx, bytes, hipMemcpyDeviceToHost, s));
  TGX_HIP(hipStreamSynchronize(s));
  return 0;
}
}  // namespace

extern "C" void tgx_trsolv_f64(const double *a, const double *b, const double *c, const double *f,
                               double *x, int lev0, int lev1, int k1, int k2, int lon0, int lon1,
                               int lonmax, int lat, int idebug) {
  (void)lat;
  (void)idebug;  // debug addfld dumps (trsolv.F:57-62) are not reproduced
  tgx_clear_error();
  int nlev = lev1 - lev0 + 1, ncols = lon1 - lon0 + 1;
  if (nlev <= 0 || ncols <= 0 || k1 < lev0 || k2 > lev1) {
    tgx_set_error(1, "tgx_trsolv_f64: bad bounds lev0=%d lev1=%d k1=%d k2=%d lon0=%d lon1=%d",
                  lev0, lev1, k1, k2, lon0, lon1);
    return;
  }
  int lonbeg = lon0 == 1 ? 3 : lon0;          // trsolv.F:32-35
  int lonend = lon1 == lonmax ? lon1 - 2 : lon1;
  if (lonend < lonbeg || k2 < k1) return;
  host_scalar(a, b, c, f, x, nlev, ncols, k1 - lev0, k2 - lev0, lonbeg - lon0, lonend - lonbeg + 1);
}

extern "C" void tgx_trsolv_ng_f64(const double *a, const double *b, const double *c,
                                  const double *f, double *x, int nk, int nlond, int nlatd, int k2) {
  tgx_clear_error();
  if (nk <= 0 || nlond <= 0 || nlatd <= 0 || k2 < 1 || k2 > nk) {
    tgx_set_error(1, "tgx_trsolv_ng_f64: bad shape nk=%d nlond=%d nlatd=%d k2=%d", nk, nlond, nlatd, k2);
    return;
  }
  host_scalar(a, b, c, f, x, nk, nlond * nlatd, 0, k2 - 1, 0, nlond * nlatd);
}

extern "C" void tgx_blktri_f64(const double *a, const double *b, const double *c, const double *f,
                               double *beta, double *gamma, double *y, double *x, int lev0,
                               int lev1, int lon0, int lon1, int lonmax) {
  (void)beta;
  (void)gamma;
  (void)y;  // work arrays of the reference; its callers discard them (duv.F:375)
  tgx_clear_error();
  int nlev = lev1 - lev0 + 1, ncols = lon1 - lon0 + 1;
  if (nlev < 2 || ncols <= 0) {
    tgx_set_error(1, "tgx_blktri_f64: bad bounds");
    return;
  }
  int lonbeg = lon0 == 1 ? 3 : lon0;          // duv.F:825-828
  int lonend = lon1 == lonmax ? lonmax - 2 : lon1;
  if (lonend < lonbeg) return;
  size_t nm = 4 * (size_t)nlev * ncols, nv = 2 * (size_t)nlev * ncols;
  void *st = nullptr;
  hipStream_t s;
  if (get_stage(sizeof(double) * (3 * nm + 2 * nv), &st, &s)) return;
  double *d = (double *)st;
  double *da = d, *db = d + nm, *dc = d + 2 * nm, *df = d + 3 * nm, *dx = d + 3 * nm + nv;
#define CK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { tgx_set_error(2, "%s: %s", #call, hipGetErrorString(e_)); return; } } while (0)
  CK(hipMemcpyAsync(da, a, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(db, b, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dc, c, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(df, f, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dx, x, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  if (tgx_blktri2_dev_f64(da, db, dc, df, dx, nlev, nlev - 1, lonbeg - lon0, lonend - lonbeg + 1, s)) return;
  CK(hipMemcpyAsync(x, dx, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
#undef CK
}

extern "C" void tgx_blktri3_f64(const double *p, const double *q, const double *r, const double *f,
                                double *x, int nk, int ncol, int k2) {
  tgx_clear_error();
  if (nk <= 0 || ncol <= 0 || k2 < 1 || k2 > nk) {
    tgx_set_error(1, "tgx_blktri3_f64: bad shape");
    return;
  }
  size_t nm = 9 * (size_t)nk * ncol, nv = 3 * (size_t)nk * ncol;
  void *st = nullptr;
  hipStream_t s;
  if (get_stage(sizeof(double) * (3 * nm + 2 * nv), &st, &s)) return;
  double *d = (double *)st;
  double *dp = d, *dq = d + nm, *dr = d + 2 * nm, *df = d + 3 * nm, *dx = d + 3 * nm + nv;
#define CK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { tgx_set_error(2, "%s: %s", #call, hipGetErrorString(e_)); return; } } while (0)
  CK(hipMemcpyAsync(dp, p, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dq, q, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dr, r, sizeof(double) * nm, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(df, f, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dx, x, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  if (tgx_blktri3_dev_f64(dp, dq, dr, df, dx, nk, k2, 0, ncol, s)) return;
  CK(hipMemcpyAsync(x, dx, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
#undef CK
}

extern "C" void tgx_ringfilter_f64(double *f, int lev0, int lev1, int order, const char *name, int lat,
                                   int nlon, int nlat, double dlamda, int nlat_filter, const int *chunk) {
  tgx_clear_error();
  const int nlev = lev1 - lev0 + 1;
  if (!f || nlev <= 0 || nlon < 4 || lat < 1 || lat > nlat || nlat_filter < 0 || (nlat_filter > 0 && !chunk)) {
    tgx_set_error(1, "tgx_ringfilter_f64: bad arguments (lev0 %d lev1 %d nlon %d lat %d nlat %d)", lev0, lev1,
                  nlon, lat, nlat);
    return;
  }
  int nchnk = 0;  // ringfilter.F:23-34
  if (lat <= nlat_filter) nchnk = chunk[lat - 1];
  else if (lat >= nlat - nlat_filter + 1) nchnk = chunk[nlat - lat];
  if (nchnk == 0) return;
  if (order < 0 || order > TGX_RING_MAX_ORDER || order >= nlon || nchnk > nlon) {  // ringfilter.F:36
    tgx_set_error(1, "tgx_ringfilter_f64: order %d (0..%d, < nlon %d) or %d chunks", order, TGX_RING_MAX_ORDER, nlon,
                  nchnk);
    return;
  }
  const int oplus = name && strcmp(name, "OPLUS") == 0;
  const size_t bytes = sizeof(double) * (size_t)(nlon + 4) * nlev;
  void *st = nullptr;
  hipStream_t s;
  if (get_stage(bytes, &st, &s)) return;
#define CK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { tgx_set_error(2, "%s: %s", #call, hipGetErrorString(e_)); return; } } while (0)
  CK(hipMemcpyAsync(st, f, bytes, hipMemcpyHostToDevice, s));
  const long off = 0;
  if (tgx_launch_ringfilter((double *)st, nlev, nlon + 4, 1, &off, &nchnk, nlon, order, oplus, dlamda, s)) return;
  CK(hipMemcpyAsync(f, st, bytes, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
#undef CK
}

// lbc.F:527-594 (bndcmp): b = (e/dz - f/2)^-1 (e/dz + f/2), fb = (e/dz - f/2)^-1 g,
// b returned row-major b[3*(l-1)+(m-1)] = b(l,m).
extern "C" void tgx_bndcmp(double dz, double *b, double *fb) {
  const double alfa = 0.22 + 0.014, pshelb = 0.1154E-5;
  double e[3][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 0}};
  double fm[3][3] = {{1, 1, 0}, {0, -1, 0}, {0, 0, 1}};
  double w1[3][3], w2[3][3];
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) {
      w1[l][m] = e[l][m] / dz - fm[l][m] / 2.;
      w2[l][m] = e[l][m] / dz + fm[l][m] / 2.;
    }
  double d = w1[0][0] * w1[1][1] * w1[2][2] - w1[0][0] * w1[1][2] * w1[2][1] -
             w1[0][1] * w1[1][0] * w1[2][2] + w1[0][1] * w1[1][2] * w1[2][0] +
             w1[0][2] * w1[1][0] * w1[2][1] - w1[0][2] * w1[1][1] * w1[2][0];
  double adj[3][3] = {
      {(w1[1][1] * w1[2][2] - w1[1][2] * w1[2][1]), -(w1[0][1] * w1[2][2] - w1[0][2] * w1[2][1]),
       (w1[0][1] * w1[1][2] - w1[0][2] * w1[1][1])},
      {-(w1[1][0] * w1[2][2] - w1[1][2] * w1[2][0]), (w1[0][0] * w1[2][2] - w1[0][2] * w1[2][0]),
       -(w1[0][0] * w1[1][2] - w1[0][2] * w1[1][0])},
      {(w1[1][0] * w1[2][1] - w1[1][1] * w1[2][0]), -(w1[0][0] * w1[2][1] - w1[0][1] * w1[2][0]),
       (w1[0][0] * w1[1][1] - w1[0][1] * w1[1][0])}};
  double w3[3][3];
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) w3[l][m] = adj[l][m] / d;
  double g[3] = {-alfa, -0., -pshelb};
  for (int l = 0; l < 3; ++l) {
    for (int m = 0; m < 3; ++m) b[3 * l + m] = w3[l][0] * w2[0][m] + w3[l][1] * w2[1][m] + w3[l][2] * w2[2][m];
    fb[l] = w3[l][0] * g[0] + w3[l][1] * g[1] + w3[l][2] * g[2];
  }
}
