// Coupling between nest levels on the device: the interpolation of
// interp.F90:1-128 used by map_in, set_bndry and map_out
// (levels_inner_ng.F90:131-410) and the time interpolation of the lateral and
// lower-boundary data between two outer steps (interp_fields_ng.F90:17-63).
//
// interp3d of the reference interpolates every source level horizontally
// (along y, then x) and then each target column in z.  A target point only
// needs the two source levels and the 2x2 source columns around it, so one
// thread computes one target point from its brackets with the reference's
// formula ((x1-x)*f0 + (x-x0)*f1)/(x1-x0) in the reference's order: the same
// values bit for bit, except that the log form's logarithms and exponentials
// are the device's (a few ulp).
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "../../include/tgx_step.h"
#include "tgx_common.h"

namespace tgx {
namespace nest {

// interp.F90:105-126 (find): 1-based bracket, 0 below xp(1), nx at or above
// xp(nx), bisection in between
__device__ __forceinline__ int find(const double *xp, int nx, double x0) {
  if (x0 < xp[0]) return 0;
  if (x0 >= xp[nx - 1]) return nx;
  int i0 = 1, i1 = nx, i = 0;
  while (i0 + 1 < i1) {
    i = (i0 + i1) / 2;
    if (xp[i - 1] <= x0) i0 = i;
    else i1 = i;
  }
  return (i0 + i1) / 2;
}

// interp.F90:88-101: the 0-based bracket (i0, i1) of x, the end intervals
// extrapolated
__device__ __forceinline__ void bracket(const double *xp, int nxp, double x, int &i0, int &i1) {
  int a = find(xp, nxp, x);
  int b = a + 1;
  if (a <= 0) {
    a = 1;
    b = 2;
  }
  if (b >= nxp + 1) {
    a = nxp - 1;
    b = nxp;
  }
  i0 = a - 1;
  i1 = b - 1;
}

__device__ __forceinline__ double lin(double x, double x0, double x1, double f0, double f1) {
  return ((x1 - x) * f0 + (x - x0) * f1) / (x1 - x0);
}

struct Src {
  const double *f;  // level planes: f[k*plane + j*ld + i]
  int nzp, nxp, nyp, ld;
  long plane;
  const double *zp, *xp, *yp;
};
struct Dst {
  double *f;
  int nz, nx, ny;
  long dk, di, dj;
  const double *z, *x, *y;
};

// interp2d at (x, y) of source level k, with the brackets already found
__device__ __forceinline__ double at_xy(const Src &S, int k, int ix0, int ix1, int jy0, int jy1, double x, double y) {
  const double *p = S.f + (size_t)k * S.plane;
  const double y0 = S.yp[jy0], y1 = S.yp[jy1];
  const double a0 = lin(y, y0, y1, p[(size_t)jy0 * S.ld + ix0], p[(size_t)jy1 * S.ld + ix0]);
  const double a1 = lin(y, y0, y1, p[(size_t)jy0 * S.ld + ix1], p[(size_t)jy1 * S.ld + ix1]);
  return lin(x, S.xp[ix0], S.xp[ix1], a0, a1);
}

__global__ void __launch_bounds__(256) interp3d_kernel(Src S, Dst D, int zlog) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)D.nz * D.nx * D.ny;
  if (t >= n) return;
  const int k = (int)(t % D.nz);
  const long r = t / D.nz;
  const int i = (int)(r % D.nx), j = (int)(r / D.nx);
  const double x = D.x[i], y = D.y[j], z = D.z[k];
  int ix0, ix1, jy0, jy1, kz0, kz1;
  bracket(S.xp, S.nxp, x, ix0, ix1);
  bracket(S.yp, S.nyp, y, jy0, jy1);
  bracket(S.zp, S.nzp, z, kz0, kz1);
  const double f0 = at_xy(S, kz0, ix0, ix1, jy0, jy1, x, y);
  const double f1 = at_xy(S, kz1, ix0, ix1, jy0, jy1, x, y);
  const double v = zlog ? exp(lin(z, S.zp[kz0], S.zp[kz1], log(f0), log(f1))) : lin(z, S.zp[kz0], S.zp[kz1], f0, f1);
  D.f[(size_t)k * D.dk + (size_t)i * D.di + (size_t)j * D.dj] = v;
}

__global__ void __launch_bounds__(256) interp2d_kernel(Src S, Dst D) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)D.nx * D.ny) return;
  const int i = (int)(t % D.nx), j = (int)(t / D.nx);
  const double x = D.x[i], y = D.y[j];
  int ix0, ix1, jy0, jy1;
  bracket(S.xp, S.nxp, x, ix0, ix1);
  bracket(S.yp, S.nyp, y, jy0, jy1);
  D.f[(size_t)i * D.di + (size_t)j * D.dj] = at_xy(S, 0, ix0, ix1, jy0, jy1, x, y);
}

// interp_fields_ng.F90:17-63: f at t = ta..tb between f0 (t = 0) and f1 (t = t1)
__global__ void __launch_bounds__(256) time_interp_kernel(const double *__restrict__ f0,
                                                          const double *__restrict__ f1, double *__restrict__ f,
                                                          long n, int t1, int ta, int nt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double a = f0[p], b = f1[p], x1 = (double)t1;
  for (int s = 0; s < nt; ++s) f[p + n * (long)s] = lin((double)(ta + s), 0., x1, a, b);
}

// Batches: up to kMaxJobs fields with the same geometry in one launch (the
// coupling of one nest level takes 22 fields per map_out, 11 per set_bndry).
constexpr int kMaxJobs = 32;
struct Jobs3 {
  const double *src[kMaxJobs];
  double *dst[kMaxJobs];
  unsigned zlog;  // bit q: job q log-linear in z
  int n;
};

__global__ void __launch_bounds__(256) interp3d_batch_kernel(Src S, Dst D, Jobs3 J) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long np = (long)D.nz * D.nx * D.ny;
  const int q = (int)blockIdx.y;
  if (t >= np || q >= J.n) return;
  const int k = (int)(t % D.nz);
  const long r = t / D.nz;
  const int i = (int)(r % D.nx), j = (int)(r / D.nx);
  const double x = D.x[i], y = D.y[j], z = D.z[k];
  int ix0, ix1, jy0, jy1, kz0, kz1;
  bracket(S.xp, S.nxp, x, ix0, ix1);
  bracket(S.yp, S.nyp, y, jy0, jy1);
  bracket(S.zp, S.nzp, z, kz0, kz1);
  Src Sq = S;
  Sq.f = J.src[q];
  const double f0 = at_xy(Sq, kz0, ix0, ix1, jy0, jy1, x, y);
  const double f1 = at_xy(Sq, kz1, ix0, ix1, jy0, jy1, x, y);
  const double v = (J.zlog >> q) & 1u ? exp(lin(z, S.zp[kz0], S.zp[kz1], log(f0), log(f1)))
                                      : lin(z, S.zp[kz0], S.zp[kz1], f0, f1);
  J.dst[q][(size_t)k * D.dk + (size_t)i * D.di + (size_t)j * D.dj] = v;
}

__global__ void __launch_bounds__(256) interp2d_batch_kernel(Src S, Dst D, Jobs3 J) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = (int)blockIdx.y;
  if (t >= (long)D.nx * D.ny || q >= J.n) return;
  const int i = (int)(t % D.nx), j = (int)(t / D.nx);
  const double x = D.x[i], y = D.y[j];
  int ix0, ix1, jy0, jy1;
  bracket(S.xp, S.nxp, x, ix0, ix1);
  bracket(S.yp, S.nyp, y, jy0, jy1);
  Src Sq = S;
  Sq.f = J.src[q];
  J.dst[q][(size_t)i * D.di + (size_t)j * D.dj] = at_xy(Sq, 0, ix0, ix1, jy0, jy1, x, y);
}

// Time interpolations (time_interp_kernel) and plain copies (copy: dst = f1)
// of fields of different sizes in one launch: blockIdx.y is the job.
struct TJob {
  const double *f0, *f1;
  double *dst;
  long n;
  int t1, ta, nt, copy;
};
struct TJobs {
  TJob j[kMaxJobs];
  int n;
};

__global__ void __launch_bounds__(256) time_batch_kernel(TJobs J) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = (int)blockIdx.y;
  if (q >= J.n) return;
  const TJob &b = J.j[q];
  if (p >= b.n) return;
  if (b.copy) {
    b.dst[p] = b.f1[p];
    return;
  }
  const double f0 = b.f0[p], f1 = b.f1[p], x1 = (double)b.t1;
  for (int s = 0; s < b.nt; ++s) b.dst[p + b.n * (long)s] = lin((double)(b.ta + s), 0., x1, f0, f1);
}

}  // namespace nest
}  // namespace tgx

namespace {
int check_axis(const char *what, int n, const double *p, int min_n) {
  if (n < min_n || !p) return tgx_set_error(TGX_ERR_ARG, "%s: %d points (at least %d) or NULL", what, n, min_n);
  return 0;
}
unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }
}  // namespace

extern "C" int tgx_ng_interp3d_dev(const double *src, int nzp, int nxp, int nyp, int ld, const double *zp,
                                   const double *xp, const double *yp, double *dst, int nz, int nx, int ny,
                                   long dk, long di, long dj, const double *z, const double *x, const double *y,
                                   int zlog, void *stream) {
  using namespace tgx::nest;
  if (int rc = check_axis("tgx_ng_interp3d_dev zp", nzp, zp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp3d_dev xp", nxp, xp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp3d_dev yp", nyp, yp, 2)) return rc;
  if (!src || !dst || !z || !x || !y || ld < nxp || nz < 1 || nx < 1 || ny < 1)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp3d_dev: bad arguments");
  const Src S{src, nzp, nxp, nyp, ld, (long)ld * nyp, zp, xp, yp};
  const Dst D{dst, nz, nx, ny, dk, di, dj, z, x, y};
  const long n = (long)nz * nx * ny;
  hipLaunchKernelGGL(interp3d_kernel, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, S, D, zlog);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_ng_interp2d_dev(const double *src, int nxp, int nyp, int ld, const double *xp, const double *yp,
                                   double *dst, int nx, int ny, long di, long dj, const double *x, const double *y,
                                   void *stream) {
  using namespace tgx::nest;
  if (int rc = check_axis("tgx_ng_interp2d_dev xp", nxp, xp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp2d_dev yp", nyp, yp, 2)) return rc;
  if (!src || !dst || !x || !y || ld < nxp || nx < 1 || ny < 1)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp2d_dev: bad arguments");
  const Src S{src, 1, nxp, nyp, ld, (long)ld * nyp, nullptr, xp, yp};
  const Dst D{dst, 1, nx, ny, 0, di, dj, nullptr, x, y};
  hipLaunchKernelGGL(interp2d_kernel, dim3(blocks((long)nx * ny)), dim3(256), 0, (hipStream_t)stream, S, D);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_ng_time_interp_dev(const double *f0, const double *f1, double *dst, long n, int t1, int ta,
                                      int tb, void *stream) {
  if (!f0 || !f1 || !dst || n < 0 || t1 < 1 || ta < 0 || tb < ta || tb > t1)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_time_interp_dev: bad arguments (n %ld t1 %d ta %d tb %d)", n, t1, ta,
                         tb);
  if (n == 0) return 0;
  hipLaunchKernelGGL(tgx::nest::time_interp_kernel, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, f0, f1,
                     dst, n, t1, ta, tb - ta + 1);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_ng_interp3d_batch_dev(int njobs, const double *const *src, double *const *dst, const int *zlog,
                                         int nzp, int nxp, int nyp, int ld, const double *zp, const double *xp,
                                         const double *yp, int nz, int nx, int ny, long dk, long di, long dj,
                                         const double *z, const double *x, const double *y, void *stream) {
  using namespace tgx::nest;
  if (njobs < 0 || njobs > kMaxJobs || (njobs > 0 && (!src || !dst || !zlog)))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp3d_batch_dev: %d jobs (0..%d)", njobs, kMaxJobs);
  if (njobs == 0) return 0;
  if (int rc = check_axis("tgx_ng_interp3d_batch_dev zp", nzp, zp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp3d_batch_dev xp", nxp, xp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp3d_batch_dev yp", nyp, yp, 2)) return rc;
  if (!z || !x || !y || ld < nxp || nz < 1 || nx < 1 || ny < 1)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp3d_batch_dev: bad arguments");
  Jobs3 J = {};
  for (int q = 0; q < njobs; ++q) {
    if (!src[q] || !dst[q]) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp3d_batch_dev: job %d has a NULL array", q);
    J.src[q] = src[q];
    J.dst[q] = dst[q];
    if (zlog[q]) J.zlog |= 1u << q;
  }
  J.n = njobs;
  const Src S{nullptr, nzp, nxp, nyp, ld, (long)ld * nyp, zp, xp, yp};
  const Dst D{nullptr, nz, nx, ny, dk, di, dj, z, x, y};
  const long n = (long)nz * nx * ny;
  hipLaunchKernelGGL(interp3d_batch_kernel, dim3(blocks(n), (unsigned)njobs), dim3(256), 0, (hipStream_t)stream, S,
                     D, J);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_ng_interp2d_batch_dev(int njobs, const double *const *src, double *const *dst, int nxp, int nyp,
                                         int ld, const double *xp, const double *yp, int nx, int ny, long di,
                                         long dj, const double *x, const double *y, void *stream) {
  using namespace tgx::nest;
  if (njobs < 0 || njobs > kMaxJobs || (njobs > 0 && (!src || !dst)))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp2d_batch_dev: %d jobs (0..%d)", njobs, kMaxJobs);
  if (njobs == 0) return 0;
  if (int rc = check_axis("tgx_ng_interp2d_batch_dev xp", nxp, xp, 2)) return rc;
  if (int rc = check_axis("tgx_ng_interp2d_batch_dev yp", nyp, yp, 2)) return rc;
  if (!x || !y || ld < nxp || nx < 1 || ny < 1)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp2d_batch_dev: bad arguments");
  Jobs3 J = {};
  for (int q = 0; q < njobs; ++q) {
    if (!src[q] || !dst[q]) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_interp2d_batch_dev: job %d has a NULL array", q);
    J.src[q] = src[q];
    J.dst[q] = dst[q];
  }
  J.n = njobs;
  const Src S{nullptr, 1, nxp, nyp, ld, (long)ld * nyp, nullptr, xp, yp};
  const Dst D{nullptr, 1, nx, ny, 0, di, dj, nullptr, x, y};
  hipLaunchKernelGGL(interp2d_batch_kernel, dim3(blocks((long)nx * ny), (unsigned)njobs), dim3(256), 0,
                     (hipStream_t)stream, S, D, J);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_ng_time_interp_batch_dev(int njobs, const double *const *f0, const double *const *f1,
                                            double *const *dst, const long *n, const int *t1, const int *ta,
                                            const int *tb, void *stream) {
  using namespace tgx::nest;
  if (njobs < 0 || njobs > kMaxJobs || (njobs > 0 && (!f0 || !f1 || !dst || !n || !t1 || !ta || !tb)))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_time_interp_batch_dev: %d jobs (0..%d)", njobs, kMaxJobs);
  TJobs J = {};
  long nmax = 0;
  for (int q = 0; q < njobs; ++q) {
    const bool copy = t1[q] == 0;
    if (!f1[q] || !dst[q] || n[q] < 0 || (!copy && (!f0[q] || t1[q] < 1 || ta[q] < 0 || tb[q] < ta[q] || tb[q] > t1[q])))
      return tgx_set_error(TGX_ERR_ARG, "tgx_ng_time_interp_batch_dev: job %d: n %ld t1 %d ta %d tb %d", q, n[q],
                           t1[q], ta[q], tb[q]);
    J.j[q] = TJob{f0[q], f1[q], dst[q], n[q], t1[q], ta[q], copy ? 1 : tb[q] - ta[q] + 1, copy ? 1 : 0};
    if (n[q] > nmax) nmax = n[q];
  }
  J.n = njobs;
  if (njobs == 0 || nmax == 0) return 0;
  hipLaunchKernelGGL(time_batch_kernel, dim3(blocks(nmax), (unsigned)njobs), dim3(256), 0, (hipStream_t)stream, J);
  TGX_HIP(hipGetLastError());
  return 0;
}
