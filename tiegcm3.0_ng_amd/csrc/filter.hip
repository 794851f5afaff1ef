// The high-latitude longitude ring filter of the global model
// (ringfilter.F:1-111), applied after the column solves of dt, duv, minor,
// comp and oplus (filter_tn dt.F:431-511 and its siblings) to the latitudes
// next to each pole.  One wavefront filters one (latitude row, level): the
// row's longitudes are staged in LDS, the Fourier coefficients of the kept
// harmonics are wavefront reductions, and the chunk averages and the
// reconstruction are read back from LDS.
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "tgx_common.h"

namespace tgx {
namespace ring {

constexpr int kMaxOrder = TGX_RING_MAX_ORDER;

// The rows of one launch, by value in the kernel arguments: row r is at
// f + k * plane + off[r] and is filtered with n[r] chunks.
struct Rows {
  long off[TGX_RING_MAX_ROWS];
  int n[TGX_RING_MAX_ROWS];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Each row holds nlon + 4 longitudes (two periodic points on each side).
// LDS: wm[nlon] (then the rebuilt row), fx[nlon], avg[nlon].
__global__ void __launch_bounds__(64) ringfilter_kernel(double *__restrict__ f, long plane, Rows R, int nlon,
                                                        int order, int oplus, double dlamda) {
  extern __shared__ double sh[];
  double *wm = sh, *fx = sh + nlon, *avg = sh + 2 * nlon;
  const int r = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
  const int nchnk = R.n[r];
  double *row = f + (size_t)k * plane + R.off[r];
  // :40-42 the interior longitudes and their mean
  double s0 = 0.;
  for (int i = lane; i < nlon; i += 64) {
    const double w = row[2 + i];
    fx[i] = w;
    s0 += w;
  }
  const double a0 = wave_sum(s0) / nlon;
  // :43-53 Fourier coefficients of the kept harmonics (ang = i dlamda, i = 1..nlon)
  double a[kMaxOrder + 1], b[kMaxOrder + 1];
  for (int m = 1; m <= order; ++m) {
    double sa = 0., sb = 0.;
    for (int i = lane; i < nlon; i += 64) {
      const double ang = (i + 1) * dlamda;
      double sn, cn;
      sincos(m * ang, &sn, &cn);
      sa += fx[i] * cn;
      sb += fx[i] * sn;
    }
    a[m] = 2. * wave_sum(sa) / nlon;
    b[m] = 2. * wave_sum(sb) / nlon;
  }
  // :55-63 the harmonic part and the residual
  for (int i = lane; i < nlon; i += 64) {
    const double ang = (i + 1) * dlamda;
    double v = a0;
    for (int m = 1; m <= order; ++m) {
      double sn, cn;
      sincos(m * ang, &sn, &cn);
      v = v + a[m] * cn + b[m] * sn;
    }
    wm[i] = v;
    fx[i] = fx[i] - v;
  }
  __syncthreads();
  // :65-69 chunk averages, each summed in longitude order
  const int n = nlon / nchnk;
  for (int c = lane; c < nchnk; c += 64) {
    double s = 0.;
    for (int m = 0; m < n; ++m) s = s + fx[c * n + m];
    avg[c] = s / n;
  }
  __syncthreads();
  // :71-97 reconstruction from the four neighbouring chunk averages; the
  // rebuilt value of longitude i replaces wm[i], which only this lane reads
  for (int i = lane; i < nlon; i += 64) {
    const int c = i / n;
    double x = fx[i];
    if (c < nchnk) {  // points past nchnk * n keep their residual (:84-93 stop at nchnk * n)
      const int m = i - c * n + 1;
      const int im2 = (c - 2 + 2 * nchnk) % nchnk, im1 = (c - 1 + nchnk) % nchnk;
      const int ip1 = (c + 1) % nchnk, ip2 = (c + 2) % nchnk;
      const double A0 = avg[c];
      const double left = (-avg[im2] + 7. * avg[im1] + 7. * A0 - avg[ip1]) / 12.;
      const double right = (-avg[im1] + 7. * A0 + 7. * avg[ip1] - avg[ip2]) / 12.;
      if (oplus) {
        const double fa = (left + right - 2. * A0) / (double)(n * n);
        const double fb = (3. * A0 - 2. * left - right) / n;
        x = fa * (3. * (m * m) - 3. * m + 1.) + fb * (2. * m - 1.) + left;
      } else {
        x = (left * (n - m) + right * m) / n;
      }
    }
    wm[i] = x + wm[i];
    row[2 + i] = wm[i];
  }
  __syncthreads();
  // :99-101 periodic points
  if (lane == 0) row[0] = wm[nlon - 2];
  if (lane == 1) row[1] = wm[nlon - 1];
  if (lane == 2) row[nlon + 2] = wm[0];
  if (lane == 3) row[nlon + 3] = wm[1];
}

}  // namespace ring
}  // namespace tgx

// Launcher shared by the device entry point and the host drop-in: nrows rows
// (host arrays rowoff, nchnk, at most TGX_RING_MAX_ROWS) of nk levels.
extern "C" int tgx_launch_ringfilter(double *f, int nk, long plane, int nrows, const long *rowoff,
                                     const int *nchnk, int nlon, int order, int oplus, double dlamda,
                                     hipStream_t s) {
  if (nrows <= 0 || nk <= 0) return 0;
  if (nrows > TGX_RING_MAX_ROWS) return tgx_set_error(TGX_ERR_ARG, "ring filter: %d rows > %d", nrows, TGX_RING_MAX_ROWS);
  if (nlon < 4 || nlon > TGX_RING_MAX_NLON || order < 0 || order > tgx::ring::kMaxOrder || order >= nlon)
    return tgx_set_error(TGX_ERR_ARG, "ring filter: nlon %d (4..%d) or order %d (0..%d, < nlon)", nlon,
                         TGX_RING_MAX_NLON, order, tgx::ring::kMaxOrder);
  tgx::ring::Rows R;
  for (int r = 0; r < nrows; ++r) {
    if (nchnk[r] < 1 || nchnk[r] > nlon)
      return tgx_set_error(TGX_ERR_ARG, "ring filter: %d chunks for %d longitudes", nchnk[r], nlon);
    R.off[r] = rowoff[r];
    R.n[r] = nchnk[r];
  }
  const size_t lds = sizeof(double) * 3 * (size_t)nlon;
  hipLaunchKernelGGL(tgx::ring::ringfilter_kernel, dim3((unsigned)nrows, (unsigned)nk), dim3(64), lds, s, f,
                     plane, R, nlon, order, oplus, dlamda);
  TGX_HIP(hipGetLastError());
  return 0;
}

// The latitudes next to each pole that the chunk table filters
// (ringfilter.F:23-34), in the level-plane layout of tgx_step.h.
extern "C" int tgx_ringfilter_dev(double *f, int nk, int nlon, int nlat, int ld, int order, int oplus,
                                  double dlamda, int nlat_filter, const int *chunk, void *stream) {
  if (!f || nk <= 0 || nlat <= 0 || ld < nlon + 4 || nlat_filter < 0 || nlat_filter > TGX_RING_MAX_ROWS / 2 ||
      (nlat_filter > 0 && !chunk))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ringfilter_dev: bad arguments (nk %d nlon %d nlat %d ld %d nlat_filter %d)",
                         nk, nlon, nlat, ld, nlat_filter);
  long off[TGX_RING_MAX_ROWS];
  int nc[TGX_RING_MAX_ROWS];
  int nrows = 0;
  for (int lat = 1; lat <= nlat; ++lat) {
    int c = 0;
    if (lat <= nlat_filter) c = chunk[lat - 1];
    else if (lat >= nlat - nlat_filter + 1) c = chunk[nlat - lat];
    if (c == 0) continue;
    off[nrows] = (long)(lat + 1) * ld;  // latitude lat is row lat + 1 of the nlat + 4 rows
    nc[nrows] = c;
    ++nrows;
  }
  return tgx_launch_ringfilter(f, nk, (long)ld * (nlat + 4), nrows, off, nc, nlon, order, oplus, dlamda,
                               (hipStream_t)stream);
}
