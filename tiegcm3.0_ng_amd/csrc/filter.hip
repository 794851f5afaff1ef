// The high-latitude longitude ring filter of the global model
// (ringfilter.F:1-111), applied after the column solves of dt, duv, minor,
// comp and oplus (filter_tn dt.F:431-511 and its siblings) to the latitudes
// next to each pole.  A wavefront filters one level of one latitude row at a
// time: the row's longitudes in registers, the Fourier coefficients of the
// kept harmonics as wavefront reductions over a table of cos(m ang) and
// sin(m ang) (the only transcendental work), the residual and the chunk
// averages in its own LDS region.  Two forms of the launch:
//  - tgx_ringfilter_dev and the host drop-in: a block of four wavefronts
//    takes one row and a run of levels and first builds the table in LDS;
//  - the global routines (tgx_ringfilter_post_dev): the table is built once
//    per call in the workspace, a block is one wavefront and one level (small
//    blocks that find room beside the other routines' column kernels), all
//    the routine's fields go in one launch, and the routine's post pass (time
//    smoothing, top level, floors) is applied to each filtered value as it is
//    written.
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "tgx_common.h"

namespace tgx {
namespace ring {

constexpr int kMaxOrder = TGX_RING_MAX_ORDER;
constexpr int kLdsMax = 65536;  // dynamic LDS of one block
static_assert(64 * 32 >= TGX_RING_MAX_NLON, "ringfilter_kernel<32> covers the longest row");

// The rows of one launch, by value in the kernel arguments: row r is at
// f + k * plane + off[r] and is filtered with n[r] chunks.
struct Rows {
  long off[TGX_RING_MAX_ROWS];
  int n[TGX_RING_MAX_ROWS];
};

// The post pass that follows the filter in each global routine, on the
// filtered latitudes: dt.F:396-424 (POST_DT), duv.F:501-514 (POST_DUV),
// minor.F:556-597 (POST_MINOR), oplus.F:911-934 (POST_OPLUS).
enum PostKind { POST_NONE = 0, POST_DT, POST_DUV, POST_MINOR, POST_OPLUS };
constexpr double DTSMOOTH = 0.95, DTSMOOTH_DIV2 = 0.5 * (1. - DTSMOOTH), SPVAL = 1.e36;

// The fields of one launch: f[q] is filtered in place on levels 0..nlev-1;
// with a post pass, xnm[q] receives the n-1 field from the step inputs fin[q]
// (n) and fnmi[q] (n-1), and a level nlev < nk (the top) is set to spval or
// zero without filtering.
struct Fields {
  double *f[4], *xnm[4];
  const double *fin[4], *fnmi[4];
  int nf, nlev, kind;
};

// The post pass of one filtered value y with the step inputs fi (n) and fm
// (n-1): the solution xo and the n-1 field xn, in the reference's order.
__device__ __forceinline__ void post_point(int kind, double y, double fi, double fm, double &xo, double &xn) {
  if (kind == POST_DT) {
    xn = DTSMOOTH_DIV2 * (fm + y) + DTSMOOTH * fi;
    xo = y < 100. ? 100. : y;
  } else if (kind == POST_DUV) {
    xn = DTSMOOTH_DIV2 * (fm + y) + DTSMOOTH * fi;
    xo = y;
  } else if (kind == POST_MINOR) {
    const double small = 1.e-12;
    const double nm = DTSMOOTH * fi + DTSMOOTH_DIV2 * (fm + y);
    xo = y < small ? small : y;
    xn = nm < small ? small : nm;
  } else {  // POST_OPLUS
    const double nm = DTSMOOTH * fi + DTSMOOTH_DIV2 * (fm + y);
    xo = y < 0. ? 0. : y;
    xn = nm < 0. ? 0. : nm;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// NPL: longitudes per lane (nlon <= 64 NPL); lane l holds i = l + 64 t.
// LDS: the table cs[(m-1) nlon + i], sn[...] for m = 1..order, then per
// wavefront fx[nlon] (the residual) and avg[nlon].  The launch covers nkp
// levels of each of the P.nf fields, field-major; wavefront w of block
// (r, y) takes the levels (y lpw + kk) W + w, kk < lpw, of that order.
// GTAB: the table is given in device memory (gtab: cos of harmonics
// 1..tord at (m-1) nlon + i, then their sin; ring_table_kernel) instead of
// being built in LDS; the LDS then holds the wavefronts' fx and avg only.
// The kernels below wrap this body.
template <int NPL, bool GTAB>
__device__ __forceinline__ void ringfilter_body(const Fields &P, long plane, const Rows &R, int nlon, int nkp,
                                                int order, int oplus, double dlamda, int lpw,
                                                const double *__restrict__ gtab, int tord) {
  extern __shared__ double sh[];
  const int W = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x;
  const int nchnk = R.n[r];
  const int ntab = GTAB ? 0 : order * nlon;
  const double *cs = GTAB ? gtab : sh, *sn = GTAB ? gtab + (size_t)tord * nlon : sh + ntab;
  double *fx = sh + 2 * ntab + (size_t)wave * 2 * nlon, *avg = fx + nlon;
  for (int t = threadIdx.x; t < ntab; t += blockDim.x) {
    const int m = t / nlon + 1, i = t - (m - 1) * nlon;
    const double ang = (i + 1) * dlamda;
    double s, c;
    sincos(m * ang, &s, &c);
    sh[t] = c;
    sh[ntab + t] = s;
  }
  auto trig = [&](int m, int i, double &c, double &s) {
    c = cs[(m - 1) * nlon + i];
    s = sn[(m - 1) * nlon + i];
  };
  const int n = nlon / nchnk;
  for (int kk = 0; kk < lpw; ++kk) {
    const int kq = ((int)blockIdx.y * lpw + kk) * W + wave;
    const bool live = kq < P.nf * nkp;  // every wavefront reaches every barrier
    const int q = live ? kq / nkp : 0, k = live ? kq - q * nkp : 0;
    const bool act = live && k < P.nlev;  // a level to filter (else the top, or nothing)
    const size_t off = (size_t)k * plane + R.off[r];
    double *row = P.f[q] + off;
    double w[NPL], v[NPL];
    __syncthreads();  // the table is written; this wavefront's fx, avg are free
    if (act) {
      // :40-42 the interior longitudes and their mean
      double s0 = 0.;
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        const int i = lane + 64 * t;
        w[t] = i < nlon ? row[2 + i] : 0.;
        if (i < nlon) s0 += w[t];
      }
      // :43-53 the Fourier sums of the kept harmonics (ang = i dlamda,
      // i = 1..nlon), then every sum's wavefront reduction, their steps
      // interleaved
      double sum[2 * kMaxOrder + 1];
      sum[0] = s0;
#pragma unroll
      for (int m = 1; m <= kMaxOrder; ++m) {
        sum[2 * m - 1] = sum[2 * m] = 0.;
        if (m > order) continue;
        double sa = 0., sb = 0.;
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
          const int i = lane + 64 * t;
          if (i < nlon) {
            double cn, sg;
            trig(m, i, cn, sg);
            sa += w[t] * cn;
            sb += w[t] * sg;
          }
        }
        sum[2 * m - 1] = sa;
        sum[2 * m] = sb;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int q = 0; q <= 2 * kMaxOrder; ++q)
          if (q <= 2 * order) sum[q] += __shfl_xor(sum[q], o, 64);
      }
      // :42, :53-61 the mean and the harmonic part, harmonics in ascending
      // order as the reference adds them
      const double a0 = sum[0] / nlon;
#pragma unroll
      for (int t = 0; t < NPL; ++t) v[t] = a0;
#pragma unroll
      for (int m = 1; m <= kMaxOrder; ++m) {
        if (m > order) continue;
        const double am = 2. * sum[2 * m - 1] / nlon, bm = 2. * sum[2 * m] / nlon;
#pragma unroll
        for (int t = 0; t < NPL; ++t) {
          const int i = lane + 64 * t;
          if (i < nlon) {
            double cn, sg;
            trig(m, i, cn, sg);
            v[t] = v[t] + am * cn + bm * sg;
          }
        }
      }
      // :63 the residual
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        const int i = lane + 64 * t;
        if (i >= nlon) continue;
        w[t] = w[t] - v[t];
        fx[i] = w[t];
      }
    }
    __syncthreads();
    // :65-69 chunk averages, each summed in longitude order; the values are
    // read from LDS sixteen at a time so that the reads overlap the additions
    if (act) {
      for (int c = lane; c < nchnk; c += 64) {
        const double *q = fx + c * n;
        double s = 0.;
        int m = 0;
        for (; m + 16 <= n; m += 16) {
          double x[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) x[u] = q[m + u];
#pragma unroll
          for (int u = 0; u < 16; ++u) s = s + x[u];
        }
        for (; m < n; ++m) s = s + q[m];
        avg[c] = s / n;
      }
    }
    __syncthreads();
    if (!act) {
      if (live && P.kind != POST_NONE) {  // the top level: spval (dt.F:407-408, duv.F:385-386, :512-513), zero (oplus.F:922-923)
        const double top = P.kind == POST_OPLUS ? 0. : SPVAL;
        double *xr = P.xnm[q] + off;
        for (int i = lane; i < nlon + 4; i += 64) {
          row[i] = top;
          xr[i] = top;
        }
      }
      continue;
    }
    const bool post = P.kind != POST_NONE;
    const double *fin = post ? P.fin[q] + off : row, *fnmi = post ? P.fnmi[q] + off : row;  // the step inputs
    double *xr = post ? P.xnm[q] + off : row;
    // :71-101 reconstruction from the four neighbouring chunk averages, the
    // harmonic part added back, the periodic points.  Longitude i = lane + 64 t
    // is point m (1-based) of chunk c; both advance by 64 = q64 n + r64 per t.
    const int q64 = 64 / n, r64 = 64 - q64 * n;
    int c = lane / n, m = lane - c * n + 1;
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
      const int i = lane + 64 * t;
      if (t > 0) {
        c += q64;
        m += r64;
        if (m > n) {
          m -= n;
          ++c;
        }
      }
      if (i >= nlon) continue;
      double x = w[t];
      if (c < nchnk) {  // points past nchnk * n keep their residual (:84-93 stop at nchnk * n)
        int im2 = c - 2, im1 = c - 1, ip1 = c + 1, ip2 = c + 2;  // :73-81, periodic in the chunks
        if (im2 < 0) im2 += nchnk;
        if (im2 < 0) im2 += nchnk;
        if (im1 < 0) im1 += nchnk;
        if (ip1 >= nchnk) ip1 -= nchnk;
        if (ip2 >= nchnk) ip2 -= nchnk;
        if (ip2 >= nchnk) ip2 -= nchnk;
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
      const double y = x + v[t];
      double xo = y, xn = 0.;
      if (post) {
        post_point(P.kind, y, fin[2 + i], fnmi[2 + i], xo, xn);
        xr[2 + i] = xn;
        if (i == nlon - 2) xr[0] = xn;
        if (i == nlon - 1) xr[1] = xn;
        if (i == 0) xr[nlon + 2] = xn;
        if (i == 1) xr[nlon + 3] = xn;
      }
      row[2 + i] = xo;
      if (i == nlon - 2) row[0] = xo;
      if (i == nlon - 1) row[1] = xo;
      if (i == 0) row[nlon + 2] = xo;
      if (i == 1) row[nlon + 3] = xo;
    }
  }
}

template <int NPL>
__global__ void __launch_bounds__(256) ringfilter_kernel(Fields P, long plane, Rows R, int nlon, int nkp, int order,
                                                         int oplus, double dlamda, int lpw) {
  ringfilter_body<NPL, false>(P, plane, R, nlon, nkp, order, oplus, dlamda, lpw, nullptr, 0);
}

// With the table in device memory: one wavefront per block, at most 128
// VGPRs, so that four wavefronts of it fit on a SIMD beside the column
// kernels of the routines running at the same time.
template <int NPL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
ringfilter_gtab_kernel(Fields P, long plane, Rows R, int nlon, int nkp, int order, int oplus, double dlamda,
                       const double *__restrict__ gtab, int tord) {
  ringfilter_body<NPL, true>(P, plane, R, nlon, nkp, order, oplus, dlamda, 1, gtab, tord);
}

// The table of ringfilter_gtab_kernel: cos(m ang) and sin(m ang) of
// longitude i (ang = (i + 1) dlamda, i < nlon) for m = 1..tord, the values
// the LDS form computes.
__global__ void __launch_bounds__(256) ring_table_kernel(double *__restrict__ tab, int nlon, int tord,
                                                         double dlamda) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tord * nlon) return;
  const int m = t / nlon + 1, i = t - (m - 1) * nlon;
  const double ang = (i + 1) * dlamda;
  double s, c;
  sincos(m * ang, &s, &c);
  tab[t] = c;
  tab[(size_t)tord * nlon + t] = s;
}

// The same filter where the table does not fit in LDS (long rows, many
// harmonics): one wavefront per (row, level), sincos evaluated where needed.
// LDS: wm[nlon] (then the rebuilt row), fx[nlon], avg[nlon].
__global__ void __launch_bounds__(64) ringfilter_direct_kernel(double *__restrict__ f, long plane, Rows R, int nlon,
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

// Whether the table and four wavefronts' rows fit in a block's LDS; if not,
// the launcher filters with ringfilter_direct_kernel (one field, no post pass).
inline bool table_fits(int nlon, int order, int *waves) {
  const size_t tab = sizeof(double) * 2 * (size_t)order * nlon, per = sizeof(double) * 2 * (size_t)nlon;
  int W = 4;
  while (W > 1 && tab + W * per > kLdsMax) W >>= 1;
  if (waves) *waves = W;
  return tab + W * per <= kLdsMax;
}

template <int NPL>
int launch(const Fields &P, int nkp, long plane, int nrows, const Rows &R, int nlon, int order, int oplus,
           double dlamda, const double *gtab, int tord, hipStream_t s) {
  int W;
  if (gtab) {
    // one wavefront per block and level: small blocks (9 KB of LDS at 576
    // longitudes) that find room beside the column kernels of the other
    // routines running at the same time
    const size_t lds = sizeof(double) * 2 * (size_t)nlon;
    const unsigned ny = (unsigned)((long)P.nf * nkp);
    hipLaunchKernelGGL(ringfilter_gtab_kernel<NPL>, dim3((unsigned)nrows, ny), dim3(64), lds, s, P, plane, R, nlon,
                       nkp, order, oplus, dlamda, gtab, tord);
  } else if (!table_fits(nlon, order, &W)) {
    if (P.nf != 1 || P.kind != POST_NONE || nkp != P.nlev)
      return tgx_set_error(TGX_ERR_ARG, "ring filter: fused post pass needs the LDS table (nlon %d, order %d)",
                           nlon, order);
    hipLaunchKernelGGL(ringfilter_direct_kernel, dim3((unsigned)nrows, (unsigned)nkp), dim3(64),
                       sizeof(double) * 3 * (size_t)nlon, s, P.f[0], plane, R, nlon, order, oplus, dlamda);
  } else {
    // wavefronts per block and levels per wavefront: the table in LDS beside
    // each wavefront's fx and avg, at most kLdsMax bytes
    const size_t lds = sizeof(double) * (2 * (size_t)order * nlon + 2 * (size_t)W * nlon);
    const int lpw = 3;
    const long nlev = (long)P.nf * nkp;
    const unsigned ny = (unsigned)((nlev + W * lpw - 1) / (W * lpw));
    hipLaunchKernelGGL(ringfilter_kernel<NPL>, dim3((unsigned)nrows, ny), dim3(64 * W), lds, s, P, plane, R, nlon,
                       nkp, order, oplus, dlamda, lpw);
  }
  TGX_HIP(hipGetLastError());
  return 0;
}

inline int launch_any(const Fields &P, int nkp, long plane, int nrows, const Rows &R, int nlon, int order,
                      int oplus, double dlamda, const double *gtab, int tord, hipStream_t s) {
  if (nlon <= 64 * 2) return launch<2>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
  if (nlon <= 64 * 3) return launch<3>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
  if (nlon <= 64 * 5) return launch<5>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
  if (nlon <= 64 * 9) return launch<9>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
  if (nlon <= 64 * 18) return launch<18>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
  return launch<32>(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
}

}  // namespace ring
}  // namespace tgx

namespace {
// Validates a launch of nrows rows (host arrays rowoff, nchnk, at most
// TGX_RING_MAX_ROWS) and launches it.
int launch_rows(const tgx::ring::Fields &P, int nkp, long plane, int nrows, const long *rowoff, const int *nchnk,
                int nlon, int order, int oplus, double dlamda, hipStream_t s, const double *gtab = nullptr,
                int tord = 0) {
  if (nrows <= 0 || nkp <= 0 || P.nf <= 0) return 0;
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
  if (gtab && order > tord)
    return tgx_set_error(TGX_ERR_ARG, "ring filter: order %d beyond the table's %d", order, tord);
  return tgx::ring::launch_any(P, nkp, plane, nrows, R, nlon, order, oplus, dlamda, gtab, tord, s);
}

// The latitudes next to each pole that the chunk table filters
// (ringfilter.F:23-34), as row offsets in the level-plane layout of
// tgx_step.h (latitude lat is row lat + 1 of the nlat + 4 rows).
int polar_rows(int nlat, int ld, int nlat_filter, const int *chunk, long *off, int *nc) {
  int nrows = 0;
  for (int lat = 1; lat <= nlat; ++lat) {
    int c = 0;
    if (lat <= nlat_filter) c = chunk[lat - 1];
    else if (lat >= nlat - nlat_filter + 1) c = chunk[nlat - lat];
    if (c == 0) continue;
    off[nrows] = (long)(lat + 1) * ld;
    nc[nrows] = c;
    ++nrows;
  }
  return nrows;
}

tgx::ring::Fields one_field(double *f, int nk) {
  tgx::ring::Fields P = {};
  P.f[0] = f;
  P.nf = 1;
  P.nlev = nk;
  P.kind = tgx::ring::POST_NONE;
  return P;
}
}  // namespace

// Launcher of the host drop-in: nrows rows of nk levels of one field.
extern "C" int tgx_launch_ringfilter(double *f, int nk, long plane, int nrows, const long *rowoff,
                                     const int *nchnk, int nlon, int order, int oplus, double dlamda,
                                     hipStream_t s) {
  return launch_rows(one_field(f, nk), nk, plane, nrows, rowoff, nchnk, nlon, order, oplus, dlamda, s);
}

extern "C" int tgx_ringfilter_dev(double *f, int nk, int nlon, int nlat, int ld, int order, int oplus,
                                  double dlamda, int nlat_filter, const int *chunk, void *stream) {
  if (!f || nk <= 0 || nlat <= 0 || ld < nlon + 4 || nlat_filter < 0 || nlat_filter > TGX_RING_MAX_ROWS / 2 ||
      (nlat_filter > 0 && !chunk))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ringfilter_dev: bad arguments (nk %d nlon %d nlat %d ld %d nlat_filter %d)",
                         nk, nlon, nlat, ld, nlat_filter);
  long off[TGX_RING_MAX_ROWS];
  int nc[TGX_RING_MAX_ROWS];
  const int nrows = polar_rows(nlat, ld, nlat_filter, chunk, off, nc);
  return launch_rows(one_field(f, nk), nk, (long)ld * (nlat + 4), nrows, off, nc, nlon, order, oplus, dlamda,
                     (hipStream_t)stream);
}

// The cos/sin table of tgx_ringfilter_post_dev (harmonics 1..tord of nlon
// longitudes, 2 tord nlon doubles at tab), on the stream.
extern "C" int tgx_ring_table_dev(double *tab, int nlon, int tord, double dlamda, hipStream_t s) {
  if (!tab || nlon < 4 || nlon > TGX_RING_MAX_NLON || tord < 1 || tord > tgx::ring::kMaxOrder)
    return tgx_set_error(TGX_ERR_ARG, "ring filter table: nlon %d (4..%d), order %d (1..%d)", nlon, TGX_RING_MAX_NLON,
                         tord, tgx::ring::kMaxOrder);
  const int n = tord * nlon;
  hipLaunchKernelGGL(tgx::ring::ring_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tab, nlon,
                     tord, dlamda);
  TGX_HIP(hipGetLastError());
  return 0;
}

// The ring filters of one global routine in one launch: the nf (<= 4) fields
// f[q] on their polar latitudes, levels 0..nlev-1, each filtered value then
// passed through the routine's post pass `kind` (tgx::ring::PostKind) with
// the step inputs fin[q], fnmi[q] into f[q] and xnm[q]; with a post pass and
// nlev < nk, the top level nlev is set as the routine sets it.  The periodic
// points of the filtered rows are rewritten; the other latitudes are not
// touched.  gtab: the table of tgx_ring_table_dev for harmonics 1..tord
// (order <= tord), filled on the same stream.  Used by the global routines
// (step_api.cpp).
extern "C" int tgx_ringfilter_post_dev(int nf, double *const *f, double *const *xnm, const double *const *fin,
                                       const double *const *fnmi, int kind, int nlev, int nk, int nlon, int nlat,
                                       int ld, int order, int oplus, double dlamda, int nlat_filter,
                                       const int *chunk, const double *gtab, int tord, hipStream_t s) {
  if (nf < 1 || nf > 4 || kind < tgx::ring::POST_NONE || kind > tgx::ring::POST_OPLUS || nlev < 1 || nlev > nk ||
      (kind != tgx::ring::POST_NONE && nk - nlev > 1) || nlat <= 0 || ld < nlon + 4 || nlat_filter < 0 ||
      nlat_filter > TGX_RING_MAX_ROWS / 2 || (nlat_filter > 0 && !chunk) || !gtab)
    return tgx_set_error(TGX_ERR_ARG, "ring filter + post: bad arguments (nf %d kind %d nlev %d nk %d)", nf, kind,
                         nlev, nk);
  tgx::ring::Fields P = {};
  for (int q = 0; q < nf; ++q) {
    P.f[q] = f[q];
    if (kind != tgx::ring::POST_NONE) {
      P.xnm[q] = xnm[q];
      P.fin[q] = fin[q];
      P.fnmi[q] = fnmi[q];
    }
  }
  P.nf = nf;
  P.nlev = nlev;
  P.kind = kind;
  long off[TGX_RING_MAX_ROWS];
  int nc[TGX_RING_MAX_ROWS];
  const int nrows = polar_rows(nlat, ld, nlat_filter, chunk, off, nc);
  return launch_rows(P, kind != tgx::ring::POST_NONE ? nk : nlev, (long)ld * (nlat + 4), nrows, off, nc, nlon,
                     order, oplus, dlamda, s, gtab, tord);
}
