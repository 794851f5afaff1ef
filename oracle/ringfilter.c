/*
 * Oracle (test infrastructure only): the high-latitude longitude ring filter
 * of the global model, ringfilter.F:1-111, restated in the reference's
 * evaluation order.  The chunk table (cons.F:331-355, nlat_filter and
 * chunk_array per resolution) is an input.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tgx_oracle.h"

extern int tgx_o_mutation_id(void);
#define MUTR(id) (tgx_o_mutation_id() == (id))

/* ringfilter.F:1-111 on one latitude row.  f is the reference's
 * f(nlonp4, lev0:lev1): nlon+4 longitudes (two periodic points on each side)
 * fastest, nlev levels.  nchnk = 0 leaves the row untouched (:23-34); order
 * Fourier harmonics are kept; oplus selects the O+ reconstruction (:82-89).
 * Returns 0, or -1 when order >= nlon (the reference stops the model). */
int tgx_o_ringfilter_row(double *f, int nlon, int nlev, int order, int oplus, int nchnk, double dlamda) {
  if (nchnk == 0) return 0;
  if (order >= nlon) return -1;
  const int nlonp4 = nlon + 4, n = nlon / nchnk;
  double *w = malloc(sizeof(double) * nlon), *wm = malloc(sizeof(double) * nlon);
  double *fx = malloc(sizeof(double) * nlon), *avg = malloc(sizeof(double) * nchnk);
  double *a = malloc(sizeof(double) * (order + 1)), *b = malloc(sizeof(double) * (order + 1));
  for (int l = 0; l < nlev; ++l) {
    double *row = f + (size_t)nlonp4 * l;
    for (int i = 0; i < nlon; ++i) w[i] = row[2 + i];             /* :40 w = f(3:nlonp2,:) */
    double a0 = 0.;
    for (int i = 0; i < nlon; ++i) a0 = a0 + w[i];
    a0 = a0 / nlon;                                                /* :42 */
    for (int m = 1; m <= order; ++m) a[m] = b[m] = 0.;
    for (int i = 1; i <= nlon; ++i) {                              /* :45-51 */
      double ang = i * dlamda;
      for (int m = 1; m <= order; ++m) {
        a[m] = a[m] + w[i - 1] * cos(m * ang);
        b[m] = b[m] + w[i - 1] * sin(m * ang);
      }
    }
    for (int m = 1; m <= order; ++m) {
      a[m] = 2. * a[m] / nlon;
      b[m] = 2. * b[m] / nlon;
    }
    for (int i = 1; i <= nlon; ++i) {                              /* :55-61 */
      double ang = i * dlamda;
      double v = a0;
      for (int m = 1; m <= order; ++m) v = v + a[m] * cos(m * ang) + b[m] * sin(m * ang);
      wm[i - 1] = v;
    }
    for (int i = 0; i < nlon; ++i) fx[i] = w[i] - wm[i];           /* :63 */
    for (int c = 0; c < nchnk; ++c) {                              /* :67-69 */
      double s = 0.;
      for (int m = 0; m < n; ++m) s = s + fx[c * n + m];
      avg[c] = s / n;
    }
    for (int c = 1; c <= nchnk; ++c) {                             /* :71-95 */
      int im2 = c - 2, im1 = c - 1, ip1 = c + 1, ip2 = c + 2;
      if (im2 <= 0) im2 = im2 + nchnk;
      if (im1 <= 0) im1 = im1 + nchnk;
      if (ip1 >= nchnk + 1) ip1 = ip1 - nchnk;
      if (ip2 >= nchnk + 1) ip2 = ip2 - nchnk;
      const double A2 = avg[im2 - 1], A1 = avg[im1 - 1], A0 = avg[c - 1], P1 = avg[ip1 - 1], P2 = avg[ip2 - 1];
      const double seven = MUTR(701) ? 6. : 7.;
      double left = (-A2 + seven * A1 + 7. * A0 - P1) / 12.;
      double right = (-A1 + 7. * A0 + 7. * P1 - P2) / 12.;
      double fa = (left + right - 2. * A0) / (MUTR(704) ? (double)n : (double)(n * n));
      double fb = (3. * A0 - 2. * left - right) / n;
      if (MUTR(702)) fb = -fb;
      if (oplus) {
        for (int m = 1; m <= n; ++m)
          fx[(c - 1) * n + m - 1] = fa * (3. * (m * m) - 3. * m + 1.) + fb * (2. * m - 1.) + left;
      } else {
        for (int m = 1; m <= n; ++m) fx[(c - 1) * n + m - 1] = (left * (n - m) + right * m) / n;
      }
    }
    for (int i = 0; i < nlon; ++i) w[i] = fx[i] + wm[i];           /* :97 */
    for (int i = 0; i < nlon; ++i) row[2 + i] = w[i];              /* :99-101 */
    const int sh = MUTR(703) ? 1 : 0;
    row[0] = w[nlon - 2 - sh];
    row[1] = w[nlon - 1 - sh];
    row[nlonp4 - 2] = w[0];
    row[nlonp4 - 1] = w[1];
  }
  free(w); free(wm); free(fx); free(avg); free(a); free(b);
  return 0;
}

/* The chunk count of latitude lat (1-based) for a table of nlat_filter
 * entries (ringfilter.F:23-34; the O+ branch assigns nothing). */
int tgx_o_ringfilter_chunks(int lat, int nlat, int nlat_filter, const int *chunk) {
  if (lat <= nlat_filter) return chunk[lat - 1];
  if (lat >= nlat - nlat_filter + 1) return chunk[nlat - lat];
  return 0;
}
