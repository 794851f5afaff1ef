/*
 * Oracle (test infrastructure only): the interpolation between nest levels,
 * interp.F90:1-128 (interp1d with its bracket search find, interp2d, interp3d),
 * restated in the reference's evaluation order, and the linear time
 * interpolation of interp_fields_ng.F90:17-63 built on interp1d.  Arrays are
 * in the reference's Fortran layout (first index fastest).
 */
#include <math.h>
#include <stdlib.h>

#include "tgx_oracle.h"

extern int tgx_o_mutation_id(void);
#define MUTI(id) (tgx_o_mutation_id() == (id))

/* interp.F90:105-126 (find): 0 below xp(1), nx at or above xp(nx), else the
 * 1-based i with xp(i) <= x0 < xp(i+1) by bisection. */
int tgx_o_interp_find(const double *xp, int nx, double x0) {
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

/* interp.F90:72-103 (interp1d) at one point x: linear in the bracket, the end
 * intervals extrapolated. */
static double interp1(double x, const double *xp, int nxp, const double *fp, long stride) {
  int i0 = tgx_o_interp_find(xp, nxp, x);
  int i1 = i0 + 1;
  if (i0 <= 0) {
    i0 = 1;
    i1 = 2;
  }
  if (i1 >= nxp + 1) {
    i0 = nxp - 1;
    i1 = nxp;
  }
  if (MUTI(801)) i1 = i0;  /* planted: degenerate bracket */
  const double x0 = xp[i0 - 1], x1 = xp[i1 - 1];
  const double f0 = fp[(long)(i0 - 1) * stride], f1 = fp[(long)(i1 - 1) * stride];
  if (MUTI(802)) return ((x - x0) * f0 + (x1 - x) * f1) / (x1 - x0);  /* planted: weights exchanged */
  return ((x1 - x) * f0 + (x - x0) * f1) / (x1 - x0);
}

void tgx_o_interp1d(const double *x, int nx, const double *xp, int nxp, const double *fp, double *f) {
  for (int i = 0; i < nx; ++i) f[i] = interp1(x[i], xp, nxp, fp, 1);
}

/* interp.F90:50-70 (interp2d): fp(nxp, nyp); first along y for every xp, then
 * along x.  f(nx, ny). */
void tgx_o_interp2d(const double *x, int nx, const double *y, int ny, const double *xp, int nxp, const double *yp,
                    int nyp, const double *fp, double *f) {
  double *aux = malloc(sizeof(double) * (size_t)nxp * ny);
  for (int i = 0; i < nxp; ++i)
    for (int j = 0; j < ny; ++j) aux[i + (long)nxp * j] = interp1(y[j], yp, nyp, fp + i, nxp);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) f[i + (long)nx * j] = interp1(x[i], xp, nxp, aux + (long)nxp * j, 1);
  free(aux);
}

/* interp.F90:6-48 (interp3d): fp(nzp, nxp, nyp); every source level
 * interpolated horizontally (interp2d), then each column in z, in log space
 * when zlog (f = exp(interp1d(z, zp, log(aux)))).  f(nz, nx, ny). */
void tgx_o_interp3d(const double *z, int nz, const double *x, int nx, const double *y, int ny, const double *zp,
                    int nzp, const double *xp, int nxp, const double *yp, int nyp, const double *fp, int zlog,
                    double *f) {
  double *lev = malloc(sizeof(double) * (size_t)nxp * nyp);
  double *out = malloc(sizeof(double) * (size_t)nx * ny);
  double *aux = malloc(sizeof(double) * (size_t)nzp * nx * ny);  /* aux(nzp, nx, ny) */
  for (int k = 0; k < nzp; ++k) {
    for (int j = 0; j < nyp; ++j)
      for (int i = 0; i < nxp; ++i) lev[i + (long)nxp * j] = fp[k + (long)nzp * (i + (long)nxp * j)];
    tgx_o_interp2d(x, nx, y, ny, xp, nxp, yp, nyp, lev, out);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) aux[k + (long)nzp * (i + (long)nx * j)] = out[i + (long)nx * j];
  }
  double *col = malloc(sizeof(double) * (size_t)nzp);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double *a = aux + (long)nzp * (i + (long)nx * j);
      double *o = f + (long)nz * (i + (long)nx * j);
      if (zlog && !MUTI(803)) {  /* planted: z interpolation linear for a log field */
        for (int k = 0; k < nzp; ++k) col[k] = log(a[k]);
        for (int k = 0; k < nz; ++k) o[k] = exp(interp1(z[k], zp, nzp, col, 1));
      } else {
        for (int k = 0; k < nz; ++k) o[k] = interp1(z[k], zp, nzp, a, 1);
      }
    }
  free(col);
  free(aux);
  free(out);
  free(lev);
}

/* interp_fields_ng.F90:17-63: the values at t = ta .. tb (0 <= ta <= tb <= t1)
 * between those at 0 and t1, interp1d on the two points (0, t1): the step
 * blocks 1..t1-1 of the lateral data, and the O+ sub-step blocks 0..t1*nsub.
 * f0, f1 hold n values each; f(:, t - ta). */
void tgx_o_time_interp(const double *f0, const double *f1, long n, int t1, int ta, int tb, double *f) {
  const double tp[2] = {0., (double)t1};
  double fp[2];
  for (long p = 0; p < n; ++p) {
    fp[0] = f0[p];
    fp[1] = f1[p];
    for (int t = ta; t <= tb; ++t) f[p + n * (long)(t - ta)] = interp1((double)t, tp, 2, fp, 1);
  }
}
