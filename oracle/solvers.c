/*
 * Oracle (test infrastructure only): column solvers, small-matrix inverses and
 * the horizontal stencils of the nested grid.  Restated from the reference in
 * its arithmetic order; see tgx_oracle.h for layout and semantics.
 */
#include "tgx_oracle.h"
#include <stdlib.h>
#include <string.h>

/* matutil.F90:57-73 (matdet3) and :110-118 (matinv3 = adj/det); column-major. */
#define A3(r, c) a[((r) - 1) + 3 * ((c) - 1)]
void tgx_o_matinv3(const double *a, double *b) {
  double d = A3(1,1)*A3(2,2)*A3(3,3) - A3(1,1)*A3(2,3)*A3(3,2)
           - A3(1,2)*A3(2,1)*A3(3,3) + A3(1,2)*A3(2,3)*A3(3,1)
           + A3(1,3)*A3(2,1)*A3(3,2) - A3(1,3)*A3(2,2)*A3(3,1);
  double adj[9];
  /* matutil.F90:59-73 (matadj3), B(r,c) stored adj[(r-1)+3*(c-1)] */
  adj[0] =  (A3(2,2)*A3(3,3) - A3(2,3)*A3(3,2));
  adj[1] = -(A3(2,1)*A3(3,3) - A3(2,3)*A3(3,1));
  adj[2] =  (A3(2,1)*A3(3,2) - A3(2,2)*A3(3,1));
  adj[3] = -(A3(1,2)*A3(3,3) - A3(1,3)*A3(3,2));
  adj[4] =  (A3(1,1)*A3(3,3) - A3(1,3)*A3(3,1));
  adj[5] = -(A3(1,1)*A3(3,2) - A3(1,2)*A3(3,1));
  adj[6] =  (A3(1,2)*A3(2,3) - A3(1,3)*A3(2,2));
  adj[7] = -(A3(1,1)*A3(2,3) - A3(1,3)*A3(2,1));
  adj[8] =  (A3(1,1)*A3(2,2) - A3(1,2)*A3(2,1));
  for (int n = 0; n < 9; ++n) b[n] = adj[n] / d;
}
#undef A3

/* matutil.F90:19-29 (matdet2), :44-55 (matadj2), :100-108 (matinv2). */
void tgx_o_matinv2(const double *a, double *b) {
  double d = a[0] * a[3] - a[2] * a[1];
  b[0] = a[3] / d;
  b[1] = -a[1] / d;
  b[2] = -a[2] / d;
  b[3] = a[0] / d;
}

/* lbc.F:527-594 (bndcmp): b = (e/dz - f/2)^-1 (e/dz + f/2), fb = (e/dz - f/2)^-1 g. */
void tgx_o_bndcmp(double dz, double b[3][3], double fb[3]) {
  const double pshelb = 0.1154E-5, alfa = 0.22 + 0.014;
  double e[9] = {0}, f[9] = {0}, wm1[9], wm2[9], wm3[9];
  double g[3] = {-alfa, -0.0, -pshelb};
  e[1 + 3 * 1] = 1.;
  f[0 + 3 * 0] = 1.;
  f[0 + 3 * 1] = 1.;
  f[1 + 3 * 1] = -1.;
  f[2 + 3 * 2] = 1.;
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) {
      wm1[l + 3 * m] = e[l + 3 * m] / dz - f[l + 3 * m] / 2.;
      wm2[l + 3 * m] = e[l + 3 * m] / dz + f[l + 3 * m] / 2.;
    }
  tgx_o_matinv3(wm1, wm3);
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m)
      b[l][m] = wm3[l + 0] * wm2[0 + 3 * m] + wm3[l + 3] * wm2[1 + 3 * m] +
                wm3[l + 6] * wm2[2 + 3 * m];
  for (int l = 0; l < 3; ++l)
    fb[l] = wm3[l + 0] * g[0] + wm3[l + 3] * g[1] + wm3[l + 6] * g[2];
}

/* trsolv.F:31-101.  a..x are (lev0:lev1, lon0:lon1); rows k1..k2 of columns
 * lonbeg..lonend are solved; periodic columns are skipped when lon0==1 or
 * lon1==lonmax (trsolv.F:32-35).  w1 starts at row k1 (the reference sets
 * w1(lev0); every caller passes k1=lev0, trsolv.F:39). */
void tgx_o_trsolv(const double *a, const double *b, const double *c, const double *f,
                  double *x, int lev0, int lev1, int k1, int k2, int lon0, int lon1,
                  int lonmax) {
  int nlev = lev1 - lev0 + 1;
  int lonbeg = lon0, lonend = lon1;
  if (lon0 == 1) lonbeg = 3;
  if (lon1 == lonmax) lonend = lon1 - 2;
  double *w1 = malloc(sizeof(double) * nlev), *w2 = malloc(sizeof(double) * nlev),
         *w3 = malloc(sizeof(double) * nlev);
  for (int i = lonbeg; i <= lonend; ++i) {
    size_t o = (size_t)(i - lon0) * nlev;
    const double *A = a + o, *B = b + o, *C = c + o, *F = f + o;
    double *X = x + o;
#define R(k) ((k) - lev0)
    w1[R(k1)] = B[R(k1)];
    for (int k = k1 + 1; k <= k2; ++k) {
      w2[R(k - 1)] = C[R(k - 1)] / w1[R(k - 1)];
      w1[R(k)] = A[R(k)] * w2[R(k - 1)];
      w1[R(k)] = B[R(k)] - w1[R(k)];
    }
    w3[R(k1)] = F[R(k1)] / w1[R(k1)];
    for (int k = k1 + 1; k <= k2; ++k) {
      w3[R(k)] = A[R(k)] * w3[R(k - 1)];
      w3[R(k)] = F[R(k)] - w3[R(k)];
      w3[R(k)] = w3[R(k)] / w1[R(k)];
    }
    X[R(k2)] = w3[R(k2)];
    for (int kk = k1 + 1; kk <= k2; ++kk) {
      int k = k1 + k2 - kk;
      X[R(k)] = w2[R(k)] * X[R(k + 1)];
      X[R(k)] = w3[R(k)] - X[R(k)];
    }
#undef R
  }
  free(w1); free(w2); free(w3);
}

/* trsolv_ng.F90:14-26.  Every column (halos included), rows 1..k2. */
void tgx_o_trsolv_ng(const double *a, const double *b, const double *c, const double *f,
                     double *x, int nk, int ncol, int k2) {
  double *w1 = malloc(sizeof(double) * nk), *w2 = malloc(sizeof(double) * nk),
         *w3 = malloc(sizeof(double) * nk);
  for (int i = 0; i < ncol; ++i) {
    size_t o = (size_t)i * nk;
    const double *A = a + o, *B = b + o, *C = c + o, *F = f + o;
    double *X = x + o;
    w1[0] = B[0];
    for (int k = 1; k < k2; ++k) {
      w2[k - 1] = C[k - 1] / w1[k - 1];
      w1[k] = B[k] - A[k] * w2[k - 1];
    }
    w3[0] = F[0] / w1[0];
    for (int k = 1; k < k2; ++k) w3[k] = (F[k] - A[k] * w3[k - 1]) / w1[k];
    X[k2 - 1] = w3[k2 - 1];
    for (int k = k2 - 2; k >= 0; --k) X[k] = w3[k] - w2[k] * X[k + 1];
  }
  free(w1); free(w2); free(w3);
}

/* One 2x2 block-Thomas column in the operation order of duv.F:830-906 and
 * duv_ng.F90:218-248 (they agree term by term).  Rows 0..nrow-1. */
static void blk2_column(const double *a, const double *b, const double *c, const double *f,
                        double *x, int nrow, double *beta, double *gam, double *y) {
#define M(arr, r, cc, k) arr[((r) - 1) + 2 * ((cc) - 1) + 4 * (k)]
#define V(arr, r, k) arr[((r) - 1) + 2 * (k)]
  double det = M(b, 1, 1, 0) * M(b, 2, 2, 0) - M(b, 1, 2, 0) * M(b, 2, 1, 0);
  M(beta, 1, 1, 0) = M(b, 2, 2, 0) / det;
  M(beta, 1, 2, 0) = -M(b, 1, 2, 0) / det;
  M(beta, 2, 1, 0) = -M(b, 2, 1, 0) / det;
  M(beta, 2, 2, 0) = M(b, 1, 1, 0) / det;
  V(y, 1, 0) = M(beta, 1, 1, 0) * V(f, 1, 0) + M(beta, 1, 2, 0) * V(f, 2, 0);
  V(y, 2, 0) = M(beta, 2, 1, 0) * V(f, 1, 0) + M(beta, 2, 2, 0) * V(f, 2, 0);
  for (int k = 1; k < nrow; ++k) {
    M(gam, 1, 1, k - 1) = M(beta, 1, 1, k - 1) * M(c, 1, 1, k - 1) + M(beta, 1, 2, k - 1) * M(c, 2, 1, k - 1);
    M(gam, 1, 2, k - 1) = M(beta, 1, 1, k - 1) * M(c, 1, 2, k - 1) + M(beta, 1, 2, k - 1) * M(c, 2, 2, k - 1);
    M(gam, 2, 1, k - 1) = M(beta, 2, 1, k - 1) * M(c, 1, 1, k - 1) + M(beta, 2, 2, k - 1) * M(c, 2, 1, k - 1);
    M(gam, 2, 2, k - 1) = M(beta, 2, 1, k - 1) * M(c, 1, 2, k - 1) + M(beta, 2, 2, k - 1) * M(c, 2, 2, k - 1);
    M(gam, 1, 1, k) = M(b, 1, 1, k) - M(a, 1, 1, k) * M(gam, 1, 1, k - 1) - M(a, 1, 2, k) * M(gam, 2, 1, k - 1);
    M(gam, 1, 2, k) = M(b, 1, 2, k) - M(a, 1, 1, k) * M(gam, 1, 2, k - 1) - M(a, 1, 2, k) * M(gam, 2, 2, k - 1);
    M(gam, 2, 1, k) = M(b, 2, 1, k) - M(a, 2, 1, k) * M(gam, 1, 1, k - 1) - M(a, 2, 2, k) * M(gam, 2, 1, k - 1);
    M(gam, 2, 2, k) = M(b, 2, 2, k) - M(a, 2, 1, k) * M(gam, 1, 2, k - 1) - M(a, 2, 2, k) * M(gam, 2, 2, k - 1);
    double dk = M(gam, 1, 1, k) * M(gam, 2, 2, k) - M(gam, 1, 2, k) * M(gam, 2, 1, k);
    M(beta, 1, 1, k) = M(gam, 2, 2, k) / dk;
    M(beta, 1, 2, k) = -M(gam, 1, 2, k) / dk;
    M(beta, 2, 1, k) = -M(gam, 2, 1, k) / dk;
    M(beta, 2, 2, k) = M(gam, 1, 1, k) / dk;
    double x1 = V(f, 1, k) - M(a, 1, 1, k) * V(y, 1, k - 1) - M(a, 1, 2, k) * V(y, 2, k - 1);
    double x2 = V(f, 2, k) - M(a, 2, 1, k) * V(y, 1, k - 1) - M(a, 2, 2, k) * V(y, 2, k - 1);
    V(y, 1, k) = M(beta, 1, 1, k) * x1 + M(beta, 1, 2, k) * x2;
    V(y, 2, k) = M(beta, 2, 1, k) * x1 + M(beta, 2, 2, k) * x2;
  }
  V(x, 1, nrow - 1) = V(y, 1, nrow - 1);
  V(x, 2, nrow - 1) = V(y, 2, nrow - 1);
  for (int k = nrow - 2; k >= 0; --k) {
    V(x, 1, k) = V(y, 1, k) - M(gam, 1, 1, k) * V(x, 1, k + 1) - M(gam, 1, 2, k) * V(x, 2, k + 1);
    V(x, 2, k) = V(y, 2, k) - M(gam, 2, 1, k) * V(x, 1, k + 1) - M(gam, 2, 2, k) * V(x, 2, k + 1);
  }
#undef M
#undef V
}

/* duv.F:801-908 (blktri): a,b,c (2,2,lev0:lev1,lon0:lon1), f,x (2,lev0:lev1,lon0:lon1);
 * rows lev0..lev1-1 of columns lonbeg..lonend (duv.F:825-828, lonmax = nlonp4). */
void tgx_o_blktri(const double *a, const double *b, const double *c, const double *f,
                  double *x, int lev0, int lev1, int lon0, int lon1, int lonmax) {
  int nlev = lev1 - lev0 + 1;
  int lonbeg = lon0, lonend = lon1;
  if (lon0 == 1) lonbeg = 3;
  if (lon1 == lonmax) lonend = lonmax - 2;
  double *beta = malloc(sizeof(double) * 4 * nlev), *gam = malloc(sizeof(double) * 4 * nlev),
         *y = malloc(sizeof(double) * 2 * nlev);
  for (int i = lonbeg; i <= lonend; ++i) {
    size_t o = (size_t)(i - lon0) * nlev;
    blk2_column(a + 4 * o, b + 4 * o, c + 4 * o, f + 2 * o, x + 2 * o, nlev - 1, beta, gam, y);
  }
  free(beta); free(gam); free(y);
}

/* duv_ng.F90:218-248: rows 1..nk-1 of every column. */
void tgx_o_blktri_ng(const double *pp, const double *qq, const double *rr,
                     const double *ss, double *xx, int nk, int ncol) {
  double *beta = malloc(sizeof(double) * 4 * nk), *gam = malloc(sizeof(double) * 4 * nk),
         *y = malloc(sizeof(double) * 2 * nk);
  for (int i = 0; i < ncol; ++i) {
    size_t o = (size_t)i * nk;
    blk2_column(pp + 4 * o, qq + 4 * o, rr + 4 * o, ss + 2 * o, xx + 2 * o, nk - 1, beta, gam, y);
  }
  free(beta); free(gam); free(y);
}

/* The 3x3 block elimination of comp_ng.F90:281-326 (= comp.F:663-806) on
 * pre-assembled blocks: p,q,r (3,3,nk,ncol), f,x (3,nk,ncol), rows 1..k2.
 * gama(1) = zz(1) = 0, so p(1) drops out exactly as in comp. */
void tgx_o_blktri3(const double *p, const double *q, const double *r, const double *f,
                   double *x, int nk, int ncol, int k2) {
  double *gama = malloc(sizeof(double) * 9 * (nk + 1)), *zz = malloc(sizeof(double) * 3 * (nk + 1));
  for (int col = 0; col < ncol; ++col) {
    const double *P = p + (size_t)9 * nk * col, *Q = q + (size_t)9 * nk * col,
                 *Rr = r + (size_t)9 * nk * col, *F = f + (size_t)3 * nk * col;
    double *X = x + (size_t)3 * nk * col;
    for (int n = 0; n < 9; ++n) gama[n] = 0.;
    for (int n = 0; n < 3; ++n) zz[n] = 0.;
    for (int k = 0; k < k2; ++k) {
      double qk[9], wkm1[9], wkv1[3];
      const double *pk = P + 9 * k, *rk = Rr + 9 * k;
      for (int n = 0; n < 9; ++n) qk[n] = Q[9 * k + n];
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp)
            qk[isp + 3 * m] = qk[isp + 3 * m] - pk[isp + 3 * n] * gama[9 * k + n + 3 * m];
      tgx_o_matinv3(qk, wkm1);
      for (int isp = 0; isp < 3; ++isp) wkv1[isp] = F[3 * k + isp];
      for (int n = 0; n < 9; ++n) gama[9 * (k + 1) + n] = 0.;
      for (int m = 0; m < 3; ++m) {
        for (int isp = 0; isp < 3; ++isp) wkv1[isp] = wkv1[isp] - pk[isp + 3 * m] * zz[3 * k + m];
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp)
            gama[9 * (k + 1) + isp + 3 * m] = gama[9 * (k + 1) + isp + 3 * m] + wkm1[isp + 3 * n] * rk[n + 3 * m];
      }
      for (int isp = 0; isp < 3; ++isp) zz[3 * (k + 1) + isp] = 0.;
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp)
          zz[3 * (k + 1) + isp] = zz[3 * (k + 1) + isp] + wkm1[isp + 3 * m] * wkv1[m];
    }
    double u_next[3] = {0., 0., 0.};
    for (int k = k2 - 1; k >= 0; --k) {
      double u[3];
      for (int isp = 0; isp < 3; ++isp) {
        u[isp] = zz[3 * (k + 1) + isp];
        for (int m = 0; m < 3; ++m) u[isp] = u[isp] - gama[9 * (k + 1) + isp + 3 * m] * u_next[m];
      }
      for (int isp = 0; isp < 3; ++isp) { X[3 * k + isp] = u[isp]; u_next[isp] = u[isp]; }
    }
  }
  free(gama); free(zz);
}

/* ---- horizontal stencils -------------------------------------------- */
#define F3(arr, k, i, j) arr[(size_t)(k) + (size_t)nk * ((size_t)(i) + (size_t)nlond * (size_t)(j))]

/* dffm_ng.F90:62-112 (df_3d), single task so sync_var3d_* is a no-op. */
void tgx_o_df3d(const double *f, double *df, int order, int nk, int nlond, int nlatd) {
  int L0 = 0, L1 = nlond - 1, J0 = 0, J1 = nlatd - 1;
  if (order == 1 || order == 2) {
    for (int j = 0; j < nlatd; ++j)
      for (int k = 0; k < nk; ++k) {
        F3(df, k, L0, j) = F3(f, k, L0 + 1, j) - F3(f, k, L0, j);
        F3(df, k, L1, j) = F3(f, k, L1, j) - F3(f, k, L1 - 1, j);
        if (order == 2) {
          F3(df, k, L0 + 1, j) = (F3(f, k, L0 + 2, j) - F3(f, k, L0, j)) / 2.;
          F3(df, k, L1 - 1, j) = (F3(f, k, L1, j) - F3(f, k, L1 - 2, j)) / 2.;
        }
        for (int i = L0 + order; i <= L1 - order; ++i)
          F3(df, k, i, j) = (F3(f, k, i + order, j) - F3(f, k, i - order, j)) / (double)(order * 2);
      }
  } else {
    int o = -order;
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        F3(df, k, i, J0) = F3(f, k, i, J0 + 1) - F3(f, k, i, J0);
        F3(df, k, i, J1) = F3(f, k, i, J1) - F3(f, k, i, J1 - 1);
        if (o == 2) {
          F3(df, k, i, J0 + 1) = (F3(f, k, i, J0 + 2) - F3(f, k, i, J0)) / 2.;
          F3(df, k, i, J1 - 1) = (F3(f, k, i, J1) - F3(f, k, i, J1 - 2)) / 2.;
        }
        for (int j = J0 + o; j <= J1 - o; ++j)
          F3(df, k, i, j) = (F3(f, k, i, j + o) - F3(f, k, i, j - o)) / (double)(o * 2);
      }
  }
}

/* dffm_ng.F90:167-217 (fm_3d). */
void tgx_o_fm3d(const double *f, double *fm, int order, int nk, int nlond, int nlatd) {
  int L0 = 0, L1 = nlond - 1, J0 = 0, J1 = nlatd - 1;
  if (order == 1 || order == 2) {
    for (int j = 0; j < nlatd; ++j)
      for (int k = 0; k < nk; ++k) {
        F3(fm, k, L0, j) = F3(f, k, L0, j);
        F3(fm, k, L1, j) = F3(f, k, L1, j);
        if (order == 2) {
          F3(fm, k, L0 + 1, j) = (F3(f, k, L0 + 2, j) + F3(f, k, L0, j)) / 2.;
          F3(fm, k, L1 - 1, j) = (F3(f, k, L1, j) + F3(f, k, L1 - 2, j)) / 2.;
        }
        for (int i = L0 + order; i <= L1 - order; ++i)
          F3(fm, k, i, j) = (F3(f, k, i + order, j) + F3(f, k, i - order, j)) / 2.;
      }
  } else {
    int o = -order;
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        F3(fm, k, i, J0) = F3(f, k, i, J0);
        F3(fm, k, i, J1) = F3(f, k, i, J1);
        if (o == 2) {
          F3(fm, k, i, J0 + 1) = (F3(f, k, i, J0 + 2) + F3(f, k, i, J0)) / 2.;
          F3(fm, k, i, J1 - 1) = (F3(f, k, i, J1) + F3(f, k, i, J1 - 2)) / 2.;
        }
        for (int j = J0 + o; j <= J1 - o; ++j)
          F3(fm, k, i, j) = (F3(f, k, i, j - o) + F3(f, k, i, j + o)) / 2.;
      }
  }
}

/* dffm_ng.F90:8-60 (df_2d, order 1 or -1 as used by oplus divb). */
void tgx_o_df2d(const double *f, double *df, int order, int nlond, int nlatd) {
  const int nk = 1;
  tgx_o_df3d(f, df, order, nk, nlond, nlatd);
}

/* advec_ng.F90:1-52. */
void tgx_o_advec(const tgx_ogrid *g, const double *f, const double *un, const double *vn,
                 double *hadvec) {
  int nk = g->nk, nlond = g->nlon + 4, nlatd = g->nlat + 4, nx = g->nlon, ny = g->nlat;
  size_t n = (size_t)nk * nlond * nlatd;
  const double re_inv = 1. / 6.37122e8;
  double *ub1 = malloc(8 * n), *ub2 = malloc(8 * n), *dx1 = malloc(8 * n), *dx2 = malloc(8 * n);
  double *vb1 = malloc(8 * n), *vb2 = malloc(8 * n), *dy1 = malloc(8 * n), *dy2 = malloc(8 * n);
  tgx_o_fm3d(un, ub1, 1, nk, nlond, nlatd);
  tgx_o_fm3d(un, ub2, 2, nk, nlond, nlatd);
  tgx_o_df3d(f, dx1, 1, nk, nlond, nlatd);
  tgx_o_df3d(f, dx2, 2, nk, nlond, nlatd);
  tgx_o_fm3d(vn, vb1, -1, nk, nlond, nlatd);
  tgx_o_fm3d(vn, vb2, -2, nk, nlond, nlatd);
  tgx_o_df3d(f, dy1, -1, nk, nlond, nlatd);
  tgx_o_df3d(f, dy2, -2, nk, nlond, nlatd);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        size_t p = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * j);
        double wk1 = dx1[p] * 2. * 2. / 3. * ub1[p] - dx2[p] * 4. / 12. * ub2[p];
        /* Fortran lon -1:0 -> i 0:1, nx+1:nx+2 -> i nx+2:nx+3 */
        if (i <= 1 || i >= nx + 2) wk1 = dx1[p] * ub1[p];
        wk1 = wk1 * g->racs[j];
        double wk2 = dy1[p] * 2. * 2. / 3. * vb1[p] - dy2[p] * 4. / 12. * vb2[p];
        if (j <= 1 || j >= ny + 2) wk2 = dy1[p] * vb1[p];
        hadvec[p] = wk1 / g->dlamda + wk2 * re_inv / g->dphi;
      }
  free(ub1); free(ub2); free(dx1); free(dx2); free(vb1); free(vb2); free(dy1); free(dy2);
}

/* smooth_ng.F90:1-42: lat pass (5-point, 3-point next to the edge, copy at
 * the edge), then lon pass on the result. */
void tgx_o_smooth(const tgx_ogrid *g, const double *fin, double *fout, double shapiro) {
  int nk = g->nk, nlond = g->nlon + 4, nlatd = g->nlat + 4, nx = g->nlon, ny = g->nlat;
  size_t n = (size_t)nk * nlond * nlatd;
  double *fm1 = malloc(8 * n), *fm2 = malloc(8 * n), *ftmp = malloc(8 * n);
  tgx_o_fm3d(fin, fm1, -1, nk, nlond, nlatd);
  tgx_o_fm3d(fin, fm2, -2, nk, nlond, nlatd);
  for (size_t p = 0; p < n; ++p) ftmp[p] = fin[p] - shapiro * (2. * fm2[p] - 8. * fm1[p] + 6. * fin[p]);
  for (int i = 0; i < nlond; ++i)
    for (int k = 0; k < nk; ++k) {
      /* lat 0 -> j=1, lat -1 -> j=0, ny+1 -> j=ny+2, ny+2 -> j=ny+3 */
      size_t p1 = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * 1);
      size_t p0 = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * 0);
      ftmp[p1] = fin[p1] - shapiro * (-8. * fm1[p1] + 8. * fin[p1]);
      ftmp[p0] = fin[p0];
      size_t q1 = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * (ny + 2));
      size_t q0 = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * (ny + 3));
      ftmp[q1] = fin[q1] - shapiro * (-8. * fm1[q1] + 8. * fin[q1]);
      ftmp[q0] = fin[q0];
    }
  tgx_o_fm3d(ftmp, fm1, 1, nk, nlond, nlatd);
  tgx_o_fm3d(ftmp, fm2, 2, nk, nlond, nlatd);
  for (size_t p = 0; p < n; ++p) fout[p] = ftmp[p] - shapiro * (2. * fm2[p] - 8. * fm1[p] + 6. * ftmp[p]);
  for (int j = 0; j < nlatd; ++j)
    for (int k = 0; k < nk; ++k) {
      size_t p1 = (size_t)k + (size_t)nk * (1 + (size_t)nlond * j);
      size_t p0 = (size_t)k + (size_t)nk * (0 + (size_t)nlond * j);
      fout[p1] = ftmp[p1] - shapiro * (-8. * fm1[p1] + 8. * ftmp[p1]);
      fout[p0] = ftmp[p0];
      size_t q1 = (size_t)k + (size_t)nk * ((size_t)(nx + 2) + (size_t)nlond * j);
      size_t q0 = (size_t)k + (size_t)nk * ((size_t)(nx + 3) + (size_t)nlond * j);
      fout[q1] = ftmp[q1] - shapiro * (-8. * fm1[q1] + 8. * ftmp[q1]);
      fout[q0] = ftmp[q0];
    }
  free(fm1); free(fm2); free(ftmp);
}

/* div_ng.F90:1-43. */
void tgx_o_div(const tgx_ogrid *g, const double *f, double *df, int dir) {
  int nk = g->nk, nlond = g->nlon + 4, nlatd = g->nlat + 4, nx = g->nlon, ny = g->nlat;
  size_t n = (size_t)nk * nlond * nlatd;
  double *d1 = malloc(8 * n), *d2 = malloc(8 * n);
  tgx_o_df3d(f, d1, dir == 1 ? 1 : -1, nk, nlond, nlatd);
  tgx_o_df3d(f, d2, dir == 1 ? 2 : -2, nk, nlond, nlatd);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        size_t p = (size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * j);
        double v = d1[p] * 2. * 2. / 3. - d2[p] * 4. / 12.;
        if (dir == 1) { if (i <= 1 || i >= nx + 2) v = d1[p]; }
        else { if (j <= 1 || j >= ny + 2) v = d1[p]; }
        df[p] = v;
      }
  free(d1); free(d2);
}
