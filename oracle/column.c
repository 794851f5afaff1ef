/*
 * Oracle (test infrastructure only): the _ng column-step routines
 * dt_ng, duv_ng, minor_ng (+ comp_{no,n4s,n2d,ar}_ng), oplus_ng, comp_o2o_ng
 * and comp_ng, restated in the reference's evaluation order.  Diagnostics
 * written with addfld are not reproduced (they feed no output).
 */
#include "tgx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cons.F:25-46 and :98-115; set in init_cons (cons.F:271-330). */
static const double RE = 6.37122e8, AVO = 6.023e23, P0 = 5.0e-4, GASK = 8.314e7;
static const double BRN2D = 0.6, DTSMOOTH = 0.95, GRAV = 870.0;
#define TSURPLUS (5.11 * 1.602e-12)
#define DTSMOOTH_DIV2 (0.5 * (1. - DTSMOOTH))
#define DZGRAV (GRAV / GASK)
#define RE_INV (1. / RE)
static const double RMO2 = 32., RMO1 = 16., RMHE = 4., RMN2 = 28., RMN4S = 14., RMNO = 30.;
static const double RMAR = 40., RMOP = 16., RMN2D = 14.;
#define RIO2 (1. / RMO2)
#define RIO1 (1. / RMO1)
#define RIHE (1. / RMHE)
#define RIN2 (1. / RMN2)
#define RIN4S (1. / RMN4S)
#define RINO (1. / RMNO)
#define RIN2D (1. / RMN2D)
/* chemrates.F:29-50 (temperature-independent rates). */
static const double RK4 = 1.0E-10, RK5 = 4.4E-10, RK6 = 4.0E-10, RK7 = 2.0E-10, RK8 = 1.0E-12;
static const double RK9 = 6.0E-11, RK10 = 1.3E-10, RK16 = 4.8E-10, RK17 = 1.0E-10, RK18 = 4.0E-10;
static const double RK21 = 0.047, RK22 = 0.171, RK23 = 8.E-10, RK24 = 5.0E-12, RK26 = 7.E-10;
static const double RK27 = 7.7E-5, BETA2 = 5.0E-12, BETA4 = 7.0E-13, BETA6 = 7.0E-11, BETA7 = 1.06E-5;
/* minor_ng.F90:25-28 and comp_ng.F90:22-25 */
static const double T00 = 273., TAU = 1.86e+3;
static const double SS[3] = {1.710, 1.749, 1.718};
/* phi(3,4) column-major: PHI[r][c] = phi(r+1,c+1) */
static const double PHI[3][4] = {{0., 1.35, 2.16, 1.11}, {0.673, 0., 1.616, 0.769}, {0.270, 0.404, 0., 0.322}};

#define DIMS                                    \
  const int nk = g->nk;                         \
  const int nlond = g->nlon + 4, nlatd = g->nlat + 4; \
  const int nx = g->nlon, ny = g->nlat;         \
  const size_t N3 = (size_t)nk * nlond * nlatd; \
  (void)nx; (void)ny; (void)N3
#define IX(k, i, j) ((size_t)(k) + (size_t)nk * ((size_t)(i) + (size_t)nlond * (size_t)(j)))
#define I2(i, j) ((size_t)(i) + (size_t)nlond * (size_t)(j))
#define ALLOC3(v) double *v = (double *)malloc(sizeof(double) * N3)
#define COLBUF(v) double *v = (double *)malloc(sizeof(double) * (nk + 2))

static double pi_(void) { return 4.0 * atan(1.0); }

/* ---- test hooks --------------------------------------------------------
 * Row capture: when a buffer is set, each routine writes the rows it hands to
 * its column solve, after the boundary rows are applied, in the layout of the
 * library's capture mode: cap[(plane*ncol + col)*nk + k], col = j*(nlon+4)+i,
 * planes 0-3 dt a,b,c,f; 4-7 O+ (last sub-step); 8+4s..11+4s minor species s
 * (0 NO, 1 N4S, 2 N2D, 3 Ar); 24-35 duv A, B, C (2x2 row-major), 36-37 F;
 * 38-64 comp A, B, C (3x3 row-major), 65-67 F.
 * Planted mistakes: tgx_o_set_mutation(id) with id != 0 makes one site of one
 * routine wrong on purpose (each site below is marked MUT(id)), so that the
 * tests can show that their checks fail on a wrong oracle. */
static double *g_cap = NULL;
static int g_mut = 0;
void tgx_o_set_capture(double *cap) { g_cap = cap; }
void tgx_o_set_mutation(int id) { g_mut = id; }
int tgx_o_mutation_id(void) { return g_mut; }
#define MUT(id) (g_mut == (id))
#define CAPW(pl, col, k, v) (g_cap[((size_t)(pl) * (size_t)nlond * (size_t)nlatd + (size_t)(col)) * (size_t)nk + (size_t)(k)] = (v))

/* Dirichlet lateral-boundary hook used by every _ng routine (e.g.
 * dt_ng.F90:201-206): ghost column lon=-1 (i=0) and lon=nlon+2 (i=nlond-1)
 * from lon_b(nk,2,nlatd), then ghost rows lat=-1, nlat+2 from lat_b(nk,nlond,2). */
static void lateral(double *x, const double *lonb, const double *latb, int nk, int nlond, int nlatd) {
  if (lonb)
    for (int j = 0; j < nlatd; ++j)
      for (int k = 0; k < nk; ++k) {
        x[IX(k, 0, j)] = lonb[(size_t)k + (size_t)nk * (0 + 2 * (size_t)j)];
        x[IX(k, nlond - 1, j)] = lonb[(size_t)k + (size_t)nk * (1 + 2 * (size_t)j)];
      }
  if (latb)
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        x[IX(k, i, 0)] = latb[(size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * 0)];
        x[IX(k, i, nlatd - 1)] = latb[(size_t)k + (size_t)nk * ((size_t)i + (size_t)nlond * 1)];
      }
}

/* ---- global grid ---------------------------------------------------------
 * params.F:38 */
#define SPVAL 1.e36

/* mp_periodic_f3d (mpi.F:2932-3056) on the interior latitudes: the periodic
 * points from the interior longitudes they repeat (lon 1, 2 = nlon+1, nlon+2;
 * nlon+3, nlon+4 = 3, 4, i.e. i = lon - 1). */
static void glb_periodic(const tgx_ogrid *g, double *f) {
  DIMS;
  const int src0 = MUT(803) ? nx - 1 : nx;
  for (int j = 2; j < nlatd - 2; ++j)
    for (int k = 0; k < nk; ++k) {
      f[IX(k, 0, j)] = f[IX(k, src0, j)];
      f[IX(k, 1, j)] = f[IX(k, nx + 1, j)];
      f[IX(k, nx + 2, j)] = f[IX(k, 2, j)];
      f[IX(k, nx + 3, j)] = f[IX(k, 3, j)];
    }
}

/* filter_tn (dt.F:431-511) and its siblings filter_uv (duv.F:910-994),
 * filter_minor (minor.F:621-702), filter_op (oplus.F:1210-1284) and
 * filter_o2o (comp.F:1065-1145): ringfilter.F on every interior latitude the
 * chunk table filters, levels 0..nlev-1 of the field. */
static void glb_filter(const tgx_ogrid *g, double *f, int nlev, int order, int oplus) {
  DIMS;
  double *row = malloc(sizeof(double) * (size_t)nlond * (size_t)nlev);
  for (int lat = 1; lat <= ny; ++lat) {
    int nch = tgx_o_ringfilter_chunks(lat, ny, g->nlat_filter, g->chunk);
    if (nch == 0) continue;
    const int j = lat + 1;
    for (int l = 0; l < nlev; ++l)
      for (int i = 0; i < nlond; ++i) row[(size_t)l * nlond + i] = f[IX(l, i, j)];
    tgx_o_ringfilter_row(row, nx, nlev, order, oplus, nch, g->dlamda);
    for (int l = 0; l < nlev; ++l)
      for (int i = 0; i < nlond; ++i) f[IX(l, i, j)] = row[(size_t)l * nlond + i];
  }
  free(row);
}

/* minor_ng.F90:145-154 / comp_ng.F90:113-122 */
static double dfactor_of(const tgx_ogrid *g, int j) {
  const double hor = .25, pi = pi_(), dtr = pi / 180.;
  if (!(g->difhor > 0)) return 1.;
  double rlat = g->glat[j] * dtr;
  if (fabs(rlat) >= pi / 4.5) return hor + 1.;
  return hor + .5 * (1. + sin(pi * (fabs(rlat) - pi / 9.) / (pi / 4.5)));
}

/* ===================================================================== */
/* dt_ng.F90:1-215 */
void tgx_o_dt_ng(const tgx_ogrid *g, const tgx_ofields *F, double *tlbc, double *tlbc_nm,
                 double *tn_upd, double *tn_nm_upd) {
  DIMS;
  const double dz = g->dz, dtx2inv = g->dtx2inv;
  ALLOC3(horadv);
  ALLOC3(tnsmooth);
  if (MUT(107)) tgx_o_advec(g, F->tn, F->vn, F->un, horadv);
  else tgx_o_advec(g, F->tn, F->un, F->vn, horadv);
  tgx_o_smooth(g, F->tn_nm, tnsmooth, g->shapiro);
  COLBUF(veradv); COLBUF(cpi); COLBUF(kmi); COLBUF(wi); COLBUF(cptn); COLBUF(dudz); COLBUF(dvdz);
  COLBUF(qsolar); COLBUF(tni); COLBUF(gg); COLBUF(ff); COLBUF(p); COLBUF(q); COLBUF(r); COLBUF(rhs);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), o2 = I2(i, j);
      const double *tn = F->tn + o, *w = F->w_upd + o, *cp = F->cp + o, *km = F->km + o;
      const double *un = F->un + o, *vn = F->vn + o, *qtotal = F->qtotal + o;
      double t_lbc = F->t_lbc[o2];
      /* :57-65 */
      veradv[0] = (tn[0] - t_lbc) * w[0] * 2.;
      veradv[nk - 1] = 0.;
      for (int k = 1; k < nk - 1; ++k) veradv[k] = (tn[k] - tn[k - 1]) * w[k];
      for (int k = 0; k < nk - 1; ++k) veradv[k] = (veradv[k] + veradv[k + 1]) * .5;
      for (int k = 0; k < nk; ++k) veradv[k] = veradv[k] / dz;
      /* :71-78 */
      for (int k = 0; k < nk - 1; ++k) {
        cpi[k] = .5 * (cp[k] + cp[k + 1]);
        kmi[k] = .5 * (km[k] + km[k + 1]);
        wi[k] = .5 * (w[k] + w[k + 1]);
      }
      cpi[nk - 1] = 1.5 * cp[nk - 1] - .5 * cp[nk - 2];
      kmi[nk - 1] = 1.5 * km[nk - 1] - .5 * km[nk - 2];
      wi[nk - 1] = 1.5 * w[nk - 1] - .5 * w[nk - 2];
      /* :67, :80-83 */
      for (int k = 0; k < nk; ++k) {
        double advec_tn = horadv[o + k] + veradv[k];
        cptn[k] = cpi[k] * (advec_tn - dtx2inv * tnsmooth[o + k]);
        cptn[k] = cptn[k] * g->expz[k];
      }
      /* :85-94 */
      for (int k = 1; k < nk - 2; ++k) {
        dudz[k] = (un[k + 1] - un[k - 1]) / (2. * dz);
        dvdz[k] = (vn[k + 1] - vn[k - 1]) / (2. * dz);
      }
      dudz[0] = (un[0] + 1. / 3. * un[1] - 4. / 3. * F->ulbc[o2]) / dz;
      dvdz[0] = (vn[0] + 1. / 3. * vn[1] - 4. / 3. * F->vlbc[o2]) / dz;
      dudz[nk - 2] = dudz[nk - 3] / 3.;
      dvdz[nk - 2] = dvdz[nk - 3] / 3.;
      dudz[nk - 1] = dudz[nk - 2] / 3.;
      dvdz[nk - 1] = dvdz[nk - 2] / 3.;
      /* :96-99 */
      for (int k = 0; k < nk - 1; ++k) qsolar[k] = .5 * (qtotal[k] + qtotal[k + 1]);
      qsolar[nk - 1] = 1.5 * qtotal[nk - 1] - .5 * qtotal[nk - 2];
      /* :101-113 */
      for (int k = 0; k < nk; ++k) {
        double t = F->xnmbar[o + k] * F->o1[o + k] * RIO1;
        double qrecomb = TSURPLUS * F->rkm12[o + k] * (t * t) * AVO / F->mbar[o + k];
        double qm = GRAV * kmi[k] / (P0 * F->scht[o + k]) * (dudz[k] * dudz[k] + dvdz[k] * dvdz[k]);
        qm = qm / g->expz[k];
        double total_heat = qsolar[k] + F->hdt[o + k] + qrecomb + F->qji_tn[o + k] + qm;
        double total_heat_a = total_heat * g->expz[k];
        cptn[k] = cptn[k] - total_heat_a;
      }
      /* :115-119 */
      for (int k = 1; k < nk - 1; ++k) tni[k] = MUT(106) ? .5 * (tn[k] + tn[k + 1]) : .5 * (tn[k - 1] + tn[k]);
      tni[0] = tlbc[o2];
      tni[nk - 1] = tn[nk - 2];
      /* :121-139 */
      for (int k = 0; k < nk; ++k) {
        double h = F->schti[o + k];
        double rho = P0 * (MUT(104) ? g->expzmid : g->expzmid_inv) / (h * GRAV);
        rho = rho * g->expz[k];
        double gpart = h * h * rho * cp[k];
        double gpart_a = gpart * g->dift[k];
        gg[k] = GRAV * (F->kt[o + k] + gpart_a) / (P0 * h * (dz * dz));
        ff[k] = GRAV * GRAV * (h * h) * rho / (tni[k] * P0 * 2. * dz);
        if (!MUT(101)) ff[k] = ff[k] * g->dift[k];
      }
      /* :141-150 */
      for (int k = 0; k < nk - 1; ++k) {
        p[k] = gg[k] - ff[k];
        q[k] = -(gg[k + 1] + (MUT(102) ? ff[k + 1] : -ff[k + 1]) + gg[k] + ff[k]);
        r[k] = gg[k + 1] + ff[k + 1];
      }
      p[nk - 1] = gg[nk - 1] - ff[nk - 1];
      q[nk - 2] = -gg[nk - 2] - ff[nk - 2];
      q[nk - 1] = -gg[nk - 1] - ff[nk - 1];
      r[nk - 2] = 0.;
      r[nk - 1] = 0.;
      /* :152-163 */
      for (int k = 0; k < nk; ++k) {
        rhs[k] = cptn[k];
        double qpart = cpi[k] * (dtx2inv + F->cool_imp[o + k]) + (MUT(103) ? 0. : wi[k] * GASK / F->mbar[o + k]);
        rhs[k] = MUT(105) ? rhs[k] - F->cool_exp[o + k] : rhs[k] + F->cool_exp[o + k];
        double qpart_a = qpart * g->expz[k];
        q[k] = q[k] - qpart_a;
      }
      q[0] = q[0] - p[0];
      rhs[0] = rhs[0] - (MUT(108) ? 1. : 2.) * p[0] * t_lbc;
      p[0] = 0.;
      if (g_cap)
        for (int k = 0; k < nk; ++k) {
          CAPW(0, I2(i, j), k, p[k]); CAPW(1, I2(i, j), k, q[k]);
          CAPW(2, I2(i, j), k, r[k]); CAPW(3, I2(i, j), k, rhs[k]);
        }
      /* :165-167 */
      tgx_o_trsolv_ng(p, q, r, rhs, tn_upd + o, nk, 1, nk - 1);
      if (!g->glb) tn_upd[o + nk - 1] = tn_upd[o + nk - 2] * tn_upd[o + nk - 2] / tn_upd[o + nk - 3];
    }
  if (g->glb) {
    /* dt.F:389-424: filter_tn, the time smoothing below the top level, spval
     * at the top (:407-408), the lower-boundary bookkeeping, the 100 K floor */
    if (!MUT(802)) glb_filter(g, tn_upd, nk - 1, 3, 0);
    for (int j = 0; j < nlatd; ++j)
      for (int i = 0; i < nlond; ++i) {
        size_t o = IX(0, i, j);
        for (int k = 0; k < nk - 1; ++k)
          tn_nm_upd[o + k] = DTSMOOTH_DIV2 * (F->tn_nm[o + k] + tn_upd[o + k]) + DTSMOOTH * F->tn[o + k];
        tn_upd[o + nk - 1] = MUT(806) ? tn_upd[o + nk - 2] : SPVAL;
        tn_nm_upd[o + nk - 1] = SPVAL;
      }
    for (size_t n = 0; n < (size_t)nlond * nlatd; ++n) {
      tlbc_nm[n] = tlbc[n];
      tlbc[n] = F->t_lbc[n];
    }
    for (size_t n = 0; n < N3; ++n) tn_upd[n] = tn_upd[n] > 100. ? tn_upd[n] : 100.;
    glb_periodic(g, tn_upd);
    glb_periodic(g, tn_nm_upd);
    free(horadv); free(tnsmooth); free(veradv); free(cpi); free(kmi); free(wi); free(cptn); free(dudz);
    free(dvdz); free(qsolar); free(tni); free(gg); free(ff); free(p); free(q); free(r); free(rhs);
    return;
  }
  /* :201-213 */
  lateral(tn_upd, F->lonb_tn, F->latb_tn, nk, nlond, nlatd);
  for (size_t n = 0; n < N3; ++n) tn_nm_upd[n] = DTSMOOTH_DIV2 * (F->tn_nm[n] + tn_upd[n]) + DTSMOOTH * F->tn[n];
  for (size_t n = 0; n < (size_t)nlond * nlatd; ++n) {
    tlbc_nm[n] = tlbc[n];
    tlbc[n] = F->t_lbc[n];
  }
  for (size_t n = 0; n < N3; ++n) tn_upd[n] = tn_upd[n] > 100. ? tn_upd[n] : 100.;
  free(horadv); free(tnsmooth); free(veradv); free(cpi); free(kmi); free(wi); free(cptn); free(dudz);
  free(dvdz); free(qsolar); free(tni); free(gg); free(ff); free(p); free(q); free(r); free(rhs);
}

/* ===================================================================== */
/* duv_ng.F90:1-323 */
void tgx_o_duv_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *tn_upd,
                  const double *Fe, const double *Fn, const double *tlbc_nm,
                  double *ulbc, double *vlbc, double *ulbc_nm, double *vlbc_nm,
                  double *un_upd, double *unm_upd, double *vn_upd, double *vnm_upd) {
  DIMS;
  (void)tlbc_nm; /* only feeds tni, which duv_ng never uses (duv_ng.F90:137-141) */
  const double dz = g->dz, dtx2inv = g->dtx2inv, wt = 0.225;
  ALLOC3(adv_u); ALLOC3(adv_v); ALLOC3(ztmp); ALLOC3(zl); ALLOC3(zp); ALLOC3(unms); ALLOC3(vnms);
  tgx_o_advec(g, F->un, F->un, F->vn, adv_u);
  tgx_o_advec(g, F->vn, F->un, F->vn, adv_v);
  /* :73-80 */
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j);
      ztmp[o] = F->z[o];
      for (int k = 0; k < nk - 1; ++k) {
        double tbar = (F->tn[o + k] * (-2.) + F->tn_nm[o + k] + tn_upd[o + k]) * wt + F->tn[o + k];
        double dztbar = dz / DZGRAV * tbar / F->mbar[o + k];
        ztmp[o + k + 1] = dztbar + ztmp[o + k];
      }
    }
  /* :83-95 */
  tgx_o_div(g, ztmp, zl, 1);
  tgx_o_div(g, F->z, zp, -1);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j);
      for (int k = 0; k < nk - 1; ++k) {
        zl[o + k] = (zl[o + k] + zl[o + k + 1]) * .5;
        zp[o + k] = (zp[o + k] + zp[o + k + 1]) * .5;
      }
      for (int k = 0; k < nk; ++k) {
        if (!MUT(205)) zl[o + k] = zl[o + k] * g->racs[j];
        zl[o + k] = zl[o + k] * GRAV / g->dlamda;
        zp[o + k] = zp[o + k] * GRAV * RE_INV / g->dphi;
      }
    }
  /* :97-98 */
  tgx_o_smooth(g, MUT(207) ? F->vn_nm : F->un_nm, unms, g->shapiro);
  tgx_o_smooth(g, F->vn_nm, vnms, g->shapiro);
  int nk4 = 4 * nk, nk2 = 2 * nk;
  double *pp = calloc(nk4, 8), *qq = calloc(nk4, 8), *rr = calloc(nk4, 8), *ss = calloc(nk2, 8),
         *xx = calloc(nk2, 8);
  COLBUF(lxxi); COLBUF(lyyi); COLBUF(lxyi); COLBUF(lyxi); COLBUF(uii); COLBUF(vii); COLBUF(wi);
  COLBUF(ev); COLBUF(gg); COLBUF(g_1); COLBUF(dwdz);
#define MAT(arr, r, c, k) arr[((r) - 1) + 2 * ((c) - 1) + 4 * (k)]
#define VEC(arr, r, k) arr[((r) - 1) + 2 * (k)]
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), o2 = I2(i, j);
      const double *un = F->un + o, *w = F->w_upd + o;
      /* :100-119 */
#define IAVG(dst, src)                                           \
  for (int k = 0; k < nk - 1; ++k) dst[k] = .5 * (src[o + k] + src[o + k + 1]); \
  dst[nk - 1] = 1.5 * src[o + nk - 1] - .5 * src[o + nk - 2]
      IAVG(lxxi, F->lxx); IAVG(lyyi, F->lyy); IAVG(lxyi, F->lxy); IAVG(lyxi, F->lyx);
      IAVG(uii, F->ui); IAVG(vii, F->vi); IAVG(wi, F->w_upd);
#undef IAVG
      (void)w;
      /* :121-127 */
      for (int k = 0; k < nk; ++k) {
        double udc = Fe[o + k] + lxxi[k] * uii[k] + lxyi[k] * vii[k];
        double vdc = Fn[o + k] + lyyi[k] * vii[k] - lyxi[k] * uii[k];
        double s1 = dtx2inv * unms[o + k] + F->hdu[o + k] + udc - adv_u[o + k] - zl[o + k];
        double s2 = dtx2inv * vnms[o + k] + F->hdv[o + k] + vdc - adv_v[o + k] - zp[o + k];
        VEC(ss, 1, k) = s1 * g->expz[k];
        VEC(ss, 2, k) = s2 * g->expz[k];
      }
      /* :143-152 */
      for (int k = 0; k < nk; ++k) ev[k] = P0 * F->scht[o + k] / GRAV;
      for (int k = 0; k < nk - 1; ++k) ev[k] = ev[k] * g->expz[k] * g->xmue[k];
      ev[nk - 1] = sqrt(ev[nk - 2] * ev[nk - 2] * ev[nk - 2] / ev[nk - 3]);
      for (int k = nk - 1; k >= 1; --k) ev[k] = sqrt(ev[k - 1] * ev[k]);
      ev[0] = ev[0] * ev[0] / ev[1];
      /* :154-164 */
      for (int k = 0; k < nk; ++k) {
        gg[k] = GRAV * (F->km[o + k] + ev[k]) / (P0 * F->schti[o + k] * (dz * dz));
        dwdz[k] = wi[k] / (2. * dz);
        dwdz[k] = dwdz[k] * g->expz[k];
      }
      for (int k = 0; k < nk - 1; ++k) g_1[k] = MUT(204) ? gg[k] : gg[k + 1];
      g_1[nk - 1] = 2. * gg[nk - 1] - gg[nk - 2];
      /* :166-205 */
      for (int k = 0; k < nk; ++k) {
        double p11 = MUT(203) ? -gg[k] : -(gg[k] + dwdz[k]);
        double r11 = -(g_1[k] - dwdz[k]);
        MAT(pp, 1, 1, k) = p11; MAT(pp, 2, 2, k) = p11; MAT(pp, 1, 2, k) = 0.; MAT(pp, 2, 1, k) = 0.;
        MAT(rr, 1, 1, k) = r11; MAT(rr, 2, 2, k) = r11; MAT(rr, 1, 2, k) = 0.; MAT(rr, 2, 1, k) = 0.;
        double q11 = gg[k] + g_1[k];
        double qa = dtx2inv + lxxi[k];
        qa = qa * g->expz[k];
        MAT(qq, 1, 1, k) = q11 + qa;
        qa = dtx2inv + lyyi[k];
        qa = qa * g->expz[k];
        MAT(qq, 2, 2, k) = q11 + qa;
        double q12 = MUT(201) ? g->cor[j] - un[k] / RE * g->tanphi[j] : g->cor[j] + un[k] / RE * g->tanphi[j];
        double q21 = q12;
        qa = q12 - (MUT(202) ? lyxi[k] : lxyi[k]);
        qa = qa * g->expz[k];
        MAT(qq, 1, 2, k) = -qa;
        qa = q21 - (MUT(202) ? lxyi[k] : lyxi[k]);
        qa = qa * g->expz[k];
        MAT(qq, 2, 1, k) = qa;
      }
      /* :207-215 */
      for (int r = 1; r <= 2; ++r)
        for (int c = 1; c <= 2; ++c) MAT(qq, r, c, 0) = MAT(qq, r, c, 0) - MAT(pp, r, c, 0);
      VEC(ss, 1, 0) = VEC(ss, 1, 0) - (MUT(206) ? 1. : 2.) * (MAT(pp, 1, 1, 0) * F->u_lbc[o2] + MAT(pp, 1, 2, 0) * F->v_lbc[o2]);
      VEC(ss, 2, 0) = VEC(ss, 2, 0) - 2. * (MAT(pp, 2, 1, 0) * F->u_lbc[o2] + MAT(pp, 2, 2, 0) * F->v_lbc[o2]);
      for (int n = 0; n < 4; ++n) pp[n] = 0.;
      for (int n = 0; n < 4; ++n) {
        qq[4 * (nk - 2) + n] = qq[4 * (nk - 2) + n] + rr[4 * (nk - 2) + n];
        rr[4 * (nk - 2) + n] = 0.;
      }
      if (g_cap)
        for (int k = 0; k < nk; ++k) {
          for (int r = 1; r <= 2; ++r)
            for (int c = 1; c <= 2; ++c) {
              int e = 2 * (r - 1) + (c - 1);
              CAPW(24 + e, I2(i, j), k, MAT(pp, r, c, k));
              CAPW(28 + e, I2(i, j), k, MAT(qq, r, c, k));
              CAPW(32 + e, I2(i, j), k, MAT(rr, r, c, k));
            }
          CAPW(36, I2(i, j), k, VEC(ss, 1, k));
          CAPW(37, I2(i, j), k, VEC(ss, 2, k));
        }
      /* :217-253 */
      tgx_o_blktri_ng(pp, qq, rr, ss, xx, nk, 1);
      if (g->glb) { /* duv.F:385-386 */
        VEC(xx, 1, nk - 1) = SPVAL;
        VEC(xx, 2, nk - 1) = SPVAL;
      } else {
        VEC(xx, 1, nk - 1) = 2 * VEC(xx, 1, nk - 2) - VEC(xx, 1, nk - 3);
        VEC(xx, 2, nk - 1) = 2 * VEC(xx, 2, nk - 2) - VEC(xx, 2, nk - 3);
      }
      for (int k = 0; k < nk; ++k) {
        un_upd[o + k] = VEC(xx, 1, k);
        vn_upd[o + k] = VEC(xx, 2, k);
      }
    }
#undef MAT
#undef VEC
  /* :296-321 */
  if (g->glb) {
    /* duv.F:490-493 filter_uv of UN and VN below the top level */
    glb_filter(g, un_upd, nk - 1, 3, 0);
    glb_filter(g, vn_upd, nk - 1, 3, 0);
  } else {
    lateral(un_upd, F->lonb_un, F->latb_un, nk, nlond, nlatd);
    lateral(vn_upd, F->lonb_vn, F->latb_vn, nk, nlond, nlatd);
  }
  for (size_t n = 0; n < (size_t)nlond * nlatd; ++n) {
    ulbc_nm[n] = ulbc[n];
    vlbc_nm[n] = vlbc[n];
    ulbc[n] = F->u_lbc[n];
    vlbc[n] = F->v_lbc[n];
  }
  for (size_t n = 0; n < N3; ++n) {
    unm_upd[n] = DTSMOOTH_DIV2 * (F->un_nm[n] + un_upd[n]) + DTSMOOTH * F->un[n];
    vnm_upd[n] = DTSMOOTH_DIV2 * (F->vn_nm[n] + vn_upd[n]) + DTSMOOTH * F->vn[n];
  }
  if (g->glb) {
    /* duv.F:501-514: the time smoothing below the top level, spval at the top */
    for (size_t n = (size_t)nk - 1; n < N3; n += (size_t)nk) unm_upd[n] = vnm_upd[n] = SPVAL;
    glb_periodic(g, un_upd);
    glb_periodic(g, vn_upd);
    glb_periodic(g, unm_upd);
    glb_periodic(g, vnm_upd);
  }
  free(adv_u); free(adv_v); free(ztmp); free(zl); free(zp); free(unms); free(vnms);
  free(pp); free(qq); free(rr); free(ss); free(xx);
  free(lxxi); free(lyyi); free(lxyi); free(lyxi); free(uii); free(vii); free(wi); free(ev); free(gg);
  free(g_1); free(dwdz);
}

/* ===================================================================== */
/* minor_ng.F90:1-239 for one species whose source/sink, boundary data and
 * constants are given. */
static void minor_core(const tgx_ogrid *g, const tgx_ofields *F, const double *fcomp,
                       const double *fcomp_tm1, const double *sloss, const double *sprod,
                       const double *flbc /* (nlond,nlatd,3) */, const double *fubc,
                       double rmx, const double phix[4], double alfax, const double *tlbc,
                       const double *lonb, const double *latb, double *out, double *out_nm, int cap_plane) {
  DIMS;
  const double dzp = g->dz, dtx2inv = g->dtx2inv, small = 1.e-12;
  const size_t N2 = (size_t)nlond * nlatd;
  double salfax[3] = {phix[0] - phix[3], phix[1] - phix[3], phix[2] - phix[3]};
  /* the denominator of ax: phix(3) in minor_ng.F90:135-136, phix(4) in the
   * global minor.F:359-362 */
  const double phix_ax = g->glb ? (MUT(801) ? phix[2] : phix[3]) : (MUT(301) ? phix[3] : phix[2]);
  ALLOC3(ftm1s); ALLOC3(hadvec);
  tgx_o_smooth(g, fcomp_tm1, ftm1s, g->shapiro);
  tgx_o_advec(g, fcomp, F->un, F->vn, hadvec);
  COLBUF(dmdz); COLBUF(pso1); COLBUF(pso2); COLBUF(pshe); COLBUF(do1dz); COLBUF(do2dz); COLBUF(dhedz);
  COLBUF(tni); COLBUF(ex); COLBUF(ax); COLBUF(thdiff); COLBUF(dtnidz); COLBUF(wi);
  COLBUF(molp); COLBUF(molq); COLBUF(molr); COLBUF(eddyp); COLBUF(eddyq); COLBUF(eddyr);
  COLBUF(pc); COLBUF(qc); COLBUF(rc); COLBUF(fr);
  double *sai = malloc(sizeof(double) * 3 * nk);
  for (int j = 0; j < nlatd; ++j) {
    double dfac = dfactor_of(g, j);
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), o2 = I2(i, j);
      const double *o2f = F->o2 + o, *o1f = F->o1 + o, *hef = F->he + o, *mbar = F->mbar + o;
      const double *barm = F->barm + o, *tn = F->tn + o;
      /* :66-77 */
      double bo2 = g->b[0][0] * o2f[0] + g->b[0][1] * o1f[0] + g->b[0][2] * hef[0] + g->fb[0];
      double bo1 = g->b[1][0] * o2f[0] + g->b[1][1] * o1f[0] + g->b[1][2] * hef[0] + g->fb[1];
      double bhe = g->b[2][0] * o2f[0] + g->b[2][1] * o1f[0] + g->b[2][2] * hef[0] + g->fb[2];
      double xmbari = 1. / (bo2 * RIO2 + bo1 * RIO1 + (1. - bo2 - bo1 - bhe) * RIN2 + bhe * RIHE);
      dmdz[0] = (mbar[0] - xmbari) / dzp;
      pso1[0] = .5 * (bo1 + o1f[0]);
      pso2[0] = .5 * (bo2 + o2f[0]);
      pshe[0] = .5 * (bhe + hef[0]);
      do1dz[0] = (o1f[0] - bo1) / dzp;
      do2dz[0] = (o2f[0] - bo2) / dzp;
      dhedz[0] = (hef[0] - bhe) / dzp;
      /* :79-87 */
      for (int k = 1; k < nk; ++k) {
        dmdz[k] = (mbar[k] - mbar[k - 1]) / dzp;
        pso1[k] = .5 * (o1f[k] + o1f[k - 1]);
        pso2[k] = .5 * (o2f[k] + o2f[k - 1]);
        pshe[k] = .5 * (hef[k] + hef[k - 1]);
        do1dz[k] = (o1f[k] - o1f[k - 1]) / dzp;
        do2dz[k] = (o2f[k] - o2f[k - 1]) / dzp;
        dhedz[k] = (hef[k] - hef[k - 1]) / dzp;
      }
      /* :89-93 */
      tni[0] = tlbc[o2];
      tni[nk - 1] = tn[nk - 2];
      for (int k = 1; k < nk - 1; ++k) tni[k] = .5 * (tn[k] + tn[k - 1]);
      /* :97-126 */
      for (int k = 0; k < nk; ++k) {
        double df[3], alfa[9], inv[9];
        for (int n = 0; n < 3; ++n) df[n] = pow(tni[k] / T00, 1.75 - SS[n]);
#define AL(r, c) alfa[((r) - 1) + 3 * ((c) - 1)]
#define PH(r, c) PHI[(r) - 1][(c) - 1]
        AL(1, 1) = -(PH(1, 4) + (PH(1, 2) - PH(1, 4)) * pso1[k] + (df[0] * PH(1, 3) - PH(1, 4)) * pshe[k]);
        AL(2, 2) = -(PH(2, 4) + (PH(2, 1) - PH(2, 4)) * pso2[k] + (df[1] * PH(2, 3) - PH(2, 4)) * pshe[k]);
        AL(3, 3) = -(df[2] * PH(3, 4) + (df[0] * PH(3, 1) - df[2] * PH(3, 4)) * pso2[k] +
                     (df[1] * PH(3, 2) - df[2] * PH(3, 4)) * pso1[k]);
        AL(1, 2) = (PH(1, 2) - PH(1, 4)) * pso2[k];
        AL(1, 3) = (df[0] * PH(1, 3) - PH(1, 4)) * pso2[k];
        AL(2, 1) = (PH(2, 1) - PH(2, 4)) * pso1[k];
        AL(2, 3) = (df[1] * PH(2, 3) - PH(2, 4)) * pso1[k];
        AL(3, 1) = (df[0] * PH(3, 1) - df[2] * PH(3, 4)) * pshe[k];
        AL(3, 2) = (df[1] * PH(3, 2) - df[2] * PH(3, 4)) * pshe[k];
#undef AL
#undef PH
        tgx_o_matinv3(alfa, inv);
        for (int n = 0; n < 3; ++n) {
          double s = 0.;
          for (int m = 0; m < 3; ++m) s = s + salfax[m] * inv[m + 3 * n];
          sai[3 * k + n] = s;
        }
      }
      /* :128-143 */
      for (int k = 0; k < nk; ++k) {
        ex[k] = 1. - (rmx + dmdz[k]) / barm[k] +
                sai[3 * k + 0] * (do2dz[k] - (1. - (RMO2 + dmdz[k]) / barm[k]) * pso2[k]) +
                sai[3 * k + 1] * (do1dz[k] - (1. - (RMO1 + dmdz[k]) / barm[k]) * pso1[k]) +
                sai[3 * k + 2] * (dhedz[k] - (1. - (RMHE + dmdz[k]) / barm[k]) * pshe[k]);
        dmdz[k] = dmdz[k] / barm[k];
        ax[k] = -(barm[k] / (TAU * RMN2) * pow(T00 / tni[k], 0.25) /
                  (phix_ax + salfax[0] * pso2[k] + salfax[1] * pso1[k] + salfax[2] * pshe[k]));
      }
      for (int k = 1; k < nk - 1; ++k) dtnidz[k] = (tni[k + 1] - tni[k - 1]) / 2.;
      dtnidz[0] = tni[1] - tni[0];
      dtnidz[nk - 1] = tni[nk - 1] - tni[nk - 2];
      for (int k = 0; k < nk; ++k) thdiff[k] = ex[k] - alfax * dtnidz[k] / (dzp * tni[k]);
      /* :156-184 */
      for (int k = 0; k < nk - 1; ++k) wi[k] = .5 * (F->w_upd[o + k] + F->w_upd[o + k + 1]);
      wi[nk - 1] = 1.5 * F->w_upd[o + nk - 1] - .5 * F->w_upd[o + nk - 2];
      for (int k = 0; k < nk - 1; ++k) {
        molp[k] = ax[k] * (1. / dzp + .5 * thdiff[k]);
        molr[k] = ax[MUT(304) ? k : k + 1] * (1. / dzp - .5 * thdiff[k + 1]);
        molq[k] = ax[k] * (1. / dzp - .5 * thdiff[k]) + ax[k + 1] * (1. / dzp + .5 * thdiff[k + 1]);
        eddyp[k] = g->expzmid_inv * g->difk[k] * (1. / dzp + (MUT(302) ? .5 : -.5) * dmdz[k]);
        eddyr[k] = g->expzmid * g->difk[k + 1] * (1. / dzp + .5 * dmdz[k + 1]);
        eddyq[k] = g->expzmid_inv * g->difk[k] * (1. / dzp + .5 * dmdz[k]) +
                   g->expzmid * g->difk[k + 1] * (1. / dzp - .5 * dmdz[k + 1]);
      }
      for (int k = 0; k < nk - 1; ++k) {
        pc[k] = molp[k] / dzp - g->expz[k] * (eddyp[k] * dfac + .5 * wi[k]) / dzp;
        rc[k] = molr[k] / dzp - g->expz[k] * (eddyr[k] * dfac - .5 * wi[k]) / dzp;
        qc[k] = -molq[k] / dzp + g->expz[k] * (eddyq[k] * dfac / dzp + dtx2inv - sloss[o + k]);
      }
      /* :186-198 */
      for (int k = 0; k < nk; ++k) {
        double s0prod = MUT(303) ? sprod[o + k] * rmx * F->xnmbar[o + k] : sprod[o + k] * rmx / F->xnmbar[o + k];
        fr[k] = ftm1s[o + k] * dtx2inv - hadvec[o + k] + s0prod;
        fr[k] = fr[k] * g->expz[k];
      }
      double l1 = flbc[o2], l2 = flbc[o2 + N2], l3 = flbc[o2 + 2 * N2];
      qc[0] = qc[0] + pc[0] * (l1 + .5 * l2 * dzp) / (l1 - .5 * l2 * dzp);
      fr[0] = fr[0] - (MUT(306) ? .5 : 1.) * pc[0] * l3 * dzp / (l1 - .5 * l2 * dzp);
      pc[0] = 0.;
      pc[nk - 1] = 1. + (MUT(305) ? -.5 : .5) * dzp * thdiff[nk - 1];
      qc[nk - 1] = pc[nk - 1] - 2.;
      rc[nk - 1] = 0.;
      fr[nk - 1] = -GRAV * rmx * fubc[o2] * dzp / (P0 * ax[nk - 1] * AVO);
      if (g_cap)
        for (int k = 0; k < nk; ++k) {
          CAPW(cap_plane + 0, I2(i, j), k, pc[k]); CAPW(cap_plane + 1, I2(i, j), k, qc[k]);
          CAPW(cap_plane + 2, I2(i, j), k, rc[k]); CAPW(cap_plane + 3, I2(i, j), k, fr[k]);
        }
      /* :200 */
      tgx_o_trsolv_ng(pc, qc, rc, fr, out + o, nk, 1, nk);
    }
  }
  if (g->glb) {
    /* minor.F:546-597: filter_minor, the time smoothing, the periodic points,
     * the floor */
    if (MUT(805))
      for (size_t n = 0; n < N3; ++n) out_nm[n] = DTSMOOTH * fcomp[n] + DTSMOOTH_DIV2 * (fcomp_tm1[n] + out[n]);
    if (!MUT(804)) glb_filter(g, out, nk, 3, 0);
    if (!MUT(805))
      for (size_t n = 0; n < N3; ++n) out_nm[n] = DTSMOOTH * fcomp[n] + DTSMOOTH_DIV2 * (fcomp_tm1[n] + out[n]);
    glb_periodic(g, out);
    glb_periodic(g, out_nm);
    for (size_t n = 0; n < N3; ++n) {
      out[n] = out[n] > small ? out[n] : small;
      out_nm[n] = out_nm[n] > small ? out_nm[n] : small;
    }
  } else {
    /* :228-237 */
    lateral(out, lonb, latb, nk, nlond, nlatd);
    for (size_t n = 0; n < N3; ++n) {
      out_nm[n] = DTSMOOTH * fcomp[n] + DTSMOOTH_DIV2 * (fcomp_tm1[n] + out[n]);
      out[n] = out[n] > small ? out[n] : small;
      out_nm[n] = out_nm[n] > small ? out_nm[n] : small;
    }
  }
  free(ftm1s); free(hadvec); free(sai);
  free(dmdz); free(pso1); free(pso2); free(pshe); free(do1dz); free(do2dz); free(dhedz);
  free(tni); free(ex); free(ax); free(thdiff); free(dtnidz); free(wi);
  free(molp); free(molq); free(molr); free(eddyp); free(eddyq); free(eddyr);
  free(pc); free(qc); free(rc); free(fr);
}

/* comp_no_ng.F90, comp_n4s_ng.F90, comp_n2d_ng.F90, comp_ar_ng.F90: the
 * species source/sink and boundary data, then minor_ng. */
void tgx_o_minor_species(const tgx_ogrid *g, const tgx_ofields *F, int species,
                         const double *n2d_upd, const double *tlbc,
                         double *out, double *out_nm) {
  DIMS;
  const size_t N2 = (size_t)nlond * nlatd;
  ALLOC3(sloss); ALLOC3(sprod);
  double *flbc = malloc(sizeof(double) * 3 * N2), *fubc = calloc(N2, sizeof(double));
  const double *fc, *fctm1, *lonb, *latb;
  double rmx, alfax;
  double phix[4];
  if (species == 0) { /* NO: comp_no_ng.F90:43-59, comp_no.F:25-26 */
    static const double ph[4] = {0.814, 0.866, 0.662, 0.926};
    memcpy(phix, ph, sizeof ph);
    alfax = 0.;
    rmx = RMNO;
    fc = F->no; fctm1 = F->no_nm; lonb = F->lonb_no; latb = F->latb_no;
    const double nob = 4.e6;
    for (size_t n = 0; n < N2; ++n) {
      flbc[n] = 0.;
      flbc[n + N2] = 1.;
      flbc[n + 2 * N2] = -nob * RMNO / F->xnmbari[(size_t)nk * n];
    }
    for (int j = 0; j < nlatd; ++j)
      for (int i = 0; i < nlond; ++i) {
        size_t o = IX(0, i, j);
        for (int k = 0; k < nk; ++k) {
          size_t p = o + k;
          double xn = F->xnmbar[p];
          sprod[p] = xn * xn * F->o2[p] * RIO2 * (F->beta1[p] * F->n4s[p] * RIN4S + BETA2 * n2d_upd[p] * RIN2D) +
                     xn * xn * xn * F->beta17[p] * F->o1[p] * RIO1 * F->n2[p] * RIN2 * F->n4s[p] * RIN4S;
          double betai;
          if (k < nk - 1)
            betai = .5 * (F->beta8[p] + F->beta9[p] + F->beta9n[p] + F->beta8[p + 1] + F->beta9[p + 1] + F->beta9n[p + 1]);
          else
            betai = 1.5 * (F->beta8[p] + F->beta9[p] + F->beta9n[p]) - .5 * (F->beta8[p - 1] + F->beta9[p - 1] + F->beta9n[p - 1]);
          sloss[p] = -xn * (F->beta3[p] * F->n4s[p] * RIN4S + BETA6 * n2d_upd[p] * RIN2D) - RK5 * F->o2p[p] - betai;
        }
      }
  } else if (species == 1) { /* N4S: comp_n4s_ng.F90:112-140, comp_n4s.F:35-36 */
    static const double ph[4] = {0.651, 0.731, 0.627, 0.741};
    memcpy(phix, ph, sizeof ph);
    alfax = 0.;
    rmx = RMN4S;
    fc = F->n4s; fctm1 = F->n4s_nm; lonb = F->lonb_n4s; latb = F->latb_n4s;
    for (size_t n = 0; n < N2; ++n) {
      size_t p = (size_t)nk * n;
      double xb = F->xnmbari[p];
      flbc[n] = 0.;
      flbc[n + N2] = 1.;
      flbc[n + 2 * N2] =
          -(RMN4S / xb *
            (F->qtef[p] * (1. - BRN2D) / xb +
             n2d_upd[p] * RIN2D * (BETA4 * xb * F->o1[p] * RIO1 + F->beta5[p] * F->ne[p] + BETA7) +
             F->beta8[p] * F->no[p] * RINO) /
            (F->beta1[p] * F->o2[p] * RIO2 + F->beta3[p] * F->no[p] * RINO +
             xb * F->beta17[p] * F->o1[p] * RIO1 * F->n2[p] * RIN2));
    }
    for (int j = 0; j < nlatd; ++j)
      for (int i = 0; i < nlond; ++i) {
        size_t o = IX(0, i, j);
        for (int k = 0; k < nk; ++k) {
          size_t p = o + k;
          double qtefi, beta8i, nei1, nei2;
          if (k < nk - 1) {
            qtefi = .5 * (F->qtef[p] + F->qtef[p + 1]);
            beta8i = .5 * (F->beta8[p] + F->beta8[p + 1]);
            nei1 = .5 * (F->ne[p] + F->ne[p + 1]);
            nei2 = sqrt(F->ne[p] * F->ne[p + 1]);
          } else {
            /* comp_n4s_ng.F90:128 writes qtef(nk), leaving qtefi(nk) undefined;
             * it only reaches f_rhs(nk), which minor_ng overwrites (:198). */
            qtefi = 0.;
            beta8i = 1.5 * F->beta8[p] - .5 * F->beta8[p - 1];
            nei1 = 1.5 * F->ne[p] - .5 * F->ne[p - 1];
            nei2 = sqrt(F->ne[p] * F->ne[p] * F->ne[p] / F->ne[p - 1]);
          }
          double xn = F->xnmbar[p];
          sprod[p] = qtefi * (1. - BRN2D) +
                     xn * (n2d_upd[p] * RIN2D * (xn * BETA4 * F->o1[p] * RIO1 + F->beta5[p] * nei1 + BETA7) +
                           beta8i * F->no[p] * RINO) +
                     xn * (F->rk2[p] * F->op[p] * F->n2[p] * RIN2 + RK6 * F->nplus[p] * F->o2[p] * RIO2 +
                           RK8 * F->nplus[p] * F->o1[p] * RIO1) +
                     nei2 * (F->ra1[p] * F->nop[p] * 0.15 + F->ra3[p] * F->n2p[p] * 1.1);
          sloss[p] = -xn * (F->beta1[p] * F->o2[p] * RIO2 + F->beta3[p] * F->no[p] * RINO +
                            xn * F->beta17[p] * F->o1[p] * RIO1 * F->n2[p] * RIN2) -
                     RK4 * F->o2p[p];
        }
      }
  } else if (species == 2) { /* N2D: comp_n2d_ng.F90:183-205, comp_n2d.F:26-27 */
    static const double ph[4] = {0.651, 0.731, 0.627, 0.741};
    memcpy(phix, ph, sizeof ph);
    alfax = 0.;
    rmx = RMN2D;
    fc = F->n2d; fctm1 = F->n2d_nm; lonb = F->lonb_n2d; latb = F->latb_n2d;
    for (size_t n = 0; n < N2; ++n) {
      size_t p = (size_t)nk * n;
      double xb = F->xnmbari[p];
      flbc[n] = 0.;
      flbc[n + N2] = 1.;
      flbc[n + 2 * N2] = -(RMN2D / xb * F->qtef[p] * BRN2D /
                           (xb * (BETA2 * F->o2[p] * RIO2 + BETA4 * F->o1[p] * RIO1 + BETA6 * F->no[p] * RINO) +
                            BETA7 + F->beta5[p] * F->ne[p]));
    }
    for (int j = 0; j < nlatd; ++j)
      for (int i = 0; i < nlond; ++i) {
        size_t o = IX(0, i, j);
        for (int k = 0; k < nk; ++k) {
          size_t p = o + k;
          double qtefi, nei;
          if (k < nk - 1) {
            qtefi = .5 * (F->qtef[p] + F->qtef[p + 1]);
            nei = sqrt(F->ne[p] * F->ne[p + 1]);
          } else {
            qtefi = 1.5 * F->qtef[p] - .5 * F->qtef[p - 1];
            nei = sqrt(F->ne[p] * F->ne[p] * F->ne[p] / F->ne[p - 1]);
          }
          double xn = F->xnmbar[p];
          sprod[p] = qtefi * BRN2D + F->rk3[p] * F->n2p[p] * xn * F->o1[p] * RIO1 +
                     (F->ra1[p] * F->nop[p] * 0.85 + F->ra3[p] * F->n2p[p] * 0.9) * nei;
          sloss[p] = -(xn * (BETA2 * F->o2[p] * RIO2 + BETA4 * F->o1[p] * RIO1 + BETA6 * F->no[p] * RINO) + BETA7 +
                       F->beta5[p] * nei + RK10 * F->op[p]);
        }
      }
  } else { /* AR: comp_ar_ng.F90:219-237 */
    static const double ph[4] = {1.042, 1.509, 0.732, 1.176};
    memcpy(phix, ph, sizeof ph);
    alfax = 0.17;
    rmx = RMAR;
    fc = F->ar; fctm1 = F->ar_nm; lonb = F->lonb_ar; latb = F->latb_ar;
    double l3 = -sqrt(F->ar_glbm1 * F->ar_glbm2);
    for (size_t n = 0; n < N2; ++n) {
      flbc[n] = 0.;
      flbc[n + N2] = 1.;
      flbc[n + 2 * N2] = l3;
    }
    for (size_t n = 0; n < N3; ++n) { sprod[n] = 0.; sloss[n] = 0.; }
  }
  minor_core(g, F, fc, fctm1, sloss, sprod, flbc, fubc, rmx, phix, alfax, tlbc, lonb, latb, out, out_nm,
             8 + 4 * species);
  free(sloss); free(sprod); free(flbc); free(fubc);
}

/* ===================================================================== */
/* oplus_ng.F90:344-365 (bdotdh) */
static void bdotdh(const tgx_ogrid *g, const tgx_ofields *F, const double *phij, double *ans) {
  DIMS;
  ALLOC3(dx); ALLOC3(dy);
  tgx_o_df3d(phij, dx, 1, nk, nlond, nlatd);
  tgx_o_df3d(phij, dy, -1, nk, nlond, nlatd);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        size_t p = IX(k, i, j), q = I2(i, j);
        double a1 = F->bx[q] * dx[p];
        double a2 = F->by[q] * dy[p];
        a1 = a1 / g->cs[j];
        ans[p] = 1. / RE * (a1 / g->dlamda + a2 / g->dphi);
      }
  free(dx); free(dy);
}

/* oplus_ng.F90:1-342 */
void tgx_o_oplus_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *op,
                    const double *optm1, const double *lonb, const double *latb,
                    double *opout, double *optm1out, double *xiop2p, double *xiop2d,
                    double *Fe, double *Fn) {
  DIMS;
  const size_t N2 = (size_t)nlond * nlatd;
  const double pi = pi_(), rtd = 180. / pi, dz = g->dz, dtx2inv = g->dtx2inv;
  const double explic = 1., phid = 2.0e8, phin = -2.0e8, ppolar = 0.;
  const int nsub = g->nstep_sub;
  double *opflux = malloc(8 * N2), *dvb = malloc(8 * N2), *csby = malloc(8 * N2), *dbxdx = malloc(8 * N2),
         *dcsbydy = malloc(8 * N2);
  /* :77-91 */
  for (size_t n = 0; n < N2; ++n) {
    double rl = F->rlatm[n], a = 0.;
    if (fabs(rl) >= pi / 4.5) a = 1.;
    if (fabs(rl) < pi / 4.5 && fabs(rl) > pi / 18) a = .5 * (1. + cos(fabs(rl) * 6. - pi * 4. / 3.));
    if (fabs(rl) <= pi / 18) a = 0.;
    double fed = phid * a, fen = phin * a, chi = F->chi[n], of;
    if (chi >= 0.5 * pi) of = fen; else of = fed;
    if ((chi * rtd - 80.) * (chi * rtd - 100.) < 0.)
      of = .5 * (fed + fen) + .5 * (fed - fen) * cos(pi * (chi * rtd - 80.) / 20.);
    if (fabs(rl) >= pi / 3.) of = of + ppolar;
    opflux[n] = of;
  }
  /* :94-104 */
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) csby[I2(i, j)] = g->cs[j] * F->by[I2(i, j)];
  tgx_o_df2d(F->bx, dbxdx, 1, nlond, nlatd);
  tgx_o_df2d(csby, dcsbydy, -1, nlond, nlatd);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t n = I2(i, j);
      double v = dbxdx[n] / g->dlamda + dcsbydy[n] / g->dphi;
      v = v / g->cs[j];
      dvb[n] = (v + 2. * F->bz[n]) / RE;
    }
  ALLOC3(dj); ALLOC3(vni); ALLOC3(bdotu); ALLOC3(bvel); ALLOC3(diffj); ALLOC3(tp_op);
  ALLOC3(opts2); ALLOC3(djbz); ALLOC3(bd_op); ALLOC3(bd_opj); ALLOC3(bd_diff); ALLOC3(bd_bvel);
  ALLOC3(opbm2); ALLOC3(dopx); ALLOC3(dopy); ALLOC3(bd_djbz);
  COLBUF(wni); COLBUF(dtp);
  /* :106-157, per column (vertical parts) */
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), n2 = I2(i, j);
      for (int k = 0; k < nk - 1; ++k) wni[k] = .5 * (F->w[o + k] + F->w[o + k + 1]);
      wni[nk - 1] = 1.5 * F->w[o + nk - 1] - .5 * F->w[o + nk - 2];
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double tr = 0.5 * (F->tn[p] + F->ti[p]);
        double t = 1. - 0.064 * log10(tr);
        double v = 18.1 * F->o2[p] * RIO2 + F->o1[p] * RIO1 * sqrt(tr) * (t * t) * g->colfac +
                   3.6 * F->he[p] * RIHE + 18.6 * F->n2[p] * RIN2;
        dj[p] = 1.42E17 / (F->xnmbar[p] * v);
        vni[p] = 16 * 3.53E-11 * v;
        if (g->opdiffcap != 0.) {
          double cap = g->opdiffcap / (1 + pow(2., g->zpint[k] - 6));
          dj[p] = dj[p] < cap ? dj[p] : cap;
        }
        double hj = F->scht[p];
        bdotu[p] = F->bx[n2] * F->un[p] + F->by[n2] * F->vn[p] + hj * F->bz[n2] * wni[k];
        bvel[p] = bdotu[p] * op[p];
        tp_op[p] = (F->te[p] + F->ti[p]) * op[p];
      }
      for (int k = 0; k < nk - 2; ++k) dtp[k + 1] = (tp_op[o + k + 2] - tp_op[o + k]) / 2.;
      dtp[nk - 2] = tp_op[o + nk - 2] - tp_op[o + nk - 3];
      dtp[nk - 1] = tp_op[o + nk - 1] - tp_op[o + nk - 2];
      dtp[0] = tp_op[o + 1] - tp_op[o + 0];
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        diffj[p] = 1. / (F->scht[p] * dz) * dtp[k] + (MUT(405) ? -RMOP : RMOP) * GRAV / GASK * op[p];
      }
    }
  /* :159 */
  tgx_o_smooth(g, optm1, opts2, MUT(406) ? g->shapiro : g->shapiro / nsub);
  /* :161-170 */
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t n2 = I2(i, j);
      double sncsdip = sin(F->dipmag[n2]) * cos(F->dipmag[n2]);
      for (int k = 0; k < nk; ++k) {
        size_t p = IX(k, i, j);
        double wd = vni[p] * diffj[p] * dj[p];
        Fe[p] = wd * sncsdip * F->sndec[n2];
        Fn[p] = wd * sncsdip * F->csdec[n2];
        djbz[p] = dj[p] * F->bz[n2];
        opbm2[p] = op[p] / (F->bmod2[n2] * F->bmod2[n2]);
      }
    }
  /* :172-178 */
  bdotdh(g, F, diffj, bd_op);
  for (size_t n = 0; n < N3; ++n) bd_op[n] = djbz[n] * bd_op[n];
  bdotdh(g, F, tp_op, bd_opj);
  for (size_t n = 0; n < N3; ++n) bd_opj[n] = bd_opj[n] * dj[n];
  bdotdh(g, F, bd_opj, bd_diff);
  bdotdh(g, F, bvel, bd_bvel);
  tgx_o_df3d(opbm2, dopx, 1, nk, nlond, nlatd);
  tgx_o_df3d(opbm2, dopy, -1, nk, nlond, nlatd);
  bdotdh(g, F, djbz, bd_djbz);
  COLBUF(dbd); COLBUF(explicit_); COLBUF(hdz); COLBUF(hdzi); COLBUF(tp1); COLBUF(tphdz0); COLBUF(tphdz1);
  COLBUF(djint); COLBUF(hdzm); COLBUF(hdzp); COLBUF(diffp); COLBUF(diffq); COLBUF(diffr);
  COLBUF(pcf); COLBUF(qcf); COLBUF(rcf); COLBUF(oploss);
  const double gmr = GRAV * RMOP / (2. * GASK);
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), n2 = I2(i, j);
      double bz = F->bz[n2], dvbz;
      /* :181-192 */
      for (int k = 1; k < nk - 2; ++k) dbd[k] = (bd_opj[o + k + 1] - bd_opj[o + k - 1]) / 2.;
      dbd[nk - 2] = bd_opj[o + nk - 2] - bd_opj[o + nk - 3];
      dbd[nk - 1] = bd_opj[o + nk - 1] - bd_opj[o + nk - 2];
      dbd[0] = bd_opj[o + 1] - bd_opj[o + 0];
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double bdzdvb = bz / (F->scht[p] * dz) * dbd[k] + dvb[n2] * bd_opj[p];
        double diffexp = bdzdvb + bd_diff[p] + bd_op[p];
        explicit_[k] = -explic * diffexp;
      }
      /* :194-211 */
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double uii = k < nk - 1 ? .5 * (F->ui[p] + F->ui[p + 1]) : 1.5 * F->ui[p] - .5 * F->ui[p - 1];
        double vii = k < nk - 1 ? .5 * (F->vi[p] + F->vi[p + 1]) : 1.5 * F->vi[p] - .5 * F->vi[p - 1];
        double vx = uii * dopx[p], vy = vii * dopy[p];
        vx = vx / g->cs[j];
        double vh = 1. / RE * (vx / g->dlamda + vy / g->dphi);
        vh = vh * (F->bmod2[n2] * F->bmod2[n2]);
        explicit_[k] = explicit_[k] + vh + bd_bvel[p];
      }
      /* :213-248 */
      dvbz = dvb[n2] / bz;
      for (int k = 0; k < nk; ++k) {
        hdz[k] = 1. / (F->scht[o + k] * dz);
        tp1[k] = 0.5 * (F->ti[o + k] + F->te[o + k]);
      }
      for (int k = 0; k < nk - 2; ++k) hdzi[k + 1] = 0.5 * (hdz[k] + hdz[k + 1]);
      hdzi[0] = 1.5 * hdz[0] - 0.5 * hdz[1];
      hdzi[nk - 1] = 1.5 * hdz[nk - 2] - 0.5 * hdz[nk - 3];
      for (int k = 0; k < nk; ++k) tphdz1[k] = 2. * tp1[k] * hdzi[k] + gmr;
      for (int k = 0; k < nk - 1; ++k) tphdz0[k + 1] = 2. * tp1[k] * hdzi[k + 1] - gmr;
      tphdz1[nk - 1] = 2. * (2. * tp1[nk - 2] - tp1[nk - 3]) * hdzi[nk - 1] + gmr;
      tphdz0[0] = 2. * (2. * tp1[0] - tp1[1]) * hdzi[0] - gmr;
      for (int k = 0; k < nk - 2; ++k) djint[k + 1] = 0.5 * (dj[o + k] + dj[o + k + 1]);
      djint[0] = 1.5 * dj[o + 0] - 0.5 * dj[o + 1];
      djint[nk - 1] = 1.5 * dj[o + nk - 2] - 0.5 * dj[o + nk - 3];
      for (int k = 0; k < nk; ++k) {
        double divbz = dvbz + bd_djbz[o + k] / (dj[o + k] * (bz * bz));
        hdzm[k] = hdz[k] - 0.5 * divbz;
        hdzp[k] = hdz[k] + 0.5 * divbz;
        hdzm[k] = hdzm[k] * (bz * bz);
        hdzp[k] = hdzp[k] * (bz * bz);
      }
      /* :250 */
      for (int k = 0; k < nk; ++k) explicit_[k] = explicit_[k] - opts2[o + k] * dtx2inv * nsub;
      /* :252-278 */
      for (int k = 0; k < nk; ++k) diffp[k] = hdzm[k] * djint[k] * tphdz0[k];
      for (int k = 0; k < nk - 1; ++k) {
        diffq[k] = -(hdzp[k] * djint[k + 1] * tphdz0[k + 1] + hdzm[k] * djint[k] * tphdz1[k]);
        diffr[k] = hdzp[k] * djint[k + 1] * tphdz1[k + 1];
      }
      diffq[nk - 1] = -(hdzp[nk - 1] * (2. * djint[nk - 1] * tphdz0[nk - 1] - djint[nk - 2] * tphdz0[nk - 2]) +
                        hdzm[nk - 1] * djint[nk - 1] * tphdz1[nk - 1]);
      diffr[nk - 1] = hdzp[nk - 1] * (2. * djint[nk - 1] * tphdz1[nk - 1] - djint[nk - 2] * tphdz1[nk - 2]);
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double wii = k < nk - 1 ? .5 * (F->wi[p] + F->wi[p + 1]) : 1.5 * F->wi[p] - .5 * F->wi[p - 1];
        double driftp = wii * 0.5 * hdz[k];
        double driftq = (MUT(401) ? wii : -wii) * 6. / RE;
        double driftr = -wii * 0.5 * hdz[k];
        double windp, windq, windr;
        double bu = bdotu[p];
        if (k == 0) windp = bz * (2. * bdotu[o] - bdotu[o + 1]) * 0.5 * hdz[0];
        else if (k <= nk - 2) windp = bz * bdotu[p - 1] * 0.5 * hdz[k];
        else windp = 0.; /* windp(nk) is never set (oplus_ng.F90:265-274) nor used */
        windq = -bu * dvbz * bz;
        if (k < nk - 2) windr = -bz * bdotu[MUT(402) ? p : p + 1] * 0.5 * hdz[k];
        else if (k == nk - 2) windr = bz * (2. * bdotu[o + nk - 2] - bdotu[o + nk - 3]) * 0.5 * hdz[nk - 2];
        else windr = bz * (2. * bdotu[o + nk - 1] - bdotu[o + nk - 2]) * 0.5 * hdz[nk - 1];
        pcf[k] = diffp[k] + driftp + windp;
        qcf[k] = diffq[k] + driftq + windq - dtx2inv * nsub;
        rcf[k] = diffr[k] + driftr + windr;
      }
      /* :280-286 */
      double ubcb = -(bz * bz) * djint[nk - 1] * tphdz0[nk - 1];
      double ubca = -(bz * bz) * djint[nk - 1] * tphdz1[nk - 1];
      qcf[nk - 2] = qcf[nk - 2] + (MUT(404) ? ubca / ubcb : ubcb / ubca) * rcf[nk - 2];
      explicit_[nk - 2] = explicit_[nk - 2] - opflux[n2] * rcf[nk - 2] / ubca;
      rcf[nk - 2] = 0.;
      /* :288-303 */
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double xn = F->xnmbar[p], ne = F->ne[p];
        double qop2pi = k < nk - 1 ? .5 * (F->qop2p[p] + F->qop2p[p + 1]) : 1.5 * F->qop2p[p] - .5 * F->qop2p[p - 1];
        double qop2di = k < nk - 1 ? .5 * (F->qop2d[p] + F->qop2d[p + 1]) : 1.5 * F->qop2d[p] - .5 * F->qop2d[p - 1];
        double qopi = k < nk - 1 ? .5 * (F->qop[p] + F->qop[p + 1]) : 1.5 * F->qop[p] - .5 * F->qop[p - 1];
        double x2p = qop2pi / (xn * ((RK16 + RK17) * F->n2[p] * RIN2 + RK18 * F->o1[p] * RIO1) +
                               (F->rk19[p] + F->rk20[p]) * ne + RK21 + RK22);
        double x2d = (qop2di + (F->rk20[p] * ne + RK22) * x2p) /
                     (xn * (RK23 * F->n2[p] * RIN2 + RK24 * F->o1[p] * RIO1 + RK26 * F->o2[p] * RIO2) +
                      F->rk25[p] * ne + RK27);
        xiop2p[p] = x2p;
        xiop2d[p] = x2d;
        oploss[k] = xn * (F->rk1[p] * F->o2[p] * RIO2 + F->rk2[p] * F->n2[p] * RIN2 + RK10 * F->n2d[p] * RIN2D);
        if (MUT(403)) oploss[k] = 2. * oploss[k];
        qcf[k] = qcf[k] - oploss[k];
        double op_prod = qopi + (F->rk19[p] * ne + RK21) * x2p + (F->rk25[p] * ne + RK27) * x2d +
                         (RK18 * x2p + RK24 * x2d) * F->o1[p] * RIO1 * xn;
        explicit_[k] = explicit_[k] - op_prod;
      }
      qcf[0] = qcf[0] - pcf[0];
      explicit_[0] = explicit_[0] - 2. * pcf[0] * F->qop[o] / (1.5 * oploss[0] - 0.5 * oploss[1]);
      pcf[0] = 0.;
      if (g_cap)
        for (int k = 0; k < nk; ++k) {
          CAPW(4, I2(i, j), k, pcf[k]); CAPW(5, I2(i, j), k, qcf[k]);
          CAPW(6, I2(i, j), k, rcf[k]); CAPW(7, I2(i, j), k, explicit_[k]);
        }
      /* :305-307 */
      tgx_o_trsolv_ng(pcf, qcf, rcf, explicit_, opout + o, nk, 1, nk - 1);
      if (!g->glb) opout[o + nk - 1] = opout[o + nk - 2] * opout[o + nk - 2] / opout[o + nk - 3];
    }
  if (g->glb) {
    /* oplus.F:896-941: filter_op ('OPLUS', the chunk-mean-preserving
     * reconstruction) below the top level, the time smoothing, zero at the
     * top level (:922-923), the floor, the periodic points */
    glb_filter(g, opout, nk - 1, 3, 1);
    for (size_t n = 0; n < N3; ++n) {
      if (n % (size_t)nk == (size_t)nk - 1) {
        opout[n] = MUT(807) ? opout[n - 1] : 0.;
        optm1out[n] = 0.;
        continue;
      }
      optm1out[n] = DTSMOOTH * op[n] + DTSMOOTH_DIV2 * (optm1[n] + opout[n]);
      opout[n] = opout[n] > 0. ? opout[n] : 0.;
      optm1out[n] = optm1out[n] > 0. ? optm1out[n] : 0.;
    }
    glb_periodic(g, opout);
    glb_periodic(g, optm1out);
  } else {
    /* :332-340 */
    lateral(opout, lonb, latb, nk, nlond, nlatd);
    for (size_t n = 0; n < N3; ++n) {
      optm1out[n] = DTSMOOTH * op[n] + DTSMOOTH_DIV2 * (optm1[n] + opout[n]);
      opout[n] = opout[n] > 0. ? opout[n] : 0.;
      optm1out[n] = optm1out[n] > 0. ? optm1out[n] : 0.;
    }
  }
  free(opflux); free(dvb); free(csby); free(dbxdx); free(dcsbydy);
  free(dj); free(vni); free(bdotu); free(bvel); free(diffj); free(tp_op); free(opts2); free(djbz);
  free(bd_op); free(bd_opj); free(bd_diff); free(bd_bvel); free(opbm2); free(dopx); free(dopy); free(bd_djbz);
  free(wni); free(dtp); free(dbd); free(explicit_); free(hdz); free(hdzi); free(tp1); free(tphdz0); free(tphdz1);
  free(djint); free(hdzm); free(hdzp); free(diffp); free(diffq); free(diffr); free(pcf); free(qcf); free(rcf);
  free(oploss);
}

/* ===================================================================== */
/* comp_o2o_ng.F90:1-83.  fs is (nk,nlond,nlatd,3,0:3). */
void tgx_o_comp_o2o_ng(const tgx_ogrid *g, const tgx_ofields *F, double *fs) {
  DIMS;
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i)
      for (int k = 0; k < nk; ++k) {
        size_t p = IX(k, i, j);
        double rji, qo2pi, qopi, beta8i, nei;
        if (k < nk - 1) {
          rji = 0.5 * (F->rj[p] + F->rj[p + 1]);
          qo2pi = 0.5 * (F->qo2p[p] + F->qo2p[p + 1]);
          qopi = 0.5 * (F->qop[p] + F->qop[p + 1]);
          beta8i = 0.5 * (F->beta8[p] + F->beta8[p + 1]);
          nei = sqrt(F->ne[p] * F->ne[p + 1]);
        } else {
          rji = 1.5 * F->rj[p] - .5 * F->rj[p - 1];
          qo2pi = 1.5 * F->qo2p[p] - .5 * F->qo2p[p - 1];
          qopi = 1.5 * F->qop[p] - .5 * F->qop[p - 1];
          beta8i = 1.5 * F->beta8[p] - .5 * F->beta8[p - 1];
          nei = sqrt(F->ne[p] * F->ne[p] * F->ne[p] / F->ne[p - 1]);
        }
        double xn = F->xnmbar[p], no = F->no[p], n4s = F->n4s[p], n2d = F->n2d[p], o2p = F->o2p[p];
        double op = F->op[p], nplus = F->nplus[p];
        double pox1 = xn * xn * (F->beta3[p] * n4s / RMN4S * no / RMNO + BETA6 * n2d / RMN2D * no / RMNO) +
                      beta8i * no / RMNO * xn + xn * (RK4 * o2p * n4s / RMN4S + RK10 * op * n2d / RMN2D) +
                      (F->ra1[p] * F->nop[p] + 2. * F->ra2[p] * o2p) * nei;
        double pox2 = xn * (F->beta1[p] * n4s / RMN4S + BETA2 * n2d / RMN2D) + F->rk1[p] * op + RK7 * nplus +
                      (MUT(502) ? 1. : 2.) * rji;
        double lox1 = (MUT(501) ? 1. : 2.) * F->rkm12[p] * xn / F->mbar[p];
        double lox2 = F->rk3[p] * F->n2p[p] + RK8 * nplus;
        double lox3 = qopi;
        double po21 = F->rkm12[p] * xn / F->mbar[p];
        double po23 = RK5 * no / RMNO * o2p * xn;
        double lo21 = xn * (F->beta1[p] * n4s / RMN4S + BETA2 * n2d / RMN2D) + F->rk1[p] * op + (RK6 + RK7) * nplus +
                      RK9 * F->n2p[p] + rji;
        double lo22 = qo2pi;
#define FS(isp, m) fs[p + N3 * ((size_t)(isp) + 3 * (size_t)(m))]
        FS(0, 1) = -lo21;
        FS(0, 2) = MUT(504) ? xn * po21 * F->o1[p] * RIO1 * RMO2 : xn * po21 * F->o1[p] * (RIO1 * RIO1) * RMO2;
        FS(1, 1) = MUT(503) ? pox2 : pox2 * RMO1 * RIO2;
        FS(1, 2) = -lox2 - lox1 * F->o1[p] * RIO1 * xn;
        FS(0, 0) = (po23 - lo22) * RMO2 / xn;
        FS(1, 0) = (pox1 - lox3) * RMO1 / xn;
        FS(0, 3) = 0.;
        FS(1, 3) = 0.;
        for (int m = 0; m < 4; ++m) FS(2, m) = 0.;
#undef FS
      }
}

/* ===================================================================== */
/* comp_ng.F90:1-459 */
void tgx_o_comp_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *fs,
                   const double *tlbc, double *flx_he,
                   double *o2_upd, double *o2nm_upd, double *o1_upd, double *o1nm_upd,
                   double *he_upd, double *henm_upd) {
  DIMS;
  const size_t N2 = (size_t)nlond * nlatd;
  const double dz = g->dz, dtx2inv = g->dtx2inv, thdiffalpha = -0.38, small = 1.e-9;
  const double rmass[3] = {RMO2, RMO1, RMHE};
  /* :72 */
  if (g->calc_helium == 0)
    for (size_t n = 0; n < N2; ++n) flx_he[n] = 0.;
  ALLOC3(o2adv); ALLOC3(o1adv); ALLOC3(headv); ALLOC3(o2s); ALLOC3(o1s); ALLOC3(hes);
  tgx_o_advec(g, F->o2, F->un, F->vn, o2adv);
  tgx_o_advec(g, F->o1, F->un, F->vn, o1adv);
  tgx_o_advec(g, F->he, F->un, F->vn, headv);
  tgx_o_smooth(g, F->o2_nm, o2s, g->shapiro);
  tgx_o_smooth(g, F->o1_nm, o1s, g->shapiro);
  tgx_o_smooth(g, F->he_nm, hes, g->shapiro);
  COLBUF(o2i); COLBUF(o1i); COLBUF(hei); COLBUF(embari); COLBUF(tni); COLBUF(dembardz); COLBUF(dtndz);
  COLBUF(wi);
  double *df = malloc(sizeof(double) * 3 * nk);
  double *gama = malloc(sizeof(double) * 9 * (nk + 1)), *zz = malloc(sizeof(double) * 3 * (nk + 1));
  double *upd = malloc(sizeof(double) * 3 * nk), *n2 = malloc(sizeof(double) * nk);
#define PH(r, c) PHI[(r) - 1][(c) - 1]
  for (int j = 0; j < nlatd; ++j) {
    double dfac = dfactor_of(g, j);
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j), n2d_ = I2(i, j);
      const double *tn = F->tn + o, *o2 = F->o2 + o, *o1 = F->o1 + o, *he = F->he + o, *mbar = F->mbar + o;
      double tl = tlbc[n2d_], fhe = flx_he[n2d_];
      /* :74-97 */
      double ps0[3];
      for (int l = 0; l < 3; ++l) ps0[l] = g->b[l][0] * o2[0] + g->b[l][1] * o1[0] + g->b[l][2] * he[0] + g->fb[l];
      double embar0 = 1. / (ps0[0] * RIO2 + ps0[1] * RIO1 + ps0[2] * RIHE + (1. - ps0[0] - ps0[1] - ps0[2]) * RIN2);
      o2i[0] = .5 * (ps0[0] + o2[0]);
      o1i[0] = .5 * (ps0[1] + o1[0]);
      hei[0] = .5 * (ps0[2] + he[0]);
      embari[0] = .5 * (embar0 + mbar[0]);
      tni[0] = tl;
      dembardz[0] = mbar[0] - embar0;
      dtndz[0] = (tn[0] - tl) * 2;
      for (int k = 1; k < nk; ++k) {
        o2i[k] = .5 * (o2[k - 1] + o2[k]);
        o1i[k] = .5 * (o1[k - 1] + o1[k]);
        hei[k] = .5 * (he[k - 1] + he[k]);
        embari[k] = .5 * (mbar[k - 1] + mbar[k]);
        tni[k] = .5 * (tn[k - 1] + tn[k]);
        dembardz[k] = mbar[k] - mbar[k - 1];
        dtndz[k] = tn[k] - tn[k - 1];
      }
      for (int k = 0; k < nk; ++k) {
        dembardz[k] = dembardz[k] / dz;
        dtndz[k] = dtndz[k] / dz;
      }
      /* the global comp.F takes the top interface's temperature from the top
       * midpoint level below it, tn(lev1-1), in diff_fac (:174-175) and wks1
       * (:446-449) */
      if (g->glb && !MUT(808)) tni[nk - 1] = tn[nk - 2];
      for (int k = 0; k < nk - 1; ++k) wi[k] = .5 * (F->w[o + k] + F->w[o + k + 1]);
      wi[nk - 1] = 1.5 * F->w[o + nk - 1] - .5 * F->w[o + nk - 2];
      /* :124-126 */
      for (int k = 0; k < nk; ++k)
        for (int n = 0; n < 3; ++n) df[3 * k + n] = pow(tni[k] / T00, 1.75 - SS[n]);
      /* alpha(k) of comp_ng.F90:139-151 / :180-192, stored column-major */
#define ALPHA(al, kk)                                                                               \
  do {                                                                                              \
    const double *d_ = df + 3 * (kk);                                                               \
    al[0] = -(PH(1, 4) + (PH(1, 2) - PH(1, 4)) * o1i[kk] + (d_[0] * PH(1, 3) - PH(1, 4)) * hei[kk]); \
    al[4] = -(PH(2, 4) + (PH(2, 1) - PH(2, 4)) * o2i[kk] + (d_[1] * PH(2, 3) - PH(2, 4)) * hei[kk]); \
    al[8] = -(d_[2] * PH(3, 4) + (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * o2i[kk] +                 \
              (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * o1i[kk]);                                     \
    al[3] = (PH(1, 2) - PH(1, 4)) * o2i[kk];                                                        \
    al[6] = (d_[0] * PH(1, 3) - PH(1, 4)) * o2i[kk];                                                \
    al[1] = (PH(2, 1) - PH(2, 4)) * o1i[kk];                                                        \
    al[7] = (d_[1] * PH(2, 3) - PH(2, 4)) * o1i[kk];                                                \
    al[2] = (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * hei[kk];                                        \
    al[5] = (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * hei[kk];                                        \
  } while (0)
      /* :128-166 */
      double ep[2][3], ak[2][9], alpha[9];
      double wks3, wks4 = dembardz[0] / (embari[0] * 2.);
      int km = 0, kp = 1;
      for (int l = 0; l < 3; ++l) ep[kp][l] = 1. - 1. / embari[0] * (rmass[l] + dembardz[0]);
      ep[kp][2] = ep[kp][2] - thdiffalpha * dtndz[0] / tni[0];
      ALPHA(alpha, 0);
      tgx_o_matinv3(alpha, ak[kp]);
      double wks1 = embari[0] * RIN2 * pow(T00 / tl, 0.25) / TAU;
      for (int n = 0; n < 9; ++n) ak[kp][n] = ak[kp][n] * wks1;
      for (int n = 0; n < 9; ++n) gama[n] = 0.;
      for (int n = 0; n < 3; ++n) zz[n] = 0.;
      double alpha22 = 0, alpha33 = 0, alpha23 = 0, alpha32 = 0, flx00 = 0;
      /* :170-315 */
      for (int k = 0; k < nk - 1; ++k) {
        int t = km; km = kp; kp = t;
        for (int l = 0; l < 3; ++l) ep[kp][l] = 1. - 1. / embari[k + 1] * (rmass[l] + dembardz[k + 1]);
        if (k != nk - 2) ep[kp][2] = ep[kp][2] - (MUT(604) ? -thdiffalpha : thdiffalpha) * dtndz[k + 1] / tni[k + 1];
        ALPHA(alpha, k + 1);
        tgx_o_matinv3(alpha, ak[kp]);
        wks1 = embari[k + 1] * RIN2 * pow(T00 / tni[k + 1], MUT(602) ? 0.5 : 0.25) / TAU;
        wks3 = wks4;
        wks4 = dembardz[k + 1] / (embari[k + 1] * 2.);
        for (int n = 0; n < 9; ++n) ak[kp][n] = ak[kp][n] * wks1;
        double molp[9], molr[9], molq[9], pk[9], rk[9], qk[9], fk[3];
        for (int isp = 0; isp < 3; ++isp)
          for (int m = 0; m < 3; ++m) {
            int e = isp + 3 * m;
            molp[e] = ak[km][e] * (1. / dz + ep[km][m] / 2.);
            molr[e] = ak[kp][e] * (1. / dz - ep[MUT(606) ? km : kp][m] / 2.);
            molq[e] = ak[km][e] * (1. / dz - ep[km][m] / 2.) + ak[kp][e] * (1. / dz + ep[kp][m] / 2.);
          }
        double eddyp = g->expzmid_inv * g->difk[k] * (1. / dz - wks3);
        double eddyr = g->expzmid * g->difk[k + 1] * (1. / dz + wks4);
        double eddyq = g->expzmid_inv * g->difk[k] * (1. / dz + (MUT(601) ? -wks3 : wks3)) +
                       g->expzmid * g->difk[k + 1] * (1. / dz - wks4);
        size_t p = o + k;
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp) {
            int e = isp + 3 * m;
            double delta = isp == m ? 1. : 0.;
            pk[e] = molp[e] / dz - g->expz[k] * delta / dz * (eddyp * dfac + .5 * wi[k]);
            rk[e] = molr[e] / dz - g->expz[k] * delta / dz * (eddyr * dfac - .5 * wi[k]);
            qk[e] = -molq[e] / dz + g->expz[k] * delta * (eddyq * dfac / dz + dtx2inv) -
                    g->expz[k] * fs[p + N3 * ((size_t)isp + 3 * (size_t)(m + 1))];
          }
        fk[0] = g->expz[k] * (o2s[p] * dtx2inv - (o2adv[p] - fs[p + N3 * 0]) + (MUT(605) ? 0. : F->hdo2[p]));
        fk[1] = g->expz[k] * (o1s[p] * dtx2inv - (o1adv[p] - fs[p + N3 * 1]) + F->hdo1[p]);
        fk[2] = g->expz[k] * (hes[p] * dtx2inv - (headv[p] - fs[p + N3 * 2]) + F->hdhe[p]);
        if (k == 0) {
          for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n)
              for (int isp = 0; isp < 3; ++isp)
                qk[isp + 3 * m] = qk[isp + 3 * m] + pk[isp + 3 * n] * (MUT(603) ? g->b[m][n] : g->b[n][m]);
          for (int m = 0; m < 3; ++m)
            for (int isp = 0; isp < 3; ++isp) fk[isp] = fk[isp] - pk[isp + 3 * m] * g->fb[m];
          for (int n = 0; n < 9; ++n) pk[n] = 0.;
        } else if (k == nk - 2) {
          for (int m = 0; m < 3; ++m)
            for (int isp = 0; isp < 3; ++isp)
              qk[isp + 3 * m] = qk[isp + 3 * m] + (1. + .5 * ep[kp][m] * dz) / (1. - .5 * ep[kp][m] * dz) * rk[isp + 3 * m];
          const double *d_ = df + 3 * (k + 1);
          alpha22 = -(PH(2, 4) + (PH(2, 1) - PH(2, 4)) * o2i[k + 1] + (d_[1] * PH(2, 3) - PH(2, 4)) * hei[k + 1]);
          alpha33 = -(d_[2] * PH(3, 4) + (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * o2i[k + 1] +
                      (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * o1i[k + 1]);
          alpha23 = (d_[1] * PH(2, 3) - PH(2, 4)) * o1i[k + 1];
          alpha32 = (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * hei[k + 1];
          /* the global comp.F:632-633 averages the mean mass over levels k and
           * k+1 (the top interface), comp_ng.F90 over k-1 and k */
          flx00 = (g->glb && !MUT(809) ? embari[k + 1] : embari[k]) * RIN2 * P0 * pow(T00 / tn[k], 0.25) /
                  (TAU * GRAV);
          for (int isp = 0; isp < 3; ++isp)
            fk[isp] = fk[isp] - rk[isp + 3 * 1] * (alpha23 - alpha22) * fhe / (flx00 * (1. / dz - 0.5 * ep[kp][1])) -
                      rk[isp + 3 * 2] * (alpha33 - alpha32) * fhe / (flx00 * (1. / dz - 0.5 * ep[kp][2]));
          for (int n = 0; n < 9; ++n) rk[n] = 0.;
        }
        if (g_cap)
          for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
              CAPW(38 + 3 * r + c, n2d_, k, pk[r + 3 * c]);
              CAPW(47 + 3 * r + c, n2d_, k, qk[r + 3 * c]);
              CAPW(56 + 3 * r + c, n2d_, k, rk[r + 3 * c]);
            }
            CAPW(65 + r, n2d_, k, fk[r]);
          }
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n)
            for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] - pk[isp + 3 * n] * gama[9 * k + n + 3 * m];
        double wkm1[9], wkv1[3];
        tgx_o_matinv3(qk, wkm1);
        for (int isp = 0; isp < 3; ++isp) wkv1[isp] = fk[isp];
        for (int n = 0; n < 9; ++n) gama[9 * (k + 1) + n] = 0.;
        for (int m = 0; m < 3; ++m) {
          for (int isp = 0; isp < 3; ++isp) wkv1[isp] = wkv1[isp] - pk[isp + 3 * m] * zz[3 * k + m];
          for (int n = 0; n < 3; ++n)
            for (int isp = 0; isp < 3; ++isp)
              gama[9 * (k + 1) + isp + 3 * m] = gama[9 * (k + 1) + isp + 3 * m] + wkm1[isp + 3 * n] * rk[n + 3 * m];
        }
        for (int isp = 0; isp < 3; ++isp) zz[3 * (k + 1) + isp] = 0.;
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp) zz[3 * (k + 1) + isp] = zz[3 * (k + 1) + isp] + wkm1[isp + 3 * m] * wkv1[m];
      }
      /* :317-332 */
      for (int isp = 0; isp < 3; ++isp) upd[3 * (nk - 1) + isp] = 0.;
      for (int k = nk - 2; k >= 0; --k)
        for (int isp = 0; isp < 3; ++isp) {
          double u = zz[3 * (k + 1) + isp];
          for (int m = 0; m < 3; ++m) u = u - gama[9 * (k + 1) + isp + 3 * m] * upd[3 * (k + 1) + m];
          upd[3 * k + isp] = u;
        }
      for (int isp = 0; isp < 3; ++isp)
        upd[3 * (nk - 1) + isp] = (1. + .5 * ep[kp][isp] * dz) / (1. - .5 * ep[kp][isp] * dz) * upd[3 * (nk - 2) + isp];
      upd[3 * (nk - 1) + 1] = upd[3 * (nk - 1) + 1] + (alpha23 - alpha22) * fhe / (flx00 * (1. / dz - 0.5 * ep[kp][1]));
      upd[3 * (nk - 1) + 2] = upd[3 * (nk - 1) + 2] + (alpha33 - alpha32) * fhe / (flx00 * (1. / dz - 0.5 * ep[kp][2]));
      for (int k = 0; k < nk; ++k) {
        o2_upd[o + k] = upd[3 * k + 0];
        o1_upd[o + k] = upd[3 * k + 1];
        he_upd[o + k] = upd[3 * k + 2];
      }
    }
  }
#undef ALPHA
#undef PH
  /* :386-417 */
  if (g->calc_helium == 0)
    for (size_t n = 0; n < N3; ++n) he_upd[n] = g->pshelb;
  if (g->glb) {
    /* comp.F:932-934 filter_o2o (order 1) of O2, O and He */
    glb_filter(g, o2_upd, nk, 1, 0);
    glb_filter(g, o1_upd, nk, 1, 0);
    glb_filter(g, he_upd, nk, 1, 0);
  } else {
    lateral(o2_upd, F->lonb_o2, F->latb_o2, nk, nlond, nlatd);
    lateral(o1_upd, F->lonb_o1, F->latb_o1, nk, nlond, nlatd);
    lateral(he_upd, F->lonb_he, F->latb_he, nk, nlond, nlatd);
  }
  for (size_t n = 0; n < N3; ++n) {
    o2nm_upd[n] = DTSMOOTH_DIV2 * (F->o2_nm[n] + o2_upd[n]) + DTSMOOTH * F->o2[n];
    o1nm_upd[n] = DTSMOOTH_DIV2 * (F->o1_nm[n] + o1_upd[n]) + DTSMOOTH * F->o1[n];
    henm_upd[n] = DTSMOOTH_DIV2 * (F->he_nm[n] + he_upd[n]) + DTSMOOTH * F->he[n];
  }
  /* :419-457 */
  for (int j = 0; j < nlatd; ++j)
    for (int i = 0; i < nlond; ++i) {
      size_t o = IX(0, i, j);
      for (int k = 0; k < nk; ++k) n2[k] = 1. - o2_upd[o + k] - o1_upd[o + k] - he_upd[o + k];
      int ktop;
      for (ktop = 0; ktop < nk; ++ktop)
        if (n2[ktop] < 0.01) break;
      /* Fortran ktop = C ktop + 1; the reference reads n2(ktop-1), so ktop=1
       * (C 0) would read outside the column; inputs must keep ktop > 1. */
      if (ktop + 1 < nk && ktop >= 1) {
        double zmtop = g->zpmid[ktop];
        double n2top = log(n2[ktop] * F->xnmbar[o + ktop]);
        double n2top1 = log(n2[ktop - 1] * F->xnmbar[o + ktop - 1]);
        double dn2dz = (n2top - n2top1) / dz;
        for (int k = ktop + 1; k < nk; ++k)
          n2[k] = exp(n2top + dn2dz * (g->zpmid[k] - zmtop)) / F->xnmbar[o + k];
      }
      for (int k = 0; k < nk; ++k) {
        size_t p = o + k;
        double a = o2_upd[p] > small ? o2_upd[p] : small;
        double b = o1_upd[p] > small ? o1_upd[p] : small;
        double c = he_upd[p] > small ? he_upd[p] : small;
        double d = n2[k] > small ? n2[k] : small;
        double norm = a + b + c + d;
        o2_upd[p] = a / norm;
        o1_upd[p] = b / norm;
        he_upd[p] = c / norm;
        double an = o2nm_upd[p] > small ? o2nm_upd[p] : small;
        double bn = o1nm_upd[p] > small ? o1nm_upd[p] : small;
        double cn = henm_upd[p] > small ? henm_upd[p] : small;
        double n2nm = 1. - o2nm_upd[p] - o1nm_upd[p] - henm_upd[p];
        double dn = n2nm > small ? n2nm : small;
        double normn = an + bn + cn + dn;
        o2nm_upd[p] = an / normn;
        o1nm_upd[p] = bn / normn;
        henm_upd[p] = cn / normn;
      }
    }
  if (g->glb) {
    glb_periodic(g, o2_upd); glb_periodic(g, o1_upd); glb_periodic(g, he_upd);
    glb_periodic(g, o2nm_upd); glb_periodic(g, o1nm_upd); glb_periodic(g, henm_upd);
  }
  free(o2adv); free(o1adv); free(headv); free(o2s); free(o1s); free(hes);
  free(o2i); free(o1i); free(hei); free(embari); free(tni); free(dembardz); free(dtndz); free(wi);
  free(df); free(gama); free(zz); free(upd); free(n2);
}
