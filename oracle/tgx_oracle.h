/*
 * CPU FP64 oracle for the TIE-GCM NG implicit vertical-column step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker.
 *
 * Every routine restates the reference algorithm in plain C, in the reference's
 * arithmetic order (left-to-right evaluation, no FMA contraction: build with
 * -ffp-contract=off).  Each function cites the reference file:line it follows.
 *
 * Layout: the reference's Fortran layout, level fastest.  A 3-D nest field is
 * f(nk, lond0:lond1, latd0:latd1) with lond0 = lon0-2, lond1 = lon1+2 (two ghost
 * columns each side, fields_ng.F90:137-140).  Index (k, i, j) is 0-based:
 *   k in [0, nk), i in [0, nlon+4), j in [0, nlat+4);  flat = k + nk*(i + nlond*j).
 * Fortran lon index l maps to i = l + 1 (lond0 = -1 -> 0); likewise for lat.
 *
 * Domain semantics: one nest domain owned by one task, so every side is a
 * domain edge (is_bndry = .true. on all four sides) and the halo exchanges in
 * sync_ng.F90 are no-ops.  This is exactly the reference running that nest on
 * one MPI task.
 */
#ifndef TGX_ORACLE_H
#define TGX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Grid, level profiles and run switches of one nest (fields_ng.F90:73-84,
 * set_const_fields_ng.F90:25-74, input switches used by the column step). */
typedef struct {
  int nk;            /* nlevp1_ng */
  int nlon, nlat;    /* interior points; arrays carry nlon+4 x nlat+4 columns */
  double dz;         /* dz(i_ng) = dlev_ng */
  double dlamda, dphi;
  double dtx2inv;    /* 1/(2*step_ng) */
  double shapiro;
  double expzmid, expzmid_inv;
  const double *expz;   /* [nk] */
  const double *difk;   /* [nk] at the run's iday */
  const double *dift;   /* [nk] */
  const double *xmue;   /* [nk] */
  const double *zpint;  /* [nk] interface levels (oplus opdiffcap) */
  const double *zpmid;  /* [nk] midpoint levels (comp N2 extrapolation) */
  const double *racs;   /* [nlat+4] 1/(re*cos(lat)) */
  const double *cs;     /* [nlat+4] cos(lat) */
  const double *tanphi; /* [nlat+4] */
  const double *cor;    /* [nlat+4] 2*omega*sin(lat) */
  const double *glat;   /* [nlat+4] geographic latitude, degrees */
  int nstep_sub;        /* O+ sub-cycles */
  int calc_helium;
  double colfac, opdiffcap, difhor;
  double b[3][3];       /* lbc b(l,m) stored b[l-1][m-1] (lbc.F:527-594) */
  double fb[3];
  double pshelb;
  /* Global grid (dt.F, duv.F, minor.F, oplus.F, comp_o2o.F, comp.F) instead of
   * a nest: glb != 0.  The ghost columns then hold the periodic points (i = 0,
   * 1 copies of i = nlon, nlon+1; i = nlon+2, nlon+3 copies of i = 2, 3) and
   * the ghost rows the latitudes across the poles, both filled by the caller
   * as the model fills them before the step (mk_polelat.F, mp_periodic_f4d);
   * no lateral Dirichlet data are applied, the solutions are ring-filtered
   * (ringfilter.F with the chunk table chunk[nlat_filter], cons.F:331-355)
   * before their time smoothing, and every output's periodic points are set
   * from the interior (mp_periodic_f3d).  Outputs on the ghost rows are not
   * defined (the model recomputes them from the interior). */
  int glb;
  int nlat_filter;
  const int *chunk;
} tgx_ogrid;

/* Model state read by the column step.  All 3-D fields are (nk, nlond, nlatd),
 * 2-D fields are (nlond, nlatd).  Suffix-free names are the itp time level,
 * "_upd" names are itc (dynamics_ng.F90:203-233). */
typedef struct {
  /* 3-D prognostic fields at itp */
  const double *tn, *tn_nm, *un, *vn, *un_nm, *vn_nm;
  const double *o2, *o2_nm, *o1, *o1_nm, *he, *he_nm;
  const double *op, *op_nm, *n2d, *n2d_nm, *n4s, *n4s_nm, *no, *no_nm, *ar, *ar_nm;
  const double *ti, *te, *ne, *o2p, *w, *z;
  const double *mbar, *barm, *xnmbar, *xnmbari, *scht, *schti, *n2;
  /* 3-D fields at itc computed earlier in the step */
  const double *w_upd;      /* swdot_ng output, flds%w(itc) */
  /* 3-D diagnostic/auxiliary inputs produced by the explicit physics */
  const double *cp, *kt, *km, *hdt, *hdu, *hdv, *hdo2, *hdo1, *hdhe;
  const double *qji_tn, *cool_imp, *cool_exp, *qtotal, *rkm12;
  const double *ui, *vi, *wi, *lxx, *lyy, *lxy, *lyx;
  const double *qop2p, *qop2d, *qop, *qo2p, *qtef, *rj;
  const double *rk1, *rk2, *rk3, *rk19, *rk20, *rk25;
  const double *beta1, *beta3, *beta5, *beta8, *beta9, *beta9n, *beta17;
  const double *ra1, *ra2, *ra3, *n2p, *nplus, *nop;
  /* 2-D fields */
  const double *t_lbc, *u_lbc, *v_lbc;          /* current lower boundary values */
  const double *ulbc, *vlbc;                    /* flds%ulbc/vlbc as dt_ng sees them (set by the previous duv_ng) */
  const double *bx, *by, *bz, *bmod2, *dipmag, *sndec, *csdec, *chi, *rlatm;
  /* lateral boundary data: lon_b (nk, 2, nlatd), lat_b (nk, nlond, 2) per field */
  const double *lonb_tn, *latb_tn, *lonb_un, *latb_un, *lonb_vn, *latb_vn;
  const double *lonb_o2, *latb_o2, *lonb_o1, *latb_o1, *lonb_he, *latb_he;
  const double *lonb_n2d, *latb_n2d, *lonb_n4s, *latb_n4s, *lonb_no, *latb_no;
  const double *lonb_ar, *latb_ar, *lonb_op, *latb_op;
  double ar_glbm1, ar_glbm2;  /* ar_glbm(1:2), comp_ar.F lower boundary */
} tgx_ofields;

/* ---- solvers -------------------------------------------------------- */
void tgx_o_trsolv(const double *a, const double *b, const double *c, const double *f,
                  double *x, int lev0, int lev1, int k1, int k2, int lon0, int lon1,
                  int lonmax);
void tgx_o_trsolv_ng(const double *a, const double *b, const double *c, const double *f,
                     double *x, int nk, int ncol, int k2);
void tgx_o_blktri(const double *a, const double *b, const double *c, const double *f,
                  double *x, int lev0, int lev1, int lon0, int lon1, int lonmax);
void tgx_o_blktri_ng(const double *pp, const double *qq, const double *rr,
                     const double *ss, double *xx, int nk, int ncol);
void tgx_o_blktri3(const double *p, const double *q, const double *r, const double *f,
                   double *x, int nk, int ncol, int k2);
void tgx_o_matinv3(const double *a, double *b);  /* 3x3, column-major like Fortran */
void tgx_o_matinv2(const double *a, double *b);
void tgx_o_bndcmp(double dz, double b[3][3], double fb[3]);

/* ---- horizontal stencils (dffm_ng.F90, advec_ng.F90, smooth_ng.F90, div_ng.F90) */
void tgx_o_df3d(const double *f, double *df, int order, int nk, int nlond, int nlatd);
void tgx_o_fm3d(const double *f, double *fm, int order, int nk, int nlond, int nlatd);
void tgx_o_df2d(const double *f, double *df, int order, int nlond, int nlatd);
void tgx_o_advec(const tgx_ogrid *g, const double *f, const double *un, const double *vn,
                 double *hadvec);
void tgx_o_smooth(const tgx_ogrid *g, const double *fin, double *fout, double shapiro);
void tgx_o_div(const tgx_ogrid *g, const double *f, double *df, int dir);

/* ---- column-step routines -------------------------------------------
 * Each routine follows its nest form (the _ng file) when g->glb == 0 and
 * its global form (dt.F, duv.F, minor.F + comp_{no,n4s,n2d,ar}.F, oplus.F,
 * comp_o2o.F, comp.F) when g->glb != 0; the global forms' differences are
 * cited where they are made. */
void tgx_o_dt_ng(const tgx_ogrid *g, const tgx_ofields *F, double *tlbc, double *tlbc_nm,
                 double *tn_upd, double *tn_nm_upd);
void tgx_o_duv_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *tn_upd,
                  const double *Fe, const double *Fn, const double *tlbc_nm,
                  double *ulbc, double *vlbc, double *ulbc_nm, double *vlbc_nm,
                  double *un_upd, double *unm_upd, double *vn_upd, double *vnm_upd);
/* Species: 0=NO 1=N4S 2=N2D 3=AR.  n2d_upd is the N2D solution of this step
 * (used by NO and N4S); tlbc is flds%tlbc as seen by the minors. */
void tgx_o_minor_species(const tgx_ogrid *g, const tgx_ofields *F, int species,
                         const double *n2d_upd, const double *tlbc,
                         double *out, double *out_nm);
void tgx_o_oplus_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *op,
                    const double *optm1, const double *lonb, const double *latb,
                    double *opout, double *optm1out, double *xiop2p, double *xiop2d,
                    double *Fe, double *Fn);
void tgx_o_comp_o2o_ng(const tgx_ogrid *g, const tgx_ofields *F, double *fs);
void tgx_o_comp_ng(const tgx_ogrid *g, const tgx_ofields *F, const double *fs,
                   const double *tlbc, double *flx_he,
                   double *o2_upd, double *o2nm_upd, double *o1_upd, double *o1nm_upd,
                   double *he_upd, double *henm_upd);

/* ---- ring filter (ringfilter.c) -------------------------------------------
 * ringfilter.F:1-111 on one latitude row f(nlon+4, nlev), longitude fastest;
 * returns -1 if order >= nlon.  tgx_o_ringfilter_chunks: the chunk count of
 * 1-based latitude lat from the table chunk[nlat_filter] (0: not filtered). */
int tgx_o_ringfilter_row(double *f, int nlon, int nlev, int order, int oplus, int nchnk, double dlamda);
int tgx_o_ringfilter_chunks(int lat, int nlat, int nlat_filter, const int *chunk);

/* ---- interpolation between nest levels (interp.c) ----------------------
 * interp.F90:1-128 and interp_fields_ng.F90:17-63, Fortran layouts:
 * tgx_o_interp_find is find (1-based, 0 below, nx at or above the range);
 * interp1d f(nx) from fp(nxp); interp2d f(nx, ny) from fp(nxp, nyp);
 * interp3d f(nz, nx, ny) from fp(nzp, nxp, nyp), log-linear in z when zlog;
 * time_interp: f(n, tb-ta+1) at t = ta..tb between f0 (t = 0) and f1 (t = t1). */
int tgx_o_interp_find(const double *xp, int nx, double x0);
void tgx_o_interp1d(const double *x, int nx, const double *xp, int nxp, const double *fp, double *f);
void tgx_o_interp2d(const double *x, int nx, const double *y, int ny, const double *xp, int nxp, const double *yp,
                    int nyp, const double *fp, double *f);
void tgx_o_interp3d(const double *z, int nz, const double *x, int nx, const double *y, int ny, const double *zp,
                    int nzp, const double *xp, int nxp, const double *yp, int nyp, const double *fp, int zlog,
                    double *f);
void tgx_o_time_interp(const double *f0, const double *f1, long n, int t1, int ta, int tb, double *f);

/* ---- test hooks (column.c) ---------------------------------------------
 * tgx_o_set_capture: rows handed to each column solve are written into cap
 * (TGX_O_CAPTURE_PLANES planes of ncol*nk doubles, the library's capture
 * layout), or not at all with NULL.  tgx_o_set_mutation: a nonzero id plants
 * one deliberate mistake (the MUT sites of column.c); 0 restores the oracle. */
#define TGX_O_CAPTURE_PLANES 68
void tgx_o_set_capture(double *cap);
void tgx_o_set_mutation(int id);

#ifdef __cplusplus
}
#endif
#endif
