/*
 * libtgx — fused implicit column step of one nest domain, device resident.
 *
 * Replaces the assembly+solve of dt_ng (dt_ng.F90:1), duv_ng (duv_ng.F90:1),
 * minor_ng with comp_{n2d,n4s,no,ar}_ng (minor_ng.F90:1, comp_*_ng.F90),
 * oplus_ng with its nstep_sub sub-cycles (oplus_ng.F90:1,
 * dynamics_ng.F90:99-121), and comp_o2o_ng + comp_ng (comp_o2o_ng.F90:1,
 * comp_ng.F90:1).  Semantics are those of the reference running the nest on
 * one task: all four sides are domain edges (lateral Dirichlet data applied,
 * one-sided stencils at the ghost points).  addfld diagnostics are not
 * produced.
 *
 * Device layout ("level planes", longitude fastest): element (k, i, j) of a
 * 3-D field is at k*plane + j*ld + i with plane = ld*(nlat+4), i in
 * [0, nlon+4) (i = Fortran lon + 1), j in [0, nlat+4), k in [0, nk);
 * 2-D fields are at j*ld + i.  ld >= nlon+4 is the row pitch in elements.
 * Lateral boundary data: lonb[(k*2 + side)*(nlat+4) + j] (side 0 = lon -1,
 * side 1 = lon nlon+2), latb[(k*2 + side)*ld + i] (side 0 = lat -1, side 1 =
 * lat nlat+2); the O+ data carries one such block per sub-step.
 *
 * All pointers are device pointers.  Every entry point is stream-ordered on
 * the hipStream_t passed as void* and returns 0 or a TGX_ERR_* code.
 */
#ifndef TGX_STEP_H
#define TGX_STEP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nk, nlon, nlat, ld;
  int nstep_sub;      /* O+ sub-cycles per step (input namelist NSTEP_SUB) */
  int calc_helium;
  int scheme;         /* column solve: TGX_SCHEME_THOMAS (thread per column, the
                         reference's operation order), TGX_SCHEME_PCR (one
                         wavefront per column, cyclic reduction, nk <= 128) or
                         TGX_SCHEME_AUTO (the faster one, timed once per shape) */
  double dz;          /* dz(i_ng) = dlev_ng */
  double dlamda, dphi, dtx2inv, shapiro, expzmid, expzmid_inv;
  double colfac, opdiffcap, difhor, pshelb, ar_glbm1, ar_glbm2;
  double b[9];        /* lbc b(l,m), row-major b[3*(l-1)+(m-1)] */
  double fb[3];
  const double *expz, *difk, *dift, *xmue, *zpint, *zpmid;   /* [nk] */
  const double *racs, *cs, *tanphi, *cor, *glat;             /* [nlat+4] */
  /* Optional (NULL in normal use): when set, the Thomas column kernels also
   * write the rows they assemble, for analysis of the column systems.  Layout:
   * capture[(plane*ncol + col)*nk + k], col = j*(nlon+4) + i, planes
   *   0-3 dt a,b,c,f;  4-7 O+ a,b,c,f (last sub-step);  8+4s..11+4s minor
   *   species s (0 NO, 1 N4S, 2 N2D, 3 Ar) a,b,c,f;  24-27 duv A, 28-31 B,
   *   32-35 C (2x2 row-major), 36-37 F;  38-46 comp A, 47-55 B, 56-64 C (3x3
   *   row-major), 65-67 F.  a x(k-1) + b x(k) + c x(k+1) = f, rows 0..nk-2
   *   (minor species 0..nk-1); lateral-boundary columns are not written. */
  double *capture;
} tgx_ng_grid;

#define TGX_NG_CAPTURE_PLANES 68

/* Model state read by the step (itp level unless noted). */
typedef struct {
  const double *tn, *tn_nm, *un, *vn, *un_nm, *vn_nm;
  const double *o2, *o2_nm, *o1, *o1_nm, *he, *he_nm;
  const double *op, *op_nm, *n2d, *n2d_nm, *n4s, *n4s_nm, *no, *no_nm, *ar, *ar_nm;
  const double *ti, *te, *ne, *o2p, *w, *z;
  const double *mbar, *barm, *xnmbar, *xnmbari, *scht, *schti, *n2;
  const double *w_upd;                         /* flds%w(itc) from swdot_ng */
  const double *cp, *kt, *km, *hdt, *hdu, *hdv, *hdo2, *hdo1, *hdhe;
  const double *qji_tn, *cool_imp, *cool_exp, *qtotal, *rkm12;
  const double *ui, *vi, *wi, *lxx, *lyy, *lxy, *lyx;
  const double *qop2p, *qop2d, *qop, *qo2p, *qtef, *rj;
  const double *rk1, *rk2, *rk3, *rk19, *rk20, *rk25;
  const double *beta1, *beta3, *beta5, *beta8, *beta9, *beta9n, *beta17;
  const double *ra1, *ra2, *ra3, *n2p, *nplus, *nop;
  /* 2-D */
  const double *t_lbc, *u_lbc, *v_lbc;   /* this step's lower-boundary values */
  const double *tlbc, *ulbc, *vlbc;      /* flds%tlbc/ulbc/vlbc at the start of the step */
  const double *flx_he;
  const double *bx, *by, *bz, *bmod2, *dipmag, *sndec, *csdec, *chi, *rlatm;
  /* lateral Dirichlet data */
  const double *lonb_tn, *latb_tn, *lonb_un, *latb_un, *lonb_vn, *latb_vn;
  const double *lonb_o2, *latb_o2, *lonb_o1, *latb_o1, *lonb_he, *latb_he;
  const double *lonb_n2d, *latb_n2d, *lonb_n4s, *latb_n4s, *lonb_no, *latb_no;
  const double *lonb_ar, *latb_ar, *lonb_op, *latb_op;
} tgx_ng_fields;

/* Outputs (itc level).  2-D: tlbc/tlbc_nm as dt_ng leaves them
 * (dt_ng.F90:210-211), ulbc..vlbc_nm as duv_ng leaves them
 * (duv_ng.F90:315-318), flx_he as comp_ng leaves it (comp_ng.F90:72). */
typedef struct {
  double *tn, *tn_nm;
  double *un, *un_nm, *vn, *vn_nm;
  double *n2d, *n2d_nm, *n4s, *n4s_nm, *no, *no_nm, *ar, *ar_nm;
  double *op, *op_nm, *xiop2p, *xiop2d, *Fe, *Fn;
  double *o2, *o2_nm, *o1, *o1_nm, *he, *he_nm;
  double *tlbc, *tlbc_nm, *ulbc, *vlbc, *ulbc_nm, *vlbc_nm, *flx_he;
} tgx_ng_out;

/* Bytes of device workspace the step needs for this grid (Thomas ratios,
 * block gammas, O+ and wind intermediates). */
unsigned long long tgx_ng_workspace_bytes(const tgx_ng_grid *g);

/* The routines, each on its own; work is a device buffer of
 * tgx_ng_workspace_bytes bytes. */
int tgx_ng_oplus(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                 void *stream);
int tgx_ng_minors(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                  void *stream);
int tgx_ng_dt(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
              void *stream);
/* needs out->tn (dt_ng) and out->Fe/Fn (oplus_ng) of the same step */
int tgx_ng_duv(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
               void *stream);
int tgx_ng_comp(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                void *stream);
/* The precision form of tgx_ng_comp (no counterpart in the reference): the
 * 3x3 block rows assembled and eliminated in FP32 with the block pivots
 * inverted in FP64, then nsweep (0..2) iterative-refinement sweeps, each
 * forming the residual with the rows assembled again in FP64 and correcting
 * with the FP32 elimination.  Same arguments and outputs as tgx_ng_comp; not
 * for capture mode (g->capture must be NULL). */
int tgx_ng_comp_mixed(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                      void *stream, int nsweep);
/* tgx_ng_step with comp in that precision form (nsweep 0..2; no counterpart
 * in the reference): the other routines as in tgx_ng_step.  Nest domains
 * only, g->capture NULL; may be captured into a hipGraph. */
int tgx_ng_step_mixed(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                      void *stream, int nsweep);
/* The whole column step (dynamics_ng.F90:99-233).  The independent routines
 * (oplus, minors, dt, comp) run concurrently on library-owned streams forked
 * from and joined back into `stream`; duv follows dt and oplus.  Results are
 * identical to the serial order; the call is stream-ordered on `stream` and
 * may be captured into a hipGraph. */
int tgx_ng_step(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                void *stream);
/* Autotuner of TGX_SCHEME_AUTO: times each routine on `stream`
 * (host-synchronising; not for graph capture) under the Thomas scheme and
 * under the wavefront-per-column scheme with 1, 2 and 4 columns per block, and
 * keeps the fastest per device, shape and routine.  tgx_ng_step calls it on
 * the first AUTO step of a shape.  tgx_ng_scheme_choice reports the scheme a
 * routine (0..4 = oplus, minors, dt, duv, comp) runs with for this grid, and
 * tgx_ng_wave_columns_choice its columns per block when that is
 * TGX_SCHEME_PCR (4 unless tuned).  No reference counterpart. */
int tgx_ng_autotune(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                    void *stream);
int tgx_ng_scheme_choice(const tgx_ng_grid *g, int routine);
int tgx_ng_wave_columns_choice(const tgx_ng_grid *g, int routine);
/* Columns per block (1, 2 or 4) of the wavefront-per-column kernels launched
 * from the calling host thread, in place of the autotuner's choice; 0 returns
 * to the choice. */
void tgx_ng_force_wave_columns(int cols);
/* Whether the TN column kernel may load its LDS tiles straight from HBM
 * (on, the default, where TN, UN, VN and TN_NM are 16-byte aligned and g.ld
 * is even) or stages them through registers (0), for launches from the
 * calling host thread.  Both give the same results.  No reference counterpart. */
void tgx_ng_set_tile_dma(int on);
/* The same step strictly in the reference's order on `stream` alone. */
int tgx_ng_step_serial(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                       void *work, void *stream);

/* Timing hooks (no counterpart in the reference; used by benchmarks and
 * profiling).  Between begin and end, each routine entry point records a
 * hipEvent pair on the stream it runs on around the whole routine and another
 * around its column kernel (oplus: the O+ solve of each sub-step; minors: the
 * N2D + Ar solve), for at most max_calls calls each.  end waits for the
 * recorded events and returns the summed milliseconds and the number of calls
 * timed: entries 0..4 the routines oplus, minors, dt, duv, comp, entries 5..9
 * their column kernels in the same order. */
int tgx_ng_profile_begin(int max_calls);
int tgx_ng_profile_end(double *ms_total /* [10] */, int *calls /* [10] */);

/* ------------------------------------------------------------------------ */
/* The model's own layout                                                    */

/* Level-fastest model arrays <-> level planes, on the device.  The model
 * holds a 3-D nest field as f(nk, lond0:lond1, latd0:latd1), level fastest
 * (fields_ng.F90:226-261); src/dst of that shape are nk*nlond*nlatd doubles,
 * the planes nk*ld*nlatd (pitch ld >= nlond).  One launch, LDS-tiled, on
 * `stream`; nk <= 240. */
int tgx_ng_pack_levels(const double *src, double *dst, int nk, int nlond, int nlatd, int ld, void *stream);
int tgx_ng_unpack_levels(const double *src, double *dst, int nk, int nlond, int nlatd, int ld, void *stream);

/* ------------------------------------------------------------------------ */
/* Coupling between nest levels                                              */

/* interp3d (interp.F90:6-48) on the device, as map_in, set_bndry and map_out
 * use it (levels_inner_ng.F90:131-410): the source field in level planes
 * (src[k*ld*nyp + j*ld + i], nzp x nyp rows of nxp values) on the axes zp,
 * xp, yp (increasing), interpolated to the target points (z[k], x[i], y[j]),
 * k < nz, i < nx, j < ny, written to dst[k*dk + i*di + j*dj]; log-linear in z
 * when zlog (the reference's zlog fields).  Bilinear in (x, y) from the
 * bracketing source columns, then linear in z, the end intervals
 * extrapolated, in the reference's order of operations.  All arrays,
 * coordinates included, are device arrays; one launch on `stream`. */
int tgx_ng_interp3d_dev(const double *src, int nzp, int nxp, int nyp, int ld, const double *zp,
                        const double *xp, const double *yp, double *dst, int nz, int nx, int ny, long dk, long di,
                        long dj, const double *z, const double *x, const double *y, int zlog, void *stream);
/* interp2d (interp.F90:50-70) on one plane src[j*ld + i], to dst[i*di + j*dj]. */
int tgx_ng_interp2d_dev(const double *src, int nxp, int nyp, int ld, const double *xp, const double *yp,
                        double *dst, int nx, int ny, long di, long dj, const double *x, const double *y,
                        void *stream);
/* interp_fields_ng.F90:17-63: the n values at t = ta..tb between f0 (t = 0) and
 * f1 (t = t1), linear in t, to dst[(t - ta)*n + p]. */
int tgx_ng_time_interp_dev(const double *f0, const double *f1, double *dst, long n, int t1, int ta, int tb,
                           void *stream);
/* The same operations on up to 32 fields of one geometry in one launch (the
 * coupling's set_bndry, map_in, map_out and init move 4 to 22 fields at a
 * time): job q reads src[q] (interp3d: log-linear in z when zlog[q]) and
 * writes dst[q], every other argument as above.  In the time batch each job
 * has its own size and times, and t1[q] = 0 makes job q a plain copy
 * dst[q] = f1[q] (f0[q], ta[q], tb[q] unused).  No counterparts in the
 * reference (which loops over fields); host arrays of device pointers. */
int tgx_ng_interp3d_batch_dev(int njobs, const double *const *src, double *const *dst, const int *zlog, int nzp,
                              int nxp, int nyp, int ld, const double *zp, const double *xp, const double *yp,
                              int nz, int nx, int ny, long dk, long di, long dj, const double *z, const double *x,
                              const double *y, void *stream);
int tgx_ng_interp2d_batch_dev(int njobs, const double *const *src, double *const *dst, int nxp, int nyp, int ld,
                              const double *xp, const double *yp, int nx, int ny, long di, long dj, const double *x,
                              const double *y, void *stream);
int tgx_ng_time_interp_batch_dev(int njobs, const double *const *f0, const double *const *f1, double *const *dst,
                                 const long *n, const int *t1, const int *ta, const int *tb, void *stream);

/* A host-layout context for one nest domain: device copies of every field of
 * tgx_ng_fields and tgx_ng_out in level planes, a workspace and a stream,
 * filled and read back by field name from arrays in the model's layout.  The
 * Fortran drop-ins with the reference's routine signatures (dt_ng, duv_ng,
 * oplus_ng, comp_ng, comp_{n2d,n4s,no,ar}_ng) are built on it; arrays are
 * copied on every call, so it is for models whose state lives on the host.
 *
 * create: g as for the step, except that its level and latitude arrays
 *   (expz .. glat) are HOST pointers, copied; g->ld and g->capture are
 *   ignored.  Returns NULL on failure (message in tgx_last_error()).
 * put: an input field by its tgx_ng_fields name: 3-D f(nk, nlon+4, nlat+4),
 *   2-D f(nlon+4, nlat+4); lateral data in the model's lon_b(nk, 2, latd) and
 *   lat_b(nk, lond, 2) slices (fields_ng.F90:304-309), lonb_op/latb_op with
 *   nstep_sub such blocks, and "lonb_op_sub"/"latb_op_sub" for the one block
 *   of a single O+ sub-step.
 * put_out: a tgx_ng_out field that a later routine reads (duv reads this
 *   step's tn, Fe, Fn).
 * put_dev: a 3-D or 2-D input from a device array of the model's layout,
 *   after the work already queued on `stream`.
 * get: a tgx_ng_out field (or a 3-D/2-D input) into a host array of the
 *   model's layout.
 * run: "oplus" (all nstep_sub sub-steps), "oplus_sub" (one sub-step with the
 *   nstep_sub time factor, lateral data of lonb_op_sub/latb_op_sub),
 *   "minors", "dt", "duv", "comp" or "step".
 * All calls return when the copies and the routine have finished. */
void *tgx_ng_ctx_create(const tgx_ng_grid *g);
int tgx_ng_ctx_destroy(void *ctx);
int tgx_ng_ctx_put(void *ctx, const char *name, const double *host);
int tgx_ng_ctx_put_out(void *ctx, const char *name, const double *host);
int tgx_ng_ctx_put_dev(void *ctx, const char *name, const double *dev, void *stream);
int tgx_ng_ctx_get(void *ctx, const char *name, double *host);
int tgx_ng_ctx_run(void *ctx, const char *routine);

#ifdef __cplusplus
}
#endif
#endif
