/*
 * libtgx — MI355X (gfx950) implicit vertical-column step of TIE-GCM NG.
 *
 * C ABI only: plain pointers and sizes, no C++ or PyTorch types.  Three groups:
 *
 *  1. Drop-in host entry points.  Same arguments, layout (level fastest,
 *     Fortran column-major) and semantics as the reference subroutines they
 *     replace; host arrays in and out, synchronous, void like the Fortran.
 *     Failures never abort: they leave a message in tgx_last_error().
 *  2. Device entry points: device pointers, stream-ordered on the given
 *     hipStream_t (passed as void*; NULL = the default stream).  They return
 *     0 on success or a TGX_ERR_* code (message in tgx_last_error()).
 *  3. The fused column step on device-resident model state (tgx_step.h).
 *
 * Device entry points that need scratch (Thomas ratios, block gammas) use a
 * library-owned buffer per calling host thread, device and stream, so calls
 * from several host threads, or on several streams, may overlap.  The
 * buffers are held until the process ends.
 */
#ifndef TGX_H
#define TGX_H

#ifdef __cplusplus
extern "C" {
#endif

#define TGX_VERSION 100

enum { TGX_OK = 0, TGX_ERR_ARG = 1, TGX_ERR_HIP = 2, TGX_ERR_NODEV = 3 };
enum { TGX_SCHEME_AUTO = 0, TGX_SCHEME_THOMAS = 1, TGX_SCHEME_PCR = 2 };

int tgx_version(void);
const char *tgx_last_error(void);   /* message of the last failure on this thread */
int tgx_last_status(void);          /* its TGX_ERR_* code, 0 if none */
void tgx_clear_error(void);

/* ------------------------------------------------------------------------ */
/* 1. Drop-in host entry points                                              */

/* Replaces trsolv (trsolv.F:2-3):
 *   subroutine trsolv(a,b,c,f,x,lev0,lev1,k1,k2,lon0,lon1,lonmax,lat,idebug)
 * a,b,c,f,x are real(lev0:lev1,lon0:lon1).  Solves
 *   a(k)x(k-1) + b(k)x(k) + c(k)x(k+1) = f(k),  k = k1..k2
 * for columns lonbeg..lonend, where lonbeg = 3 if lon0 == 1 and lonend =
 * lon1-2 if lon1 == lonmax (periodic points skipped, trsolv.F:32-35).
 * Entries of x outside the solved rows/columns are left unchanged.  lat and
 * idebug are accepted and ignored (debug history dumps, trsolv.F:57-62). */
void tgx_trsolv_f64(const double *a, const double *b, const double *c, const double *f,
                    double *x, int lev0, int lev1, int k1, int k2, int lon0, int lon1,
                    int lonmax, int lat, int idebug);

/* Replaces trsolv_ng (trsolv_ng.F90:1-9):
 *   subroutine trsolv_ng(a,b,c,f,x,k2,i_ng)
 * with the module-state dimensions passed explicitly: arrays are
 * real(nk, nlond, nlatd) (nk = nlevp1_ng(i_ng), nlond/nlatd = the subdomain
 * including its two ghost points per side, fields_ng.F90:137-140).  Rows 1..k2
 * of every column (ghosts included) are solved. */
void tgx_trsolv_ng_f64(const double *a, const double *b, const double *c, const double *f,
                       double *x, int nk, int nlond, int nlatd, int k2);

/* Replaces blktri (duv.F:801-820):
 *   subroutine blktri(a,b,c,f,beta,gamma,y,x,lev0,lev1,lon0,lon1)
 * a,b,c real(2,2,lev0:lev1,lon0:lon1), f,x real(2,lev0:lev1,lon0:lon1); rows
 * lev0..lev1-1 of columns lonbeg..lonend are solved.  lonmax is the
 * params_module value nlonp4 the reference reads from module state.  beta,
 * gamma and y are the reference's work outputs; every caller discards them
 * (duv.F:375) and this entry point leaves them unchanged. */
void tgx_blktri_f64(const double *a, const double *b, const double *c, const double *f,
                    double *beta, double *gamma, double *y, double *x, int lev0, int lev1,
                    int lon0, int lon1, int lonmax);

/* comp's 3x3 block elimination (comp_ng.F90:281-326, comp.F:663-806) on
 * assembled blocks.  The reference has no separate entry point (the solve is
 * inline in comp/comp_ng); this is the boundary after splitting assembly from
 * elimination.  p,q,r real(3,3,nk,ncol), f,x real(3,nk,ncol); rows 1..k2. */
void tgx_blktri3_f64(const double *p, const double *q, const double *r, const double *f,
                     double *x, int nk, int ncol, int k2);

/* Lower-boundary matrix of comp/minor (lbc.F:527-594, bndcmp): b(l,m) is
 * returned row-major in b[3*(l-1)+(m-1)], fb[3]. */
void tgx_bndcmp(double dz, double *b, double *fb);

/* Replaces ringfilter (ringfilter.F:1-111):
 *   subroutine ringfilter(f,lev0,lev1,order,name,lat)
 * f is real(nlonp4, lev0:lev1), longitude fastest (one latitude, the two
 * periodic points on each side included), filtered in place; name 'OPLUS'
 * selects the O+ reconstruction.  The module state the reference reads
 * (params_module nlon, nlat; cons_module dlamda, nlat_filter, chunk_array)
 * is passed explicitly.  Latitudes the chunk table does not filter are left
 * untouched, as in the reference. */
void tgx_ringfilter_f64(double *f, int lev0, int lev1, int order, const char *name, int lat, int nlon,
                        int nlat, double dlamda, int nlat_filter, const int *chunk);

/* ------------------------------------------------------------------------ */
/* 2. Device entry points (level-fastest batches)                            */

/* The ring filter on a device field in the level-plane layout of
 * tgx_step.h (element (k, lat row j, lon i) at k*ld*(nlat+4) + j*ld + i,
 * j = latitude + 1, i = longitude + 1 with the periodic points at i = 0, 1,
 * nlon+2, nlon+3): every latitude the chunk table chunk[nlat_filter]
 * (cons.F:331-355) filters, all nk levels, in one launch.  order <= TGX_RING_MAX_ORDER,
 * nlon <= TGX_RING_MAX_NLON, nlat_filter <= TGX_RING_MAX_ROWS/2. */
#define TGX_RING_MAX_ROWS 128
#define TGX_RING_MAX_NLON 2048
#define TGX_RING_MAX_ORDER 8
int tgx_ringfilter_dev(double *f, int nk, int nlon, int nlat, int ld, int order, int oplus, double dlamda,
                       int nlat_filter, const int *chunk, void *stream);
/* Column j of the batch starts at ptr + (col0+j)*ldk (scalars) and rows
 * k1..k2 (0-based, inclusive) are solved. */

/* scheme: TGX_SCHEME_THOMAS (thread per column, reference operation order),
 * TGX_SCHEME_PCR (one wavefront per column, cyclic reduction in registers,
 * at most 256 rows), TGX_SCHEME_AUTO (timed once per shape, cached). */
int tgx_trsolv_dev_f64(const double *a, const double *b, const double *c, const double *f,
                       double *x, int ldk, int k1, int k2, int col0, int ncol, int scheme,
                       void *stream);
/* The choice the autotuner made for this shape, or TGX_SCHEME_AUTO if none yet. */
int tgx_autotune_choice(int ldk, int k1, int k2, int ncol);

/* Allocation-free tgx_trsolv_dev_f64.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  Same arguments
 * and results as tgx_trsolv_dev_f64; it calls no hipMalloc, hipFree or other
 * allocating routine.  The Thomas scheme keeps its ratios in a fixed device
 * array of TGX_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) that the library holds
 * on every GPU it runs on until the process ends, so ldk*ncol must not exceed
 * TGX_NOALLOC_SCRATCH_DOUBLES (TGX_ERR_ARG otherwise; PCR needs no scratch).
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (other streams, other host threads) race and must be ordered by
 * the caller.  TGX_SCHEME_AUTO takes the autotuner's cached choice for the
 * shape if tgx_trsolv_dev_f64 made one, else Thomas; it never times. */
#define TGX_NOALLOC_SCRATCH_DOUBLES (8u * 1024u * 1024u)
int tgx_trsolv_dev_f64_noalloc(const double *a, const double *b, const double *c, const double *f,
                               double *x, int ldk, int k1, int k2, int col0, int ncol, int scheme,
                               void *stream);

/* Mixed precision: FP32 cyclic reduction of the FP64 system, FP64 residual
 * r = f - A x and nsweep (0..8) refinement sweeps reusing the FP32
 * reduction.  At most 256 rows. */
int tgx_trsolv_dev_mixed(const double *a, const double *b, const double *c, const double *f,
                         double *x, int ldk, int k1, int k2, int col0, int ncol, int nsweep,
                         void *stream);

/* 2x2 block Thomas (duv.F:830-906): a,b,c (2,2,ldk,·), f,x (2,ldk,·);
 * rows 0..nrow-1. */
int tgx_blktri2_dev_f64(const double *a, const double *b, const double *c, const double *f,
                        double *x, int ldk, int nrow, int col0, int ncol, void *stream);

/* 3x3 block elimination of comp: p,q,r (3,3,ldk,·), f,x (3,ldk,·); rows 0..nrow-1. */
int tgx_blktri3_dev_f64(const double *p, const double *q, const double *r, const double *f,
                        double *x, int ldk, int nrow, int col0, int ncol, void *stream);

/* Mixed precision for N x N block rows (N = 1, 2, 3; N = 1 is the scalar
 * system): FP32 elimination with each block pivot inverted in FP64, FP64
 * residual r = F - A x, nsweep (0..8) refinement sweeps reusing the FP32
 * factorisation (the precision variant of trsolv_ng / blktri / comp's 3x3
 * elimination).  Element e (row-major) of block A at (col, k) is at
 * A[e*pstride + col*nk + k]; rows 0..nrow-1; columns with valid[col] == 0 are
 * skipped.  work: tgx_blk_mixed_work_bytes(N, nrow, ncol) bytes of device
 * memory. */
int tgx_blk_mixed_dev(int nblk, const double *A, const double *B, const double *C, const double *F,
                      double *X, long pstride, int nk, int nrow, const unsigned char *valid, long ncol,
                      int nsweep, void *work, void *stream);
unsigned long long tgx_blk_mixed_work_bytes(int nblk, int nrow, long ncol);

#ifdef __cplusplus
}
#endif
#endif
