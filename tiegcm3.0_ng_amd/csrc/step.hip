// Fused assembly + solve of the nested-grid implicit column step.
//
// One thread per column (i, j); each thread walks its column level by level,
// assembles the tridiagonal (or block) row of that level from the model
// fields, and eliminates it at once (Thomas), so the coefficients never
// round-trip through HBM.  Fields are "level planes" with longitude fastest,
// so the 64 lanes of a wavefront read 64 consecutive doubles at every level.
// Horizontal stencils (advection, Shapiro smoothing, divergence, b.grad) read
// neighbouring columns of the same plane, which the L1/L2 caches serve.
//
// Arithmetic follows the reference's evaluation order term by term (the
// library is built with FP contraction off), citing the _ng sources.
#include <cstdlib>

#include "tgx_common.h"
#include "../../include/tgx.h"
#include "../../include/tgx_step.h"

namespace tgx {
namespace ng {

// cons.F:25-46, :98-115, :320-323 (init_cons)
constexpr double RE = 6.37122e8, AVO = 6.023e23, P0 = 5.0e-4, GASK = 8.314e7;
constexpr double BRN2D = 0.6, DTSMOOTH = 0.95, GRAV = 870.0;
constexpr double TSURPLUS = 5.11 * 1.602e-12;
constexpr double DTSMOOTH_DIV2 = 0.5 * (1. - DTSMOOTH);
constexpr double SPVAL = 1.e36;  // params.F:38
constexpr double DZGRAV = GRAV / GASK;
constexpr double RE_INV = 1. / RE;
constexpr double RMO2 = 32., RMO1 = 16., RMHE = 4., RMN2 = 28., RMN4S = 14., RMNO = 30.;
constexpr double RMAR = 40., RMOP = 16., RMN2D = 14.;
constexpr double RIO2 = 1. / RMO2, RIO1 = 1. / RMO1, RIHE = 1. / RMHE, RIN2 = 1. / RMN2;
constexpr double RIN4S = 1. / RMN4S, RINO = 1. / RMNO, RIN2D = 1. / RMN2D;
// chemrates.F:29-50
constexpr double RK4 = 1.0E-10, RK5 = 4.4E-10, RK6 = 4.0E-10, RK7 = 2.0E-10, RK8 = 1.0E-12;
constexpr double RK9 = 6.0E-11, RK10 = 1.3E-10, RK16 = 4.8E-10, RK17 = 1.0E-10, RK18 = 4.0E-10;
constexpr double RK21 = 0.047, RK22 = 0.171, RK23 = 8.E-10, RK24 = 5.0E-12, RK26 = 7.E-10;
constexpr double RK27 = 7.7E-5, BETA2 = 5.0E-12, BETA4 = 7.0E-13, BETA6 = 7.0E-11, BETA7 = 1.06E-5;
// minor_ng.F90:25-28, comp_ng.F90:22-25
constexpr double T00 = 273., TAU = 1.86e+3;
__device__ __forceinline__ double ssn(int n) { return n == 0 ? 1.710 : (n == 1 ? 1.749 : 1.718); }
// phi(r,c), 1-based as in the reference
__device__ __forceinline__ double PHI(int r, int c) {
  const double t[3][4] = {{0., 1.35, 2.16, 1.11}, {0.673, 0., 1.616, 0.769}, {0.270, 0.404, 0., 0.322}};
  return t[r - 1][c - 1];
}

struct Geo {
  int nk, nlon, nlat, nlond, nlatd, ld;
  // Global grid (dt.F ... comp.F): no lateral Dirichlet data; on the
  // latitudes the ring filter reaches (nfilt next to each pole, and the rows
  // across the poles) the column kernels write their raw solutions, which
  // the host then filters and passes through the global post pass; on the
  // other latitudes they apply the post pass themselves.  glb = 0 for a nest
  // domain.
  int glb, nfilt;
  __device__ __forceinline__ bool raw_row(int j) const {
    return glb && (j <= nfilt + 1 || j >= nlat - nfilt + 2);
  }
  size_t P;
  long ngw;  // groups of 64 columns (j * nlond + i) in a tiled workspace region
  long nwt;  // wavefronts of a tile-kernel launch (4 per column tile)
  __device__ __forceinline__ size_t a3(int k, int i, int j) const {
    return (size_t)k * P + (size_t)j * ld + i;
  }
  __device__ __forceinline__ size_t a2(int i, int j) const { return (size_t)j * ld + i; }
  // Element of value n (of nv) at level k of column (i, j) in a tiled
  // workspace region: per level and group of 64 consecutive columns, nv runs
  // of 64 doubles, column fastest, so the values a wavefront of a column
  // kernel reads or writes at one level share one block of nv * 512 bytes.
  __device__ __forceinline__ size_t tile(int k, int i, int j, int n, int nv) const {
    const long c = (long)j * nlond + i;
    return (((size_t)k * ngw + (size_t)(c >> 6)) * nv + n) * 64 + (size_t)(c & 63);
  }
  // The same for the calling thread of a tile kernel (one block of 256 per
  // TGX_TILE_X x TGX_TILE_Y column tile): per level and wavefront of the
  // launch, nv runs of 64 doubles, lane fastest.
  __device__ __forceinline__ size_t wtile(int k, int n, int nv) const {
    const size_t w = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    return (((size_t)k * nwt + w) * nv + n) * 64 + (threadIdx.x & 63);
  }
};

// Rows assembled by the Thomas column kernels, written only when the caller
// asks for them (grid.capture != NULL): plane pl of the capture buffer, level
// fastest per column.
enum { CAP_DT = 0, CAP_OP = 4, CAP_MIN = 8, CAP_UV = 24, CAP_CO = 38 };
__device__ __forceinline__ void cap_put(const tgx_ng_grid &g, const Geo &G, int pl, int i, int j, int k,
                                        double v) {
  size_t ncol = (size_t)G.nlond * G.nlatd, col = (size_t)j * G.nlond + i;
  g.capture[((size_t)pl * ncol + col) * G.nk + k] = v;
}

// Set by the global-grid entry points around their launches on the calling
// host thread (tgx_glb_* in step_api.cpp).
thread_local int t_glb_mode = 0, t_glb_nfilt = 0;
// Columns (wavefronts) per block of the wavefront-per-column kernels, set by
// step_api.cpp before each routine's launches from the autotuner's choice.
thread_local int t_wave_cols = 4;
extern "C" void tgx_ng_set_wave_columns(int cols) { t_wave_cols = cols == 1 || cols == 2 ? cols : 4; }
// Whether the dt tile kernel may stage its tiles by direct-to-LDS loads
// (launches from the calling host thread; on by default, off to compare).
thread_local int t_tile_dma = 1;
extern "C" void tgx_ng_set_tile_dma(int on) { t_tile_dma = on != 0; }
extern "C" void tgx_ng_set_global_mode(int on, int nfilt) {
  t_glb_mode = on;
  t_glb_nfilt = on ? nfilt : 0;
}

static Geo make_geo(const tgx_ng_grid &g) {
  Geo G;
  G.glb = t_glb_mode;
  G.nfilt = t_glb_nfilt;
  G.nk = g.nk; G.nlon = g.nlon; G.nlat = g.nlat;
  G.nlond = g.nlon + 4; G.nlatd = g.nlat + 4; G.ld = g.ld;
  G.P = (size_t)g.ld * (size_t)(g.nlat + 4);
  G.ngw = ((long)G.nlond * G.nlatd + 63) / 64;
  G.nwt = 4L * ((G.nlond + TGX_TILE_X - 1) / TGX_TILE_X) * ((G.nlatd + TGX_TILE_Y - 1) / TGX_TILE_Y);
  return G;
}

// ---- stencils on one level plane (dffm_ng.F90, single task) ----------------
// Branch-free: every case of the reference's edge rules is a pair of loads at
// clamped indices and a scale, so a wavefront issues all its loads at once.
// Scaling by 0.5/0.25 is exactly the reference's division by 2/4.
// A plane read either from global memory (a pointer to the level's plane) or
// from an LDS tile that holds it around the block's points.
__device__ __forceinline__ double pv(const double *f, const Geo &G, int i, int j) {
  return f[(size_t)j * G.ld + i];
}
constexpr int TILE_X = TGX_TILE_X, TILE_Y = TGX_TILE_Y, TILE_H = 2;
constexpr int TILE_W = TILE_X + 2 * TILE_H, TILE_R = TILE_Y + 2 * TILE_H;
struct LdsPlane {
  const double *t;  // TILE_R x TILE_W, row-major
  int i0, j0;       // plane indices of t[0]
};
__device__ __forceinline__ double pv(const LdsPlane &f, const Geo &G, int i, int j) {
  return f.t[(j - f.j0) * TILE_W + (i - f.i0)];
}
#define PV(f, i, j) pv(f, G, i, j)
// df_3d order 1 in longitude: edges one-sided, interior centred / 2
template <class F>
__device__ __forceinline__ double dflon1(const F &f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  int ip = i < L1 ? i + 1 : L1, im = i > 0 ? i - 1 : 0;
  double d = PV(f, ip, j) - PV(f, im, j);
  return (i == 0 || i == L1) ? d : d * 0.5;
}
// df_3d order 2: edges one-sided, next-to-edge centred over 2 cells / 2,
// interior (f(i+2)-f(i-2)) / 4
template <class F>
__device__ __forceinline__ double dflon2(const F &f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
  int ip = edge ? (i == 0 ? 1 : L1) : (near ? i + 1 : i + 2);
  int im = edge ? (i == 0 ? 0 : L1 - 1) : (near ? i - 1 : i - 2);
  double d = PV(f, ip, j) - PV(f, im, j);
  return edge ? d : (near ? d * 0.5 : d * 0.25);
}
// fm_3d order 1/2: edges return f itself ((f+f)*0.5 == f exactly)
template <class F>
__device__ __forceinline__ double fmlon1(const F &f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1);
  int ip = edge ? i : i + 1, im = edge ? i : i - 1;
  return (PV(f, ip, j) + PV(f, im, j)) * 0.5;
}
template <class F>
__device__ __forceinline__ double fmlon2(const F &f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
  int ip = edge ? i : (near ? i + 1 : i + 2), im = edge ? i : (near ? i - 1 : i - 2);
  return (PV(f, ip, j) + PV(f, im, j)) * 0.5;
}
template <class F>
__device__ __forceinline__ double dflat1(const F &f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  int jp = j < J1 ? j + 1 : J1, jm = j > 0 ? j - 1 : 0;
  double d = PV(f, i, jp) - PV(f, i, jm);
  return (j == 0 || j == J1) ? d : d * 0.5;
}
template <class F>
__device__ __forceinline__ double dflat2(const F &f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1), near = (j == 1 || j == J1 - 1);
  int jp = edge ? (j == 0 ? 1 : J1) : (near ? j + 1 : j + 2);
  int jm = edge ? (j == 0 ? 0 : J1 - 1) : (near ? j - 1 : j - 2);
  double d = PV(f, i, jp) - PV(f, i, jm);
  return edge ? d : (near ? d * 0.5 : d * 0.25);
}
template <class F>
__device__ __forceinline__ double fmlat1(const F &f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1);
  int jp = edge ? j : j + 1, jm = edge ? j : j - 1;
  return (PV(f, i, jm) + PV(f, i, jp)) * 0.5;
}
template <class F>
__device__ __forceinline__ double fmlat2(const F &f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1), near = (j == 1 || j == J1 - 1);
  int jp = edge ? j : (near ? j + 1 : j + 2), jm = edge ? j : (near ? j - 1 : j - 2);
  return (PV(f, i, jm) + PV(f, i, jp)) * 0.5;
}

// advec_ng.F90:244-265 at (i, j) of the plane
template <class F, class U>
__device__ __forceinline__ double advec(const F &f, const U &u, const U &v,
                                        const Geo &G, int i, int j, double racs, double dlamda,
                                        double dphi) {
  double ub1 = fmlon1(u, G, i, j), ub2 = fmlon2(u, G, i, j);
  double dx1 = dflon1(f, G, i, j), dx2 = dflon2(f, G, i, j);
  double wk1 = dx1 * 2. * 2. / 3. * ub1 - dx2 * 4. / 12. * ub2;
  if (i <= 1 || i >= G.nlon + 2) wk1 = dx1 * ub1;
  wk1 = wk1 * racs;
  double vb1 = fmlat1(v, G, i, j), vb2 = fmlat2(v, G, i, j);
  double dy1 = dflat1(f, G, i, j), dy2 = dflat2(f, G, i, j);
  double wk2 = dy1 * 2. * 2. / 3. * vb1 - dy2 * 4. / 12. * vb2;
  if (j <= 1 || j >= G.nlat + 2) wk2 = dy1 * vb1;
  return wk1 / dlamda + wk2 * RE_INV / dphi;
}

// advec with the constant factors folded and the divisions by the grid
// spacings taken as products with their reciprocals (rdl = 1/dlamda,
// rdp = 1/(re dphi)): a few ulp from advec, for the species the minor and
// major-species kernels advect.
template <class F, class U>
__device__ __forceinline__ double advec_r(const F &f, const U &u, const U &v, const Geo &G, int i, int j,
                                          double racs, double rdl, double rdp) {
  constexpr double C43 = 4. / 3., C13 = 1. / 3.;
  double ub1 = fmlon1(u, G, i, j), ub2 = fmlon2(u, G, i, j);
  double dx1 = dflon1(f, G, i, j), dx2 = dflon2(f, G, i, j);
  double wk1 = dx1 * C43 * ub1 - dx2 * C13 * ub2;
  if (i <= 1 || i >= G.nlon + 2) wk1 = dx1 * ub1;
  wk1 = wk1 * racs;
  double vb1 = fmlat1(v, G, i, j), vb2 = fmlat2(v, G, i, j);
  double dy1 = dflat1(f, G, i, j), dy2 = dflat2(f, G, i, j);
  double wk2 = dy1 * C43 * vb1 - dy2 * C13 * vb2;
  if (j <= 1 || j >= G.nlat + 2) wk2 = dy1 * vb1;
  return wk1 * rdl + wk2 * rdp;
}

// smooth_ng.F90:283-294: latitude pass at (i, j); rows 0/J1 copy, rows 1/J1-1
// use the 3-point form
template <class F>
__device__ __forceinline__ double smooth_lat(const F &f, const Geo &G, int i, int j, double s) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1), near = (j == 1 || j == J1 - 1);
  int j1p = j < J1 ? j + 1 : J1, j1m = j > 0 ? j - 1 : 0;
  int j2p = j + 2 <= J1 ? j + 2 : J1, j2m = j >= 2 ? j - 2 : 0;
  double fin = PV(f, i, j);
  double fm1 = (PV(f, i, j1m) + PV(f, i, j1p)) * 0.5;
  double fm2 = (PV(f, i, j2m) + PV(f, i, j2p)) * 0.5;
  double gen = fin - s * (2. * fm2 - 8. * fm1 + 6. * fin);
  double nr = fin - s * (-8. * fm1 + 8. * fin);
  return edge ? fin : (near ? nr : gen);
}
// smooth_ng.F90:296-307: longitude pass of the latitude-smoothed field
template <class F>
__device__ __forceinline__ double smooth(const F &f, const Geo &G, int i, int j, double s) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
  int i1p = i < L1 ? i + 1 : L1, i1m = i > 0 ? i - 1 : 0;
  int i2p = i + 2 <= L1 ? i + 2 : L1, i2m = i >= 2 ? i - 2 : 0;
  double t0 = smooth_lat(f, G, i, j, s);
  double fm1 = (smooth_lat(f, G, i1p, j, s) + smooth_lat(f, G, i1m, j, s)) * 0.5;
  double fm2 = (smooth_lat(f, G, i2p, j, s) + smooth_lat(f, G, i2m, j, s)) * 0.5;
  double gen = t0 - s * (2. * fm2 - 8. * fm1 + 6. * t0);
  double nr = t0 - s * (-8. * fm1 + 8. * t0);
  return edge ? t0 : (near ? nr : gen);
}

// div_ng.F90:331-346
template <class F>
__device__ __forceinline__ double divx(const F &f, const Geo &G, int i, int j) {
  double d1 = dflon1(f, G, i, j);
  double d2 = dflon2(f, G, i, j);
  double gen = d1 * 2. * 2. / 3. - d2 * 4. / 12.;
  return (i <= 1 || i >= G.nlon + 2) ? d1 : gen;
}
template <class F>
__device__ __forceinline__ double divy(const F &f, const Geo &G, int i, int j) {
  double d1 = dflat1(f, G, i, j);
  double d2 = dflat2(f, G, i, j);
  double gen = d1 * 2. * 2. / 3. - d2 * 4. / 12.;
  return (j <= 1 || j >= G.nlat + 2) ? d1 : gen;
}

// oplus_ng.F90:344-365 (bdotdh)
template <class F>
__device__ __forceinline__ double bdotdh(const F &f, const Geo &G, int i, int j, double bx,
                                         double by, double cs, double dlamda, double dphi) {
  double a1 = bx * dflon1(f, G, i, j);
  double a2 = by * dflat1(f, G, i, j);
  a1 = a1 / cs;
  return 1. / RE * (a1 / dlamda + a2 / dphi);
}
#undef PV

// Column mapping: flat column id -> (i, j), XCD-contiguous block order.
__device__ __forceinline__ bool column_of(const Geo &G, int &i, int &j) {
  int nb = gridDim.x;
  int b = tgx_xcd_block(nb);
  if (b >= nb) return false;
  long c = (long)b * blockDim.x + threadIdx.x;
  if (c >= (long)G.nlond * G.nlatd) return false;
  j = (int)(c / G.nlond);
  i = (int)(c - (long)j * G.nlond);
  return true;
}

__device__ __forceinline__ bool lateral_value(const Geo &G, const double *lonb, const double *latb,
                                              int k, int i, int j, double &v) {
  // dt_ng.F90:201-206 and the same hook in every _ng routine: lon ghosts
  // first, then lat ghosts (corners take the lat data).  A global grid has
  // none (its ghost columns are periodic points, its ghost rows lie across
  // the poles).
  if (G.glb) return false;
  if (j == 0 || j == G.nlatd - 1) {
    v = latb[(size_t)(k * 2 + (j == 0 ? 0 : 1)) * G.ld + i];
    return true;
  }
  if (i == 0 || i == G.nlond - 1) {
    v = lonb[(size_t)(k * 2 + (i == 0 ? 0 : 1)) * G.nlatd + j];
    return true;
  }
  return false;
}

// =============================================================================
// Horizontal-stencil prelude: one thread per (column, level), every term a
// column kernel needs from neighbouring columns, written as a level plane.
enum { ST_ADVEC = 0, ST_SMOOTH, ST_DIVX, ST_DIVY, ST_BDOTDH, ST_VDOTN, ST_ADVEC_R };
struct StencilJob {
  int kind;
  const double *f;
  double *out;
  double s;   // Shapiro coefficient for ST_SMOOTH
};
struct StencilJobs {
  int n;
  StencilJob job[8];
  const double *aux[3];  // STC_COMP: the fs(:,:,0) planes the right-hand sides take
  double *ci;            // STC_COMP: the tiled comp inputs (values 0-2 are written);
                         // STC_MINOR: the tiled minor right-hand-side parts (MR_NV values)
};
// How a prelude stores its jobs: each job its own plane (STC_NONE), or the
// smoothed and advected planes of one species combined into the part of the
// right-hand side they form, in the reference's order of operations:
//   STC_MINOR (jobs 0-3 smooth fcomp_tm1, 4-7 advect fcomp, minor_ng.F90:219):
//     out[s] = smooth[s]*dtx2inv - advec[s]
//   STC_COMP (jobs 0-2 advect O2/O/He, 3-5 smooth *_nm, comp_ng.F90:191-193):
//     out[s] = smooth[s]*dtx2inv - (advec[s] - fs(s,0))   (He: no fs term)
enum { STC_NONE = 0, STC_MINOR, STC_COMP };
// Tiled comp inputs: the right-hand-side parts of O2, O, He (0-2) and
// fs(:,1,1), fs(:,1,2), fs(:,2,1), fs(:,2,2) of comp_o2o_ng (3-6).
constexpr int CI_NV = 7;
// Tiled minor-species right-hand-side parts, one value per species id
// (NO, N4S, N2D, AR), and the Thomas ratios (0, 1) and forward solutions
// (2, 3) of the two species of a pass.
constexpr int MR_NV = 4, MW_NV = 4;
// Tiled minor-species sources: loss and scaled production of NO or N2D
// (values 0, 1; N2D's until its solve, then NO's) and of N4S (2, 3), and the
// six coefficients of the NO / N4S sources linear in N2D.
constexpr int MS_NV = 4, ML_NV = 6;
__host__ __device__ constexpr int ms_loss(int species) { return species == 1 ? 2 : 0; }
__host__ __device__ constexpr int ms_prod(int species) { return species == 1 ? 3 : 1; }
// Tiled 2x2 block Thomas values of the wind solve: gamma (g11, g21, g12, g22)
// and the forward solution (u, v).
constexpr int UW_NV = 6;
// Tiled Thomas values of the TN solve: the ratio and the forward solution.
constexpr int DW_NV = 2;
// Tiled Thomas values of the O+ solve: the ratio and the forward solution.
constexpr int OW_NV = 2;
// O+ factorisation kept for the later sub-steps of a step: per level the
// sub-diagonal after the boundary rows, the pivot, and at levels 0 and nk-2
// the boundary terms of the right-hand side.
constexpr int OF_NV = 3;

// Horizontal stencils: a block takes a TILE_X x TILE_Y tile of the level
// planes and a run of STENCIL_RUN levels.  Each level's inputs are loaded
// into LDS with a 2-point halo (coalesced rows); halo points outside the
// domain are never read (the stencils clamp at the domain edges).  The job
// kinds are fixed at compile time.  While a level is evaluated from LDS, the
// next level's tiles are already in flight into registers.  The latitude
// passes of all smoothed fields share one barrier.  Dynamic LDS: one tile per
// job (+ u, v when a job advects) and one latitude-pass buffer per smoothed
// field.
constexpr int STENCIL_RUN = 3;
template <int... K>
struct StencilKinds {
  static constexpr int n = sizeof...(K);
  static constexpr int kind[n] = {K...};
  static constexpr bool adv() {
    for (int q = 0; q < n; ++q)
      if (kind[q] == ST_ADVEC || kind[q] == ST_ADVEC_R) return true;
    return false;
  }
  static constexpr int nsmooth() {
    int c = 0;
    for (int q = 0; q < n; ++q) c += kind[q] == ST_SMOOTH;
    return c;
  }
  static constexpr int ntiles() { return n + (adv() ? 2 : 0); }
  static constexpr size_t lds_bytes() {
    return ((size_t)ntiles() * TILE_R * TILE_W + (size_t)nsmooth() * TILE_Y * TILE_W) * sizeof(double);
  }
};

template <class KS, int COMB>
__global__ void __launch_bounds__(256) stencil_run_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                          StencilJobs J) {
  extern __shared__ double smem[];
  constexpr int NJ = KS::n, NT = KS::ntiles(), NS = KS::nsmooth();
  constexpr bool ADV = KS::adv();
  constexpr int TS = TILE_R * TILE_W, TL = TILE_Y * TILE_W;
  const int ntx = (G.nlond + TILE_X - 1) / TILE_X, nty = (G.nlatd + TILE_Y - 1) / TILE_Y;
  const int nrun = (G.nk + STENCIL_RUN - 1) / STENCIL_RUN;
  const int b = tgx_xcd_block(gridDim.x);
  if (b >= ntx * nty * nrun) return;  // uniform per block
  const int bxi = b % ntx, rr = b / ntx, byj = rr % nty, run = rr / nty;
  const int k0 = run * STENCIL_RUN, k1 = min(k0 + STENCIL_RUN, G.nk);
  const int i0 = bxi * TILE_X, j0 = byj * TILE_Y;
  const int i = i0 + (int)threadIdx.x % TILE_X, j = j0 + (int)threadIdx.x / TILE_X;
  const bool inside = i < G.nlond && j < G.nlatd;
  const size_t c2 = (size_t)j * G.ld + i;
  const double *src[NT];
#pragma unroll
  for (int t = 0; t < NJ; ++t) src[t] = J.job[t].f;
  if (ADV) {
    src[NT - 2] = in.un;
    src[NT - 1] = in.vn;
  }
  // tile elements of this thread (a 12 x 36 tile: two per thread at most)
  bool okr[2];
  int er[2];
  size_t qr[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int e = (int)threadIdx.x + r * 256;
    const int ii = i0 - TILE_H + e % TILE_W, jj = j0 - TILE_H + e / TILE_W;
    er[r] = e;
    okr[r] = e < TS && ii >= 0 && ii < G.nlond && jj >= 0 && jj < G.nlatd;
    qr[r] = okr[r] ? (size_t)jj * G.ld + ii : 0;
  }
  double ld[2][NT];
  auto issue = [&](int k) {
    const size_t off = (size_t)k * G.P;
#pragma unroll
    for (int r = 0; r < 2; ++r)
      if (okr[r]) {
#pragma unroll
        for (int t = 0; t < NT; ++t) ld[r][t] = src[t][off + qr[r]];
      }
  };
  const double racs = inside ? g.racs[j] : 0., csj = inside ? g.cs[j] : 1.;
  const double bx = inside ? in.bx[c2] : 0., by = inside ? in.by[c2] : 0., bm = inside ? in.bmod2[c2] : 0.;
  const double rdl = 1. / g.dlamda, rdp = RE_INV / g.dphi;
  double *tl = smem + (size_t)NT * TS;  // NS latitude-pass buffers
  const LdsPlane U{smem + (NT - 2) * TS, i0 - TILE_H, j0 - TILE_H}, V{smem + (NT - 1) * TS, i0 - TILE_H, j0 - TILE_H};
  issue(k0);
  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
      if (okr[r]) {
#pragma unroll
        for (int t = 0; t < NT; ++t) smem[t * TS + er[r]] = ld[r][t];
      }
    __syncthreads();
    if (k + 1 < k1) issue(k + 1);
    asm volatile("" ::: "memory");
    if (NS > 0) {  // smooth_ng.F90:283-294 for the tile rows and their longitude halo
      int qs = 0;
#pragma unroll
      for (int q = 0; q < NJ; ++q) {
        if (KS::kind[q] != ST_SMOOTH) continue;
        const LdsPlane F{smem + q * TS, i0 - TILE_H, j0 - TILE_H};
        for (int e = threadIdx.x; e < TL; e += blockDim.x) {
          int ii = i0 - TILE_H + e % TILE_W, jj = j0 + e / TILE_W;
          if (ii >= 0 && ii < G.nlond && jj < G.nlatd) tl[qs * TL + e] = smooth_lat(F, G, ii, jj, J.job[q].s);
        }
        ++qs;
      }
      __syncthreads();
    }
    if (inside) {
      const size_t p = (size_t)k * G.P + c2;
      int qs = 0;
      double vv[NJ];
#pragma unroll
      for (int q = 0; q < NJ; ++q) {
        const LdsPlane F{smem + q * TS, i0 - TILE_H, j0 - TILE_H};
        double v;
        if (KS::kind[q] == ST_SMOOTH) {  // :296-307 from the latitude pass
          const double *row = tl + qs * TL + (j - j0) * TILE_W + TILE_H - i0;
          ++qs;
          const double sc = J.job[q].s;
          int L1 = G.nlond - 1;
          bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
          int i1p = i < L1 ? i + 1 : L1, i1m = i > 0 ? i - 1 : 0;
          int i2p = i + 2 <= L1 ? i + 2 : L1, i2m = i >= 2 ? i - 2 : 0;
          double t0 = row[i];
          double fm1 = (row[i1p] + row[i1m]) * 0.5;
          double fm2 = (row[i2p] + row[i2m]) * 0.5;
          double gen = t0 - sc * (2. * fm2 - 8. * fm1 + 6. * t0);
          double nr = t0 - sc * (-8. * fm1 + 8. * t0);
          v = edge ? t0 : (near ? nr : gen);
        } else if (KS::kind[q] == ST_ADVEC) {
          v = advec(F, U, V, G, i, j, racs, g.dlamda, g.dphi);
        } else if (KS::kind[q] == ST_ADVEC_R) {
          v = advec_r(F, U, V, G, i, j, racs, rdl, rdp);
        } else if (KS::kind[q] == ST_DIVX) {
          v = divx(F, G, i, j);
        } else if (KS::kind[q] == ST_DIVY) {
          v = divy(F, G, i, j);
        } else if (KS::kind[q] == ST_BDOTDH) {
          v = bdotdh(F, G, i, j, bx, by, csj, g.dlamda, g.dphi);
        } else {  // ST_VDOTN: oplus_ng.F90:196-209, f = op/bmod2**2
          double uii = k < G.nk - 1 ? .5 * (in.ui[p] + in.ui[p + G.P]) : 1.5 * in.ui[p] - .5 * in.ui[p - G.P];
          double vii = k < G.nk - 1 ? .5 * (in.vi[p] + in.vi[p + G.P]) : 1.5 * in.vi[p] - .5 * in.vi[p - G.P];
          double vx = uii * dflon1(F, G, i, j), vy = vii * dflat1(F, G, i, j);
          vx = vx / csj;
          double vh = 1. / RE * (vx / g.dlamda + vy / g.dphi);
          v = vh * (bm * bm);
        }
        if (COMB == STC_NONE)
          J.job[q].out[p] = v;
        else
          vv[q] = v;
      }
      if (COMB == STC_MINOR) {
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) J.ci[G.tile(k, i, j, sp, MR_NV)] = vv[sp] * g.dtx2inv - vv[4 + sp];
      } else if (COMB == STC_COMP) {
#pragma unroll
        for (int sp = 0; sp < 3; ++sp)
          J.ci[G.tile(k, i, j, sp, CI_NV)] = vv[3 + sp] * g.dtx2inv - (vv[sp] - (sp < 2 ? J.aux[sp][p] : 0.));
      }
    }
    __syncthreads();  // the next level overwrites the tiles and buffers
  }
}
using KindsOplus = StencilKinds<ST_BDOTDH, ST_BDOTDH, ST_BDOTDH, ST_VDOTN, ST_BDOTDH, ST_SMOOTH>;
using KindsDt = StencilKinds<ST_ADVEC, ST_SMOOTH>;
using KindsMinor =
    StencilKinds<ST_SMOOTH, ST_SMOOTH, ST_SMOOTH, ST_SMOOTH, ST_ADVEC_R, ST_ADVEC_R, ST_ADVEC_R, ST_ADVEC_R>;
using KindsComp = StencilKinds<ST_ADVEC_R, ST_ADVEC_R, ST_ADVEC_R, ST_SMOOTH, ST_SMOOTH, ST_SMOOTH>;
using KindsDuv = StencilKinds<ST_DIVX, ST_DIVY, ST_SMOOTH, ST_SMOOTH, ST_ADVEC, ST_ADVEC>;

// =============================================================================
// dt_ng.F90:54-213 with the horizontal stencils fused in: a block covers a TILE_X x
// TILE_Y tile of columns; at each level the tile and its 2-point halo of TN,
// UN, VN and TN_NM are staged through LDS (double-buffered, the next level's
// loads issued with the column loads), the advection of TN (advec_ng) and the
// Shapiro smoothing of TN_NM (smooth_ng, latitude pass once per point) are
// evaluated from LDS, and the column is assembled and eliminated by Thomas
// (trsolv_ng).  Every thread of the block takes part in the staging and the
// barriers; lateral-boundary and out-of-domain columns do not solve.
//
// DMA: the tiles go from HBM straight to LDS (global_load_lds_dwordx4: lane l
// of wavefront w writes tile points 128 w + 2 l and 2 l + 1, one row of the
// tile, at LDS base + 16 l), so no registers hold them in flight; the kernel
// then fits three wavefronts per SIMD and every tile of a 0.625 deg grid is
// resident in one round.  Needs 16-byte aligned TN, UN, VN, TN_NM and an even
// row pitch (the host checks).
template <bool DMA>
__device__ __forceinline__ void tile_dma16(const double *src, double *lds_wave_base) {
  __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)src,
                                   (__attribute__((address_space(3))) void *)lds_wave_base, 16, 0, 0);
}
// Waits for this wavefront's outstanding vector memory operations, the tile
// loads into LDS among them (before the barrier that publishes them).
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <bool CAP, bool DMA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DMA ? 3 : 1)))
dt_tile_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in, tgx_ng_out out, double *__restrict__ w2s) {
  constexpr int TS = TILE_R * TILE_W, NF = 4;
  constexpr int TSB = DMA ? 4 * 128 : TS;  // DMA: each of the 4 wavefronts fills 128 tile points
  static_assert(TS <= 4 * 128 && TILE_W % 2 == 0, "tile rows of whole point pairs, 128 points per wavefront");
  __shared__ double tb[2][NF][TSB];
  __shared__ double tls[TILE_Y * TILE_W];
  const int ntx = (G.nlond + TILE_X - 1) / TILE_X;
  const int bb = tgx_xcd_block(gridDim.x);
  const int i0 = (bb % ntx) * TILE_X, j0 = (bb / ntx) * TILE_Y;
  const int i = i0 + (int)threadIdx.x % TILE_X, j = j0 + (int)threadIdx.x / TILE_X;
  const bool active = i < G.nlond && j < G.nlatd;
  const bool lat = active && !G.glb && (j == 0 || j == G.nlatd - 1 || i == 0 || i == G.nlond - 1);
  const bool solve = active && !lat;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  // tile slots of this thread: e = tid and tid + 256
  int te[2];
  bool tok[2];
  size_t toff[2];
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const int e = (int)threadIdx.x + sl * 256;
    const int ii = i0 - TILE_H + e % TILE_W, jj = j0 - TILE_H + e / TILE_W;
    te[sl] = e;
    tok[sl] = e < TS && ii >= 0 && ii < G.nlond && jj >= 0 && jj < G.nlatd;
    toff[sl] = tok[sl] ? (size_t)jj * G.ld + ii : 0;
  }
  const double *tsrc[NF] = {in.tn, in.un, in.vn, in.tn_nm};
  double stage[NF][2];
  auto tile_load = [&](int k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int sl = 0; sl < 2; ++sl)
        if (tok[sl]) stage[f][sl] = tsrc[f][(size_t)k * G.P + toff[sl]];
  };
  auto tile_store = [&](int buf) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int sl = 0; sl < 2; ++sl)
        if (tok[sl]) tb[buf][f][te[sl]] = stage[f][sl];
  };
  // DMA staging: this lane's pair of tile points (a pair is inside the grid
  // or outside it as a whole on the left, top and bottom; on the right a pair
  // may end one point past the grid, in the row pitch, which is even)
  const int e2 = 2 * (int)threadIdx.x, wv = (int)threadIdx.x >> 6;
  const int di = i0 - TILE_H + e2 % TILE_W, dj = j0 - TILE_H + e2 / TILE_W;
  const bool dok = DMA && e2 < TS && di >= 0 && di < G.nlond && dj >= 0 && dj < G.nlatd;
  const unsigned doff = dok ? (unsigned)((size_t)dj * G.ld + di) : 0u;  // within one plane
  auto tile_dma = [&](int k, int buf) {
    if (dok) {
#pragma unroll
      for (int f = 0; f < NF; ++f) tile_dma16<DMA>(tsrc[f] + (size_t)k * G.P + doff, &tb[buf][f][128 * wv]);
    }
  };
  const int ic = active ? i : 0, jc = active ? j : 0;
  const size_t c2 = G.a2(ic, jc);
  double t_lbc = 0., tl = 0.;
  if (active) {
    t_lbc = in.t_lbc[c2];
    tl = in.tlbc[c2];
    out.tlbc_nm[c2] = tl;
    out.tlbc[c2] = t_lbc;
  }
  if (lat) {
    double v;
    for (int k = 0; k < nk; ++k) {
      lateral_value(G, in.lonb_tn, in.latb_tn, k, i, j, v);
      size_t p = G.a3(k, i, j);
      out.tn_nm[p] = DTSMOOTH_DIV2 * (in.tn_nm[p] + v) + DTSMOOTH * in.tn[p];
      out.tn[p] = v > 100. ? v : 100.;
    }
  }
#define F3(f, k) (f)[G.a3((k), ic, jc)]
  struct DL { double w, cp, km, qtotal, schti, kt, un, vn; };
  struct DS { double xn, o1, rkm12, mbar, scht, hdt, qji, cool_imp, cool_exp; };
  auto load_dl = [&](int L, DL &D) {
    size_t p = G.a3(L, ic, jc);
    D.w = in.w_upd[p]; D.cp = in.cp[p]; D.km = in.km[p]; D.qtotal = in.qtotal[p];
    D.schti = in.schti[p]; D.kt = in.kt[p]; D.un = in.un[p]; D.vn = in.vn[p];
  };
  auto load_ds = [&](int L, DS &S) {
    size_t p = G.a3(L, ic, jc);
    S.xn = in.xnmbar[p]; S.o1 = in.o1[p];
    S.rkm12 = in.rkm12[p]; S.mbar = in.mbar[p]; S.scht = in.scht[p]; S.hdt = in.hdt[p];
    S.qji = in.qji_tn[p]; S.cool_imp = in.cool_imp[p]; S.cool_exp = in.cool_exp[p];
  };
  auto gf = [&](const DL &D, double expzL, double diftL, double tni, double &gg, double &ff) {
    double h = D.schti;
    double rho = P0 * g.expzmid_inv / (h * GRAV);
    rho = rho * expzL;
    double gpart = h * h * rho * D.cp;
    double gpart_a = gpart * diftL;
    gg = GRAV * (D.kt + gpart_a) / (P0 * h * (dz * dz));
    ff = GRAV * GRAV * (h * h) * rho / (tni * P0 * 2. * dz);
    ff = ff * diftL;
  };
  const double ulbc = solve ? in.ulbc[c2] : 0., vlbc = solve ? in.vlbc[c2] : 0.;
  const double racs = g.racs[jc];
  DL dc = {}, dn = {};
  double tn_m = 0., tn_c = 0., tn_p = 0.;
  double gk = 0., fk = 0.;
  if (solve) {
    load_dl(0, dc);
    tn_c = F3(in.tn, 0);
    tn_p = F3(in.tn, 1);
    gf(dc, g.expz[0], g.dift[0], tl, gk, fk);
  }
  double un_m = 0., vn_m = 0.;
  double w1 = 0., w3 = 0., rprev = 0., dudz_prev = 0., dvdz_prev = 0.;
  if (DMA) {
    tile_dma(0, 0);
    wait_vmem();
  } else {
    tile_load(0);
    tile_store(0);
  }
  for (int k = 0; k <= nk - 2; ++k) {
    DS ds = {};
    double tn_pp = 0.;
    if (!DMA && k + 1 <= nk - 2) tile_load(k + 1);
    if (solve) {
      load_dl(k + 1, dn);
      load_ds(k, ds);
      tn_pp = k + 2 <= nk - 1 ? F3(in.tn, k + 2) : 0.;
    }
    const double expzk = g.expz[k], expzk1 = g.expz[k + 1], dift1 = g.dift[k + 1];
    asm volatile("" ::: "memory");
    __syncthreads();  // the level-k tiles are in buffer k&1
    // buffer (k+1)&1 was last read in iteration k-1, which every wave has left
    if (DMA && k + 1 <= nk - 2) tile_dma(k + 1, (k + 1) & 1);
    const int bk = k & 1;
    const LdsPlane Ftn{tb[bk][0], i0 - TILE_H, j0 - TILE_H}, U{tb[bk][1], i0 - TILE_H, j0 - TILE_H},
        V{tb[bk][2], i0 - TILE_H, j0 - TILE_H}, Fnm{tb[bk][3], i0 - TILE_H, j0 - TILE_H};
    // smooth_ng.F90:283-294: latitude pass of TN_NM on the tile rows and their longitude halo
    for (int e = threadIdx.x; e < TILE_W * TILE_Y; e += blockDim.x) {
      int ii = i0 - TILE_H + e % TILE_W, jj = j0 + e / TILE_W;
      if (ii >= 0 && ii < G.nlond && jj < G.nlatd) tls[e] = smooth_lat(Fnm, G, ii, jj, g.shapiro);
    }
    __syncthreads();
    if (solve) {
      // advec_ng of TN and the longitude pass of the smoothing (:296-307)
      const double hadv = advec(Ftn, U, V, G, i, j, racs, g.dlamda, g.dphi);
      double tsm;
      {
        const double *row = tls + (j - j0) * TILE_W + TILE_H - i0;
        const double s_ = g.shapiro;
        int L1 = G.nlond - 1;
        bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
        int i1p = i < L1 ? i + 1 : L1, i1m = i > 0 ? i - 1 : 0;
        int i2p = i + 2 <= L1 ? i + 2 : L1, i2m = i >= 2 ? i - 2 : 0;
        double t0 = row[i];
        double fm1 = (row[i1p] + row[i1m]) * 0.5;
        double fm2 = (row[i2p] + row[i2m]) * 0.5;
        double gen = t0 - s_ * (2. * fm2 - 8. * fm1 + 6. * t0);
        double nr = t0 - s_ * (-8. * fm1 + 8. * t0);
        tsm = edge ? t0 : (near ? nr : gen);
      }
      const double w_k = dc.w, w_k1 = dn.w;
      // veradv (:57-65)
      double raw_k = k == 0 ? (tn_c - t_lbc) * w_k * 2. : (tn_c - tn_m) * w_k;
      double raw_k1 = (k + 1 <= nk - 2) ? (tn_p - tn_c) * w_k1 : 0.;
      double veradv = (raw_k + raw_k1) * .5;
      veradv = veradv / dz;
      // interface means (:71-78)
      double cpi = .5 * (dc.cp + dn.cp);
      double kmi = .5 * (dc.km + dn.km);
      double wi = .5 * (w_k + w_k1);
      // cptn (:67, :80-83)
      double advec_tn = hadv + veradv;
      double cptn = cpi * (advec_tn - dtx2inv * tsm);
      cptn = cptn * expzk;
      // dudz (:85-94)
      double dudz, dvdz;
      if (k == 0) {
        dudz = (dc.un + 1. / 3. * dn.un - 4. / 3. * ulbc) / dz;
        dvdz = (dc.vn + 1. / 3. * dn.vn - 4. / 3. * vlbc) / dz;
      } else if (k <= nk - 3) {
        dudz = (dn.un - un_m) / (2. * dz);
        dvdz = (dn.vn - vn_m) / (2. * dz);
      } else {
        dudz = dudz_prev / 3.;
        dvdz = dvdz_prev / 3.;
      }
      dudz_prev = dudz;
      dvdz_prev = dvdz;
      // heating (:96-113)
      double qsolar = .5 * (dc.qtotal + dn.qtotal);
      double t = ds.xn * ds.o1 * RIO1;
      double qrecomb = TSURPLUS * ds.rkm12 * (t * t) * AVO / ds.mbar;
      double qm = GRAV * kmi / (P0 * ds.scht) * (dudz * dudz + dvdz * dvdz);
      qm = qm / expzk;
      double total_heat = qsolar + ds.hdt + qrecomb + ds.qji + qm;
      double total_heat_a = total_heat * expzk;
      cptn = cptn - total_heat_a;
      // coefficients (:141-163)
      double p = gk - fk, q, r;
      double g1 = 0., f1 = 0.;
      if (k <= nk - 3) {
        double tni1 = .5 * (tn_c + tn_p);
        gf(dn, expzk1, dift1, tni1, g1, f1);
        q = -(g1 - f1 + gk + fk);
        r = g1 + f1;
      } else {
        q = -gk - fk;
        r = 0.;
      }
      double rhs = cptn;
      double qpart = cpi * (dtx2inv + ds.cool_imp) + wi * GASK / ds.mbar;
      rhs = rhs + ds.cool_exp;
      double qpart_a = qpart * expzk;
      q = q - qpart_a;
      if (k == 0) {
        q = q - p;
        rhs = rhs - 2. * p * t_lbc;
        p = 0.;
      }
      if (CAP) {
        cap_put(g, G, CAP_DT + 0, i, j, k, p); cap_put(g, G, CAP_DT + 1, i, j, k, q);
        cap_put(g, G, CAP_DT + 2, i, j, k, r); cap_put(g, G, CAP_DT + 3, i, j, k, rhs);
      }
      // Thomas forward (trsolv_ng.F90:14-22)
      if (k == 0) {
        w1 = q;
        w3 = rhs / w1;
      } else {
        double w2 = rprev / w1;
        w2s[G.wtile(k - 1, 0, DW_NV)] = w2;
        w1 = q - p * w2;
        w3 = (rhs - p * w3) / w1;
      }
      w2s[G.wtile(k, 1, DW_NV)] = w3;  // forward solution
      rprev = r;
      gk = g1; fk = f1;
    }
    tn_m = tn_c; tn_c = tn_p; tn_p = tn_pp;
    un_m = dc.un; vn_m = dc.vn;
    dc = dn;
    if (k + 1 <= nk - 2) {
      if (DMA)
        wait_vmem();  // this wave's level-(k+1) tile loads have landed in LDS
      else
        tile_store((k + 1) & 1);
    }
  }
  if (!solve) return;
  // back substitution (:165-167) fused with the output stage (:208-213); the
  // next level's loads are in flight while this level is stored
  struct DB { double x, w2, tnm, tn; };
  auto load_db = [&](int k, DB &B) {
    size_t p = G.a3(k, ic, jc);
    const size_t t = G.wtile(k, 0, DW_NV);
    B.x = w2s[t + 64]; B.w2 = k < nk - 2 ? w2s[t] : 0.; B.tnm = in.tn_nm[p]; B.tn = in.tn[p];
  };
  const double tnm_top = F3(in.tn_nm, nk - 1), tn_top = F3(in.tn, nk - 1);
  double xn = w3;
  DB bc, bn;
  load_db(nk - 2, bc);
  for (int k = nk - 2; k >= 0; --k) {
    if (k > 0) load_db(k - 1, bn);
    asm volatile("" ::: "memory");
    double x = k == nk - 2 ? xn : bc.x - bc.w2 * xn;
    if (G.raw_row(jc)) {  // the raw solution; filter_tn and the post pass follow (dt.F:389-424)
      out.tn[G.a3(k, ic, jc)] = x;
      xn = x;
      bc = bn;
      continue;
    }
    if (G.glb && k == nk - 2) {  // spval at the top level (dt.F:407-408)
      size_t pt = G.a3(nk - 1, ic, jc);
      out.tn[pt] = SPVAL;
      out.tn_nm[pt] = SPVAL;
    }
    if (!G.glb && k == nk - 3) {
      double xt = xn * xn / x;
      size_t pt = G.a3(nk - 1, ic, jc);
      out.tn_nm[pt] = DTSMOOTH_DIV2 * (tnm_top + xt) + DTSMOOTH * tn_top;
      out.tn[pt] = xt > 100. ? xt : 100.;
    }
    size_t p = G.a3(k, ic, jc);
    out.tn_nm[p] = DTSMOOTH_DIV2 * (bc.tnm + x) + DTSMOOTH * bc.tn;
    out.tn[p] = x > 100. ? x : 100.;
    xn = x;
    bc = bn;
  }
#undef F3
}

// 3x3 inverse (matutil.F90:110-118), column-major, same term order as the oracle
__device__ __forceinline__ void matinv3_ng(const double *A, double *Bm) {
#define M3(r, c) A[((r) - 1) + 3 * ((c) - 1)]
  double d = M3(1,1)*M3(2,2)*M3(3,3) - M3(1,1)*M3(2,3)*M3(3,2)
           - M3(1,2)*M3(2,1)*M3(3,3) + M3(1,2)*M3(2,3)*M3(3,1)
           + M3(1,3)*M3(2,1)*M3(3,2) - M3(1,3)*M3(2,2)*M3(3,1);
  Bm[0] =  (M3(2,2)*M3(3,3) - M3(2,3)*M3(3,2)) / d;
  Bm[1] = -(M3(2,1)*M3(3,3) - M3(2,3)*M3(3,1)) / d;
  Bm[2] =  (M3(2,1)*M3(3,2) - M3(2,2)*M3(3,1)) / d;
  Bm[3] = -(M3(1,2)*M3(3,3) - M3(1,3)*M3(3,2)) / d;
  Bm[4] =  (M3(1,1)*M3(3,3) - M3(1,3)*M3(3,1)) / d;
  Bm[5] = -(M3(1,1)*M3(3,2) - M3(1,2)*M3(3,1)) / d;
  Bm[6] =  (M3(1,2)*M3(2,3) - M3(1,3)*M3(2,2)) / d;
  Bm[7] = -(M3(1,1)*M3(2,3) - M3(1,3)*M3(2,1)) / d;
  Bm[8] =  (M3(1,1)*M3(2,2) - M3(1,2)*M3(2,1)) / d;
#undef M3
}

// The same inverse with one division: the adjugate times 1/det.  Within
// 2 ulp of matinv3_ng per element; used where the assembly is evaluated with
// shared reciprocals (column kernels of the minor species and comp).
__device__ __forceinline__ void matinv3_r(const double *A, double *Bm) {
#define M3(r, c) A[((r) - 1) + 3 * ((c) - 1)]
  double d = M3(1,1)*M3(2,2)*M3(3,3) - M3(1,1)*M3(2,3)*M3(3,2)
           - M3(1,2)*M3(2,1)*M3(3,3) + M3(1,2)*M3(2,3)*M3(3,1)
           + M3(1,3)*M3(2,1)*M3(3,2) - M3(1,3)*M3(2,2)*M3(3,1);
  const double r = 1. / d;
  Bm[0] =  (M3(2,2)*M3(3,3) - M3(2,3)*M3(3,2)) * r;
  Bm[1] = -(M3(2,1)*M3(3,3) - M3(2,3)*M3(3,1)) * r;
  Bm[2] =  (M3(2,1)*M3(3,2) - M3(2,2)*M3(3,1)) * r;
  Bm[3] = -(M3(1,2)*M3(3,3) - M3(1,3)*M3(3,2)) * r;
  Bm[4] =  (M3(1,1)*M3(3,3) - M3(1,3)*M3(3,1)) * r;
  Bm[5] = -(M3(1,1)*M3(3,2) - M3(1,2)*M3(3,1)) * r;
  Bm[6] =  (M3(1,2)*M3(2,3) - M3(1,3)*M3(2,2)) * r;
  Bm[7] = -(M3(1,1)*M3(2,3) - M3(1,3)*M3(2,1)) * r;
  Bm[8] =  (M3(1,1)*M3(2,2) - M3(1,2)*M3(2,1)) * r;
#undef M3
}
// The temperature powers of the diffusion coefficients, (tni/T00)^(1.75-s_n)
// (n = 0..2, comp_ng.F90:124-126, minor_ng.F90:95-97) and (T00/tni)^0.25
// (comp_ng.F90:153, minor_ng.F90:136), from one logarithm: exp(y log x)
// differs from pow(x, y) by a few ulp for these arguments.
struct TPow { double df[3], q; };
__device__ __forceinline__ TPow tpows(double tni) {
  const double L = log(tni / T00);
  TPow t;
  for (int n = 0; n < 3; ++n) t.df[n] = exp((1.75 - ssn(n)) * L);
  t.q = exp(-0.25 * L);
  return t;
}
__device__ __forceinline__ double dfactor_at(const tgx_ng_grid &g, int j) {
  // minor_ng.F90:145-154 / comp_ng.F90:113-122
  const double pi = 4.0 * atan(1.0), hor = .25, dtr = pi / 180.;
  if (!(g.difhor > 0)) return 1.;
  double rlat = g.glat[j] * dtr;
  if (fabs(rlat) >= pi / 4.5) return hor + 1.;
  return hor + .5 * (1. + sin(pi * (fabs(rlat) - pi / 9.) / (pi / 4.5)));
}

// =============================================================================
// minor_ng.F90 for a pair of species sharing one pass over the column.
// Species ids: 0=NO 1=N4S 2=N2D 3=AR (comp_*_ng.F90).
struct SpeciesConst {
  double rmx, alfax, phix[4], salfax[3];
};
__device__ __forceinline__ SpeciesConst species_const(int s) {
  SpeciesConst c;
  if (s == 0) { c.rmx = RMNO; c.alfax = 0.; c.phix[0] = 0.814; c.phix[1] = 0.866; c.phix[2] = 0.662; c.phix[3] = 0.926; }
  else if (s == 1 || s == 2) { c.rmx = s == 1 ? RMN4S : RMN2D; c.alfax = 0.; c.phix[0] = 0.651; c.phix[1] = 0.731; c.phix[2] = 0.627; c.phix[3] = 0.741; }
  else { c.rmx = RMAR; c.alfax = 0.17; c.phix[0] = 1.042; c.phix[1] = 1.509; c.phix[2] = 0.732; c.phix[3] = 1.176; }
  for (int m = 0; m < 3; ++m) c.salfax[m] = c.phix[m] - c.phix[3];
  return c;
}

struct MinorShared {   // level quantities shared by the species (minor_ng.F90:66-126)
  double pso1, pso2, pshe, dmdz, do1dz, do2dz, dhedz, tni, barm, inv[9], q;
};
struct MinorSpecies {  // per species and level
  double ax, thdiff, sloss, s0prod, rhs;
};

struct MinorTerms {  // the prelude's right-hand-side parts smooth(fcomp_tm1)*dtx2inv
                     // - advec(fcomp), tiled (MR_NV values by species id), and
                     // planes by species id of the loss / scaled production of
                     // minor_src_kernel
  const double *rt;
  double *st;  // tiled sources (MS_NV values, ms_loss / ms_prod by species)
  // NO / N4S sources as chem_kernel's linear forms in this step's N2D
  // (minor_kernel<.., LIN = true>): the tiled coefficients (ML_NV) and the
  // solved N2D
  const double *lt, *n2d;
};

// Production and loss of species s at level L of column (i, j): the pointwise
// chemistry of comp_n2d_ng.F90:195-203, comp_n4s_ng.F90:122-138 and
// comp_no_ng.F90:48-57 (Ar has none).  n2d_upd is this step's N2D.
__device__ __forceinline__ void minor_sources(const Geo &G, const tgx_ng_fields &in,
                                              const double *__restrict__ n2d_upd, int s, int L,
                                              int i, int j, double &sloss, double &sprod) {
  const int nk = G.nk;
  size_t p = G.a3(L, i, j);
  double xn = in.xnmbar[p];
  if (s == 3) { sloss = 0.; sprod = 0.; return; }
  if (s == 2) {  // comp_n2d_ng.F90:195-203
    double qtefi, nei;
    if (L < nk - 1) {
      qtefi = .5 * (in.qtef[p] + in.qtef[p + G.P]);
      nei = sqrt(in.ne[p] * in.ne[p + G.P]);
    } else {
      qtefi = 1.5 * in.qtef[p] - .5 * in.qtef[p - G.P];
      nei = sqrt(in.ne[p] * in.ne[p] * in.ne[p] / in.ne[p - G.P]);
    }
    sprod = qtefi * BRN2D + in.rk3[p] * in.n2p[p] * xn * in.o1[p] * RIO1 +
            (in.ra1[p] * in.nop[p] * 0.85 + in.ra3[p] * in.n2p[p] * 0.9) * nei;
    sloss = -(xn * (BETA2 * in.o2[p] * RIO2 + BETA4 * in.o1[p] * RIO1 + BETA6 * in.no[p] * RINO) + BETA7 +
              in.beta5[p] * nei + RK10 * in.op[p]);
    return;
  }
  double n2d = n2d_upd[p];
  if (s == 1) {  // comp_n4s_ng.F90:122-138
    double qtefi, beta8i, nei1, nei2;
    if (L < nk - 1) {
      qtefi = .5 * (in.qtef[p] + in.qtef[p + G.P]);
      beta8i = .5 * (in.beta8[p] + in.beta8[p + G.P]);
      nei1 = .5 * (in.ne[p] + in.ne[p + G.P]);
      nei2 = sqrt(in.ne[p] * in.ne[p + G.P]);
    } else {
      qtefi = 0.;  // undefined in the reference (:128); reaches only f_rhs(nk), overwritten
      beta8i = 1.5 * in.beta8[p] - .5 * in.beta8[p - G.P];
      nei1 = 1.5 * in.ne[p] - .5 * in.ne[p - G.P];
      nei2 = sqrt(in.ne[p] * in.ne[p] * in.ne[p] / in.ne[p - G.P]);
    }
    sprod = qtefi * (1. - BRN2D) +
            xn * (n2d * RIN2D * (xn * BETA4 * in.o1[p] * RIO1 + in.beta5[p] * nei1 + BETA7) +
                  beta8i * in.no[p] * RINO) +
            xn * (in.rk2[p] * in.op[p] * in.n2[p] * RIN2 + RK6 * in.nplus[p] * in.o2[p] * RIO2 +
                  RK8 * in.nplus[p] * in.o1[p] * RIO1) +
            nei2 * (in.ra1[p] * in.nop[p] * 0.15 + in.ra3[p] * in.n2p[p] * 1.1);
    sloss = -xn * (in.beta1[p] * in.o2[p] * RIO2 + in.beta3[p] * in.no[p] * RINO +
                   xn * in.beta17[p] * in.o1[p] * RIO1 * in.n2[p] * RIN2) -
            RK4 * in.o2p[p];
    return;
  }
  // NO: comp_no_ng.F90:48-57
  sprod = xn * xn * in.o2[p] * RIO2 * (in.beta1[p] * in.n4s[p] * RIN4S + BETA2 * n2d * RIN2D) +
          xn * xn * xn * in.beta17[p] * in.o1[p] * RIO1 * in.n2[p] * RIN2 * in.n4s[p] * RIN4S;
  double betai;
  if (L < nk - 1)
    betai = .5 * (in.beta8[p] + in.beta9[p] + in.beta9n[p] + in.beta8[p + G.P] + in.beta9[p + G.P] +
                  in.beta9n[p + G.P]);
  else
    betai = 1.5 * (in.beta8[p] + in.beta9[p] + in.beta9n[p]) -
            .5 * (in.beta8[p - G.P] + in.beta9[p - G.P] + in.beta9n[p - G.P]);
  sloss = -xn * (in.beta3[p] * in.n4s[p] * RIN4S + BETA6 * n2d * RIN2D) - RK5 * in.o2p[p] - betai;
}

// sloss and s0prod = sprod*rmx/xnmbar of species SA and SB at every level
// (level-parallel; Ar has no planes).
template <int SA, int SB>
__global__ void __launch_bounds__(256) minor_src_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                        tgx_ng_out out, MinorTerms MT) {
  int nb = gridDim.x;
  long t = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  long npt = (long)G.nlond * G.nlatd * G.nk;
  if (t >= npt) return;
  int i = (int)(t % G.nlond);
  long r = t / G.nlond;
  int j = (int)(r % G.nlatd);
  int k = (int)(r / G.nlatd);
  const size_t p = G.a3(k, i, j);
  const int SP[2] = {SA, SB};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (SP[q] == 3) continue;
    double sl, sp;
    minor_sources(G, in, out.n2d, SP[q], k, i, j, sl, sp);
    const size_t ts = G.tile(k, i, j, 0, MS_NV);
    MT.st[ts + ms_loss(SP[q]) * 64] = sl;
    MT.st[ts + ms_prod(SP[q]) * 64] = sp * species_const(SP[q]).rmx / in.xnmbar[p];
  }
}

template <int SA, int SB, bool CAP, bool LIN = false>
__global__ void __launch_bounds__(256) minor_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                    tgx_ng_out out, double *__restrict__ w2t,
                                                    MinorTerms MT) {
  // Column solve of two species (minor_ng.F90:60-237); their sources come
  // from minor_src_kernel.  Each level's loads are issued together ahead of
  // its arithmetic, with a window of three levels of the shared fields.
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dzp = g.dz, dtx2inv = g.dtx2inv, rdz = 1. / g.dz;
  const size_t c2 = G.a2(i, j);
  const int SP[2] = {SA, SB};
  double *outs[2], *outs_nm[2];
  const double *fc[2], *fctm1[2], *lonb[2], *latb[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int s = SP[q];
    outs[q] = s == 0 ? out.no : s == 1 ? out.n4s : s == 2 ? out.n2d : out.ar;
    outs_nm[q] = s == 0 ? out.no_nm : s == 1 ? out.n4s_nm : s == 2 ? out.n2d_nm : out.ar_nm;
    fc[q] = s == 0 ? in.no : s == 1 ? in.n4s : s == 2 ? in.n2d : in.ar;
    fctm1[q] = s == 0 ? in.no_nm : s == 1 ? in.n4s_nm : s == 2 ? in.n2d_nm : in.ar_nm;
    lonb[q] = s == 0 ? in.lonb_no : s == 1 ? in.lonb_n4s : s == 2 ? in.lonb_n2d : in.lonb_ar;
    latb[q] = s == 0 ? in.latb_no : s == 1 ? in.latb_n4s : s == 2 ? in.latb_n2d : in.latb_ar;
  }
  const double small = 1.e-12;
  double v;
  if (lateral_value(G, in.lonb_no, in.latb_no, 0, i, j, v)) {
  #pragma unroll
  for (int q = 0; q < 2; ++q)
      for (int k = 0; k < nk; ++k) {
        lateral_value(G, lonb[q], latb[q], k, i, j, v);
        size_t p = G.a3(k, i, j);
        double nm = DTSMOOTH * fc[q][p] + DTSMOOTH_DIV2 * (fctm1[q][p] + v);
        outs[q][p] = v > small ? v : small;
        outs_nm[q][p] = nm > small ? nm : small;
      }
    return;
  }
#define F3(f, k) (f)[G.a3((k), i, j)]
  const double pi = 4.0 * atan(1.0);
  double dfac;
  {  // minor_ng.F90:145-154
    const double hor = .25, dtr = pi / 180.;
    if (!(g.difhor > 0)) dfac = 1.;
    else {
      double rlat = g.glat[j] * dtr;
      if (fabs(rlat) >= pi / 4.5) dfac = hor + 1.;
      else dfac = hor + .5 * (1. + sin(pi * (fabs(rlat) - pi / 9.) / (pi / 4.5)));
    }
  }
  const double racs = g.racs[j];
  SpeciesConst sc[2] = {species_const(SA), species_const(SB)};
  const double *n2d_upd = out.n2d;  // the N2D of this step (used by NO and N4S)

  // lower boundary of the major species (:66-69)
  const double o2_0 = F3(in.o2, 0), o1_0 = F3(in.o1, 0), he_0 = F3(in.he, 0);
  const double bo2 = g.b[0] * o2_0 + g.b[1] * o1_0 + g.b[2] * he_0 + g.fb[0];
  const double bo1 = g.b[3] * o2_0 + g.b[4] * o1_0 + g.b[5] * he_0 + g.fb[1];
  const double bhe = g.b[6] * o2_0 + g.b[7] * o1_0 + g.b[8] * he_0 + g.fb[2];
  const double xmbari = 1. / (bo2 * RIO2 + bo1 * RIO1 + (1. - bo2 - bo1 - bhe) * RIN2 + bhe * RIHE);

  struct MR { double o1, o2, he, mbar, tn, barm, w; };  // shared fields at one level
  struct SR { double rhs, sloss, s0prod, la, lb, n2d; };  // one species at one level
  auto load_mr = [&](int L, MR &R) {
    size_t p = G.a3(L, i, j);
    R.o1 = in.o1[p]; R.o2 = in.o2[p]; R.he = in.he[p]; R.mbar = in.mbar[p]; R.tn = in.tn[p];
    R.barm = in.barm[p]; R.w = in.w_upd[p];
  };
  auto load_sr = [&](int L, SR *Q) {
    size_t p = G.a3(L, i, j);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      Q[q].rhs = MT.rt[G.tile(L, i, j, SP[q], MR_NV)];
      const size_t ts = G.tile(L, i, j, 0, MS_NV), tl = G.tile(L, i, j, 0, ML_NV);
      if (SP[q] == 3) { Q[q].sloss = 0.; Q[q].s0prod = 0.; }
      else if (LIN && SP[q] == 0) {  // NO: sloss = lin2 + n2d lin3, s0prod = lin4 + n2d lin5
        Q[q].sloss = MT.lt[tl + 128]; Q[q].la = MT.lt[tl + 192];
        Q[q].s0prod = MT.lt[tl + 256]; Q[q].lb = MT.lt[tl + 320]; Q[q].n2d = MT.n2d[p];
      } else if (LIN && SP[q] == 1) {  // N4S: s0prod = lin0 + n2d lin1
        Q[q].sloss = MT.st[ts + ms_loss(1) * 64];
        Q[q].s0prod = MT.lt[tl]; Q[q].lb = MT.lt[tl + 64]; Q[q].n2d = MT.n2d[p];
      }
      else { Q[q].sloss = MT.st[ts + ms_loss(SP[q]) * 64]; Q[q].s0prod = MT.st[ts + ms_prod(SP[q]) * 64]; }
    }
  };
  // :66-126 interface quantities at L from levels A = L-1 and B = L (L >= 1),
  // or from the lower boundary and B = level 0 (L = 0)
  auto shared_at = [&](int L, const MR &A, const MR &B, MinorShared &S) {
    if (L == 0) {
      S.dmdz = (B.mbar - xmbari) * rdz;
      S.pso1 = .5 * (bo1 + o1_0);
      S.pso2 = .5 * (bo2 + o2_0);
      S.pshe = .5 * (bhe + he_0);
      S.do1dz = (o1_0 - bo1) * rdz;
      S.do2dz = (o2_0 - bo2) * rdz;
      S.dhedz = (he_0 - bhe) * rdz;
      S.tni = in.tlbc[c2];
    } else {
      S.dmdz = (B.mbar - A.mbar) * rdz;
      S.pso1 = .5 * (B.o1 + A.o1);
      S.pso2 = .5 * (B.o2 + A.o2);
      S.pshe = .5 * (B.he + A.he);
      S.do1dz = (B.o1 - A.o1) * rdz;
      S.do2dz = (B.o2 - A.o2) * rdz;
      S.dhedz = (B.he - A.he) * rdz;
      S.tni = L == nk - 1 ? A.tn : .5 * (B.tn + A.tn);  // :89-93
    }
    S.barm = B.barm;
    double al[9];
    const TPow tp = tpows(S.tni);
    const double *df = tp.df;
    S.q = tp.q;
    al[0] = -(PHI(1, 4) + (PHI(1, 2) - PHI(1, 4)) * S.pso1 + (df[0] * PHI(1, 3) - PHI(1, 4)) * S.pshe);
    al[4] = -(PHI(2, 4) + (PHI(2, 1) - PHI(2, 4)) * S.pso2 + (df[1] * PHI(2, 3) - PHI(2, 4)) * S.pshe);
    al[8] = -(df[2] * PHI(3, 4) + (df[0] * PHI(3, 1) - df[2] * PHI(3, 4)) * S.pso2 +
              (df[1] * PHI(3, 2) - df[2] * PHI(3, 4)) * S.pso1);
    al[3] = (PHI(1, 2) - PHI(1, 4)) * S.pso2;
    al[6] = (df[0] * PHI(1, 3) - PHI(1, 4)) * S.pso2;
    al[1] = (PHI(2, 1) - PHI(2, 4)) * S.pso1;
    al[7] = (df[1] * PHI(2, 3) - PHI(2, 4)) * S.pso1;
    al[2] = (df[0] * PHI(3, 1) - df[2] * PHI(3, 4)) * S.pshe;
    al[5] = (df[1] * PHI(3, 2) - df[2] * PHI(3, 4)) * S.pshe;
    matinv3_r(al, S.inv);
  };
  auto species_at = [&](int q, const MinorShared &S, double dtnidz, const SR &R, MinorSpecies &T) {
    const SpeciesConst &c = sc[q];
    double sai[3];
    for (int n = 0; n < 3; ++n) {
      double s = 0.;
      for (int m = 0; m < 3; ++m) s = s + c.salfax[m] * S.inv[m + 3 * n];
      sai[n] = s;
    }
    const double rb = 1. / S.barm;
    double ex = 1. - (c.rmx + S.dmdz) * rb +
                sai[0] * (S.do2dz - (1. - (RMO2 + S.dmdz) * rb) * S.pso2) +
                sai[1] * (S.do1dz - (1. - (RMO1 + S.dmdz) * rb) * S.pso1) +
                sai[2] * (S.dhedz - (1. - (RMHE + S.dmdz) * rb) * S.pshe);
    // phi_x(He) in the nest form's sum (minor_ng.F90:135-136), phi_x(N2) in
    // the global minor.F:359-362
    T.ax = -(S.barm * (1. / (TAU * RMN2)) * S.q /
             (c.phix[G.glb ? 3 : 2] + c.salfax[0] * S.pso2 + c.salfax[1] * S.pso1 + c.salfax[2] * S.pshe));
    T.thdiff = ex - c.alfax * dtnidz / (dzp * S.tni);
    T.sloss = R.sloss;
    T.s0prod = R.s0prod;
    if (LIN && SP[q] == 0) {
      T.sloss = R.sloss + R.n2d * R.la;
      T.s0prod = R.s0prod + R.n2d * R.lb;
    } else if (LIN && SP[q] == 1) {
      T.s0prod = R.s0prod + R.n2d * R.lb;
    }
    T.rhs = R.rhs;
  };
  // lower-boundary data flbc(:,:,1:3) of each species (comp_*_ng.F90)
  double l1[2], l2[2], l3[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int s = SP[q];
    size_t p = G.a3(0, i, j);
    double xb = in.xnmbari[p];
    l1[q] = 0.;
    l2[q] = 1.;
    if (s == 0) l3[q] = -4.e6 * RMNO / xb;
    else if (s == 1)
      l3[q] = -(RMN4S / xb *
                (in.qtef[p] * (1. - BRN2D) / xb +
                 n2d_upd[p] * RIN2D * (BETA4 * xb * in.o1[p] * RIO1 + in.beta5[p] * in.ne[p] + BETA7) +
                 in.beta8[p] * in.no[p] * RINO) /
                (in.beta1[p] * in.o2[p] * RIO2 + in.beta3[p] * in.no[p] * RINO +
                 xb * in.beta17[p] * in.o1[p] * RIO1 * in.n2[p] * RIN2));
    else if (s == 2)
      l3[q] = -(RMN2D / xb * in.qtef[p] * BRN2D /
                (xb * (BETA2 * in.o2[p] * RIO2 + BETA4 * in.o1[p] * RIO1 + BETA6 * in.no[p] * RINO) + BETA7 +
                 in.beta5[p] * in.ne[p]));
    else l3[q] = -sqrt(g.ar_glbm1 * g.ar_glbm2);
  }
  const double fubc = 0.;  // every species sets fubc = 0 (comp_*_ng.F90)

  // window of the shared fields at levels k (R0), k+1 (R1), k+2 (R2)
  MR R0, R1, R2;
  SR Y[2];
  load_mr(0, R0);
  load_mr(1, R1);
  load_sr(0, Y);
  MinorShared Sc, Sn;
  MinorSpecies Tc[2], Tn[2];
  shared_at(0, R0, R0, Sc);
  // tni at level 1 (:89-93)
  double tni_1 = 1 == nk - 1 ? R0.tn : .5 * (R1.tn + R0.tn);
#pragma unroll
  for (int q = 0; q < 2; ++q) species_at(q, Sc, tni_1 - Sc.tni, Y[q], Tc[q]);
  double w1[2] = {0., 0.}, w3[2] = {0., 0.}, rprev[2] = {0., 0.};
  double wic = .5 * (R0.w + R1.w);
  for (int k = 0; k < nk; ++k) {
    if (k + 2 <= nk - 1) load_mr(k + 2, R2);
    if (k + 1 <= nk - 1) load_sr(k + 1, Y);
    const double difk0 = g.difk[k], difk1 = k + 1 <= nk - 1 ? g.difk[k + 1] : 0., expzk = g.expz[k];
    asm volatile("" ::: "memory");
    double wi_n = 0.;
    if (k + 1 <= nk - 1) {
      shared_at(k + 1, R0, R1, Sn);
      double tni_p = 0.;
      if (k + 2 <= nk - 1) tni_p = k + 2 == nk - 1 ? R1.tn : .5 * (R2.tn + R1.tn);
      double dtnidz = (k + 1 == nk - 1) ? Sn.tni - Sc.tni : (tni_p - Sc.tni) / 2.;
#pragma unroll
      for (int q = 0; q < 2; ++q) species_at(q, Sn, dtnidz, Y[q], Tn[q]);
      wi_n = k + 1 < nk - 1 ? .5 * (R1.w + R2.w) : 1.5 * R1.w - .5 * R0.w;
    }
    const double dmdzc = Sc.dmdz / Sc.barm;
    const double dmdzn = k + 1 <= nk - 1 ? Sn.dmdz / Sn.barm : 0.;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double pc, qc, rc, fr;
      if (k <= nk - 2) {
        double molp = Tc[q].ax * (rdz + .5 * Tc[q].thdiff);
        double molr = Tn[q].ax * (rdz - .5 * Tn[q].thdiff);
        double molq = Tc[q].ax * (rdz - .5 * Tc[q].thdiff) + Tn[q].ax * (rdz + .5 * Tn[q].thdiff);
        double eddyp = g.expzmid_inv * difk0 * (rdz - .5 * dmdzc);
        double eddyr = g.expzmid * difk1 * (rdz + .5 * dmdzn);
        double eddyq = g.expzmid_inv * difk0 * (rdz + .5 * dmdzc) +
                       g.expzmid * difk1 * (rdz - .5 * dmdzn);
        pc = molp * rdz - expzk * (eddyp * dfac + .5 * wic) * rdz;
        rc = molr * rdz - expzk * (eddyr * dfac - .5 * wic) * rdz;
        qc = -molq * rdz + expzk * (eddyq * dfac * rdz + dtx2inv - Tc[q].sloss);
        fr = Tc[q].rhs + Tc[q].s0prod;
        fr = fr * expzk;
        if (k == 0) {
          qc = qc + pc * (l1[q] + .5 * l2[q] * dzp) / (l1[q] - .5 * l2[q] * dzp);
          fr = fr - pc * l3[q] * dzp / (l1[q] - .5 * l2[q] * dzp);
          pc = 0.;
        }
      } else {
        pc = 1. + .5 * dzp * Tc[q].thdiff;
        qc = pc - 2.;
        rc = 0.;
        fr = -GRAV * sc[q].rmx * fubc * dzp / (P0 * Tc[q].ax * AVO);
      }
      if (CAP) {
        const int pl = CAP_MIN + 4 * SP[q];
        cap_put(g, G, pl + 0, i, j, k, pc); cap_put(g, G, pl + 1, i, j, k, qc);
        cap_put(g, G, pl + 2, i, j, k, rc); cap_put(g, G, pl + 3, i, j, k, fr);
      }
      // Thomas forward, rows 1..nk (minor_ng.F90:200)
      if (k == 0) {
        w1[q] = qc;
        w3[q] = fr / w1[q];
      } else {
        double w2 = rprev[q] / w1[q];
        w2t[G.tile(k - 1, i, j, q, MW_NV)] = w2;
        w1[q] = qc - pc * w2;
        w3[q] = (fr - pc * w3[q]) / w1[q];
      }
      w2t[G.tile(k, i, j, 2 + q, MW_NV)] = w3[q];  // forward solution
      rprev[q] = rc;
    }
    Sc = Sn;
    Tc[0] = Tn[0];
    Tc[1] = Tn[1];
    wic = wi_n;
    R0 = R1;
    R1 = R2;
  }
  // back substitution and output stage (:234-237); the next level's loads
  // are in flight while this level is stored
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    struct BL { double x, w2, f, fm; };
    auto load_bl = [&](int k, BL &B) {
      size_t p = G.a3(k, i, j);
      B.x = w2t[G.tile(k, i, j, 2 + q, MW_NV)];
      B.w2 = k < nk - 1 ? w2t[G.tile(k, i, j, q, MW_NV)] : 0.;
      B.f = fc[q][p];
      B.fm = fctm1[q][p];
    };
    BL bc, bn;
    load_bl(nk - 1, bc);
    double xn = w3[q];
    for (int k = nk - 1; k >= 0; --k) {
      if (k > 0) load_bl(k - 1, bn);
      asm volatile("" ::: "memory");
      size_t p = G.a3(k, i, j);
      double x = k == nk - 1 ? xn : bc.x - bc.w2 * xn;
      xn = x;
      if (G.raw_row(j)) {  // the raw solution; filter_minor and the post pass follow (minor.F:546-597)
        outs[q][p] = x;
        bc = bn;
        continue;
      }
      double nm = DTSMOOTH * bc.f + DTSMOOTH_DIV2 * (bc.fm + x);
      outs[q][p] = x > small ? x : small;
      outs_nm[q][p] = nm > small ? nm : small;
      bc = bn;
    }
  }
#undef F3
}


// =============================================================================
// All pointwise chemistry of one step that does not depend on this step's
// solutions, in one level-parallel pass that reads every chemistry field once:
//   comp_o2o_ng fs planes (comp_o2o_ng.F90:44-81, levels < nk-1);
//   N2D loss and scaled production (comp_n2d_ng.F90:195-203);
//   N4S loss (comp_n4s_ng.F90:134-137);
//   NO and N4S production, and NO loss, split as a + n2d*b in this step's N2D
//   (comp_no_ng.F90:48-57, comp_n4s_ng.F90:129-133), scaled by rmx/xnmbar;
//   O+ xiop2p, xiop2d, loss and production (oplus_ng.F90:288-299).
// All values are computed before any is stored, so shared fields load once.
struct ChemOut {
  double *ci;                          // tiled comp inputs (fs(:,1:2,1:2) written)
  double *fs0[2];                      // fs(:,1,0), fs(:,2,0) planes of comp_o2o
  double *st;                          // tiled minor sources: N2D loss, s0prod, N4S loss
  double *lt;                          // tiled N4S s0prod a, b; NO sloss a, b; NO s0prod a, b
  double *oploss, *opprod;             // O+ loss and production
};

// Fields chem_kernel reads at two neighbouring levels (interface means, the
// top level's extrapolation): one level of them is carried between levels.
// Loads and stores at a 32-bit byte offset from a kernel-argument base: the
// address stays one SGPR pair plus one VGPR (every level plane of the
// largest grids is well under 4 GiB).
__device__ __forceinline__ double ldo(const double *base, unsigned off) {
  return *(const double *)((const char *)base + off);
}
__device__ __forceinline__ void sto(double *base, unsigned off, double v) {
  *(double *)((char *)base + off) = v;
}

struct ChemPair { double ne, beta8, rj, qo2p, qop, qtef, beta9, beta9n, qop2p, qop2d; };
__device__ __forceinline__ void chem_pair_load(const tgx_ng_fields &in, unsigned p, ChemPair &v) {
  v.ne = ldo(in.ne, p); v.beta8 = ldo(in.beta8, p); v.rj = ldo(in.rj, p); v.qo2p = ldo(in.qo2p, p); v.qop = ldo(in.qop, p);
  v.qtef = ldo(in.qtef, p); v.beta9 = ldo(in.beta9, p); v.beta9n = ldo(in.beta9n, p); v.qop2p = ldo(in.qop2p, p);
  v.qop2d = ldo(in.qop2d, p);
}

// Number of level runs of chem_kernel: one thread walks about CHEM_RUN levels
// of a column, upward, loading each level once.
constexpr int CHEM_RUN = 8;
__host__ __device__ inline int chem_runs(int nk) { return (nk + CHEM_RUN - 1) / CHEM_RUN; }

// One level of the chemistry pass: c = the paired fields at this level, o =
// at the level above (below at the top level).
__device__ __forceinline__ void chem_level(const tgx_ng_fields &in, const tgx_ng_out &out, const ChemOut &C,
                                           unsigned p, size_t tci, size_t tms, size_t tml, bool top,
                                           const ChemPair &c, const ChemPair &o) {
  const double xn = ldo(in.xnmbar, p), ne = c.ne;
  const double o1 = ldo(in.o1, p), o2 = ldo(in.o2, p), n2 = ldo(in.n2, p), op = ldo(in.op, p), no = ldo(in.no, p), n4s = ldo(in.n4s, p);
  const double n2d = ldo(in.n2d, p), o2p = ldo(in.o2p, p), nplus = ldo(in.nplus, p), n2p = ldo(in.n2p, p), nop = ldo(in.nop, p);
  const double rk1 = ldo(in.rk1, p), rk2 = ldo(in.rk2, p), rk3 = ldo(in.rk3, p), ra1 = ldo(in.ra1, p), ra3 = ldo(in.ra3, p);
  const double beta1 = ldo(in.beta1, p), beta3 = ldo(in.beta3, p), beta5 = ldo(in.beta5, p), beta17 = ldo(in.beta17, p);
  const double rkm12 = ldo(in.rkm12, p), mbar = ldo(in.mbar, p), ra2 = ldo(in.ra2, p);
  const double rk19 = ldo(in.rk19, p), rk20 = ldo(in.rk20, p), rk25 = ldo(in.rk25, p);
  const double beta8 = c.beta8, beta8_o = o.beta8;
  // ---- comp_o2o_ng fs (levels < nk-1)
  double fsv[6] = {0., 0., 0., 0., 0., 0.};
  if (!top) {
    double rji = 0.5 * (c.rj + o.rj);
    double qo2pi = 0.5 * (c.qo2p + o.qo2p);
    double qopi = 0.5 * (c.qop + o.qop);
    double beta8i = 0.5 * (beta8 + beta8_o);
    double nei = sqrt(ne * o.ne);
    double pox1 = xn * xn * (beta3 * n4s / RMN4S * no / RMNO + BETA6 * n2d / RMN2D * no / RMNO) +
                  beta8i * no / RMNO * xn + xn * (RK4 * o2p * n4s / RMN4S + RK10 * op * n2d / RMN2D) +
                  (ra1 * nop + 2. * ra2 * o2p) * nei;
    double pox2 = xn * (beta1 * n4s / RMN4S + BETA2 * n2d / RMN2D) + rk1 * op + RK7 * nplus + 2. * rji;
    double lox1 = 2. * rkm12 * xn / mbar;
    double lox2 = rk3 * n2p + RK8 * nplus;
    double lox3 = qopi;
    double po21 = rkm12 * xn / mbar;
    double po23 = RK5 * no / RMNO * o2p * xn;
    double lo21 = xn * (beta1 * n4s / RMN4S + BETA2 * n2d / RMN2D) + rk1 * op + (RK6 + RK7) * nplus +
                  RK9 * n2p + rji;
    double lo22 = qo2pi;
    fsv[0] = -lo21;
    fsv[1] = xn * po21 * o1 * (RIO1 * RIO1) * RMO2;
    fsv[2] = pox2 * RMO1 * RIO2;
    fsv[3] = -lox2 - lox1 * o1 * RIO1 * xn;
    fsv[4] = (po23 - lo22) * RMO2 / xn;
    fsv[5] = (pox1 - lox3) * RMO1 / xn;
  }
  // ---- N2D (comp_n2d_ng.F90:195-203)
  double n2d_sl, n2d_sp;
  {
    double qtefi, nei;
    if (!top) {
      qtefi = .5 * (c.qtef + o.qtef);
      nei = sqrt(ne * o.ne);
    } else {
      qtefi = 1.5 * c.qtef - .5 * o.qtef;
      nei = sqrt(ne * ne * ne / o.ne);
    }
    n2d_sp = qtefi * BRN2D + rk3 * n2p * xn * o1 * RIO1 + (ra1 * nop * 0.85 + ra3 * n2p * 0.9) * nei;
    n2d_sl = -(xn * (BETA2 * o2 * RIO2 + BETA4 * o1 * RIO1 + BETA6 * no * RINO) + BETA7 + beta5 * nei +
               RK10 * op);
  }
  // ---- N4S (comp_n4s_ng.F90:122-138): production a + n2d*b, loss
  double n4s_a, n4s_b, n4s_sl;
  {
    double qtefi, beta8i, nei1, nei2;
    if (!top) {
      qtefi = .5 * (c.qtef + o.qtef);
      beta8i = .5 * (beta8 + beta8_o);
      nei1 = .5 * (ne + o.ne);
      nei2 = sqrt(ne * o.ne);
    } else {
      qtefi = 0.;  // undefined in the reference (:128); reaches only f_rhs(nk), overwritten
      beta8i = 1.5 * beta8 - .5 * beta8_o;
      nei1 = 1.5 * ne - .5 * o.ne;
      nei2 = sqrt(ne * ne * ne / o.ne);
    }
    n4s_a = qtefi * (1. - BRN2D) + xn * (beta8i * no * RINO) +
            xn * (rk2 * op * n2 * RIN2 + RK6 * nplus * o2 * RIO2 + RK8 * nplus * o1 * RIO1) +
            nei2 * (ra1 * nop * 0.15 + ra3 * n2p * 1.1);
    n4s_b = xn * RIN2D * (xn * BETA4 * o1 * RIO1 + beta5 * nei1 + BETA7);
    n4s_sl = -xn * (beta1 * o2 * RIO2 + beta3 * no * RINO + xn * beta17 * o1 * RIO1 * n2 * RIN2) - RK4 * o2p;
  }
  // ---- NO (comp_no_ng.F90:48-57): production and loss a + n2d*b
  double no_spa, no_spb, no_sla, no_slb;
  {
    no_spa = xn * xn * o2 * RIO2 * (beta1 * n4s * RIN4S) + xn * xn * xn * beta17 * o1 * RIO1 * n2 * RIN2 * n4s * RIN4S;
    no_spb = xn * xn * o2 * RIO2 * (BETA2 * RIN2D);
    double betai = !top ? .5 * (beta8 + c.beta9 + c.beta9n + beta8_o + o.beta9 + o.beta9n)
                        : 1.5 * (beta8 + c.beta9 + c.beta9n) - .5 * (beta8_o + o.beta9 + o.beta9n);
    no_sla = -xn * (beta3 * n4s * RIN4S) - RK5 * o2p - betai;
    no_slb = -xn * (BETA6 * RIN2D);
  }
  // ---- O+ (oplus_ng.F90:288-299)
  double qop2pi = !top ? .5 * (c.qop2p + o.qop2p) : 1.5 * c.qop2p - .5 * o.qop2p;
  double qop2di = !top ? .5 * (c.qop2d + o.qop2d) : 1.5 * c.qop2d - .5 * o.qop2d;
  double x2p = qop2pi / (xn * ((RK16 + RK17) * n2 * RIN2 + RK18 * o1 * RIO1) + (rk19 + rk20) * ne + RK21 + RK22);
  double x2d = (qop2di + (rk20 * ne + RK22) * x2p) /
               (xn * (RK23 * n2 * RIN2 + RK24 * o1 * RIO1 + RK26 * o2 * RIO2) + rk25 * ne + RK27);
  double oploss = xn * (rk1 * o2 * RIO2 + rk2 * n2 * RIN2 + RK10 * n2d * RIN2D);
  double opprod = 0.;
  if (!top) {
    double qopi = .5 * (c.qop + o.qop);
    opprod = qopi + (rk19 * ne + RK21) * x2p + (rk25 * ne + RK27) * x2d + (RK18 * x2p + RK24 * x2d) * o1 * RIO1 * xn;
  }
  // ---- stores
  if (!top) {
    for (int q = 0; q < 4; ++q) C.ci[tci + (size_t)(3 + q) * 64] = fsv[q];
    sto(C.fs0[0], p, fsv[4]);
    sto(C.fs0[1], p, fsv[5]);
  }
  const size_t ts = tms, tl = tml;
  C.st[ts + 0 * 64] = n2d_sl;
  C.st[ts + 1 * 64] = n2d_sp * RMN2D / xn;
  C.st[ts + 2 * 64] = n4s_sl;
  C.lt[tl + 0 * 64] = n4s_a * RMN4S / xn;
  C.lt[tl + 1 * 64] = n4s_b * RMN4S / xn;
  C.lt[tl + 2 * 64] = no_sla;
  C.lt[tl + 3 * 64] = no_slb;
  C.lt[tl + 4 * 64] = no_spa * RMNO / xn;
  C.lt[tl + 5 * 64] = no_spb * RMNO / xn;
  sto(out.xiop2p, p, x2p);
  sto(out.xiop2d, p, x2d);
  sto(C.oploss, p, oploss);
  if (!top) sto(C.opprod, p, opprod);
}

// The pointwise chemistry of the whole step (comp_o2o_ng fs, the N2D / N4S /
// NO sources, O+ loss and production and xiop2p/xiop2d) in one pass.  A
// thread takes one of nrun runs of levels of one column, bottom to top, so the
// fields read at two levels are loaded once each (runs of about CHEM_RUN
// levels; shorter on small grids, which would not fill the chip).
__global__ void __launch_bounds__(256) chem_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                   tgx_ng_out out, ChemOut C, int nrun) {
  int nb = gridDim.x;
  long t = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  const int nk = G.nk;
  long npt = (long)G.nlond * G.nlatd * nrun;
  if (t >= npt) return;
  int i = (int)(t % G.nlond);
  long r = t / G.nlond;
  int j = (int)(r % G.nlatd);
  int run = (int)(r / G.nlatd);
  const int k0 = (int)((long)run * nk / nrun), k1 = (int)((long)(run + 1) * nk / nrun);
  ChemPair c, o, d;
  const unsigned PB = (unsigned)(G.P * sizeof(double));  // level stride in bytes
  chem_pair_load(in, (unsigned)(G.a3(k0, i, j) * sizeof(double)), c);
  if (k0 == nk - 1)  // a run that is only the top level
    chem_pair_load(in, (unsigned)(G.a3(k0 - 1, i, j) * sizeof(double)), d);
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    const bool top = k == nk - 1;
    const unsigned p = (unsigned)(G.a3(k, i, j) * sizeof(double));
    if (!top)
      chem_pair_load(in, p + PB, o);
    else
      o = d;
    chem_level(in, out, C, p, G.tile(k, i, j, 0, CI_NV), G.tile(k, i, j, 0, MS_NV), G.tile(k, i, j, 0, ML_NV),
               top, c, o);
    d = c;
    c = o;
  }
}

// =============================================================================
// One wavefront per column.  Lane l holds level l ("lo") and level l+64
// ("hi"), so a wave covers up to 128 levels; neighbouring levels come from
// other lanes, and the column's tridiagonal systems are solved in registers by
// parallel cyclic reduction (no ratio planes, no second pass over memory).  A
// workgroup holds consecutive columns of one latitude row, so the waves of a
// workgroup share the cache lines of every level plane.
//
// Lane exchanges: values one lane away (the neighbouring level, and the first
// reduction step) move by DPP wave rotations, which run on the VALU; longer
// distances go through ds_bpermute.  Both return, for every lane, the value
// the source lane holds, so the two forms give the same bits.  Every lane of
// the wavefront is active wherever they are used.

// x of lane (lane - 1) & 63 (wave_ror:1) and of lane (lane + 1) & 63
// (wave_rol:1), one 32-bit half at a time.
template <int CTRL>
__device__ __forceinline__ double dpp_rot(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
constexpr int DPP_WAVE_ROL1 = 0x134, DPP_WAVE_ROR1 = 0x13C;
__device__ __forceinline__ double lane_prev(double x) { return dpp_rot<DPP_WAVE_ROR1>(x); }
__device__ __forceinline__ double lane_next(double x) { return dpp_rot<DPP_WAVE_ROL1>(x); }

// Partners of a reduction step at distance d: m = lane (lane - d) & 63, p =
// lane (lane + d) & 63.
struct XchgRot {  // d = 1
  static constexpr bool same = false;
  __device__ double m(double v) const { return lane_prev(v); }
  __device__ double p(double v) const { return lane_next(v); }
};
struct XchgPerm {  // 1 < d < 64
  static constexpr bool same = false;
  int sm, sp;
  __device__ XchgPerm(int lane, int d) : sm((lane - d) & 63), sp((lane + d) & 63) {}
  __device__ double m(double v) const { return __shfl(v, sm, 64); }
  __device__ double p(double v) const { return __shfl(v, sp, 64); }
};
struct XchgXor32 {  // d = 32: both partners are lane ^ 32 (same = true: one exchange serves both)
  static constexpr bool same = true;
  int s;
  __device__ explicit XchgXor32(int lane) : s(lane ^ 32) {}
  __device__ double m(double v) const { return __shfl(v, s, 64); }
  __device__ double p(double v) const { return __shfl(v, s, 64); }
};
struct XchgOwn {  // d = 64 with two rows per lane: the partner row is in this lane
  static constexpr bool same = true;
  __device__ double m(double v) const { return v; }
  __device__ double p(double v) const { return v; }
};

// Value at level L-1 (dn) and L+1 (up) for the lo and hi rows of this lane.
// Lane 0 lo (level -1) and lane 63 hi (level 128) receive unused values.
__device__ __forceinline__ void wave_dn(double lo, double hi, int lane, double &dlo, double &dhi) {
  dlo = lane_prev(lo);
  dhi = lane_prev(lane == 63 ? lo : hi);  // lane 0 hi (level 64) takes lane 63 lo
}
__device__ __forceinline__ void wave_up(double lo, double hi, int lane, double &ulo, double &uhi) {
  ulo = lane_next(lane == 0 ? hi : lo);  // lane 63 lo (level 63) takes lane 0 hi
  uhi = lane_next(hi);
}

// The same for NH rows per lane: NH = 1 when the column has at most 64
// levels (lane l holds level l only; hi is neither read nor written).
template <int NH>
__device__ __forceinline__ void wave_dn_h(double lo, double hi, int lane, double &dlo, double &dhi) {
  if constexpr (NH == 2) {
    wave_dn(lo, hi, lane, dlo, dhi);
  } else {
    dlo = lane_prev(lo);
  }
}
template <int NH>
__device__ __forceinline__ void wave_up_h(double lo, double hi, int lane, double &ulo, double &uhi) {
  if constexpr (NH == 2) {
    wave_up(lo, hi, lane, ulo, uhi);
  } else {
    ulo = lane_next(lo);
  }
}

struct TriRow { double a, b, c, f; };  // a x(k-1) + b x(k) + c x(k+1) = f

__device__ __forceinline__ TriRow tri_ident() { return TriRow{0., 1., 0., 0.}; }

template <class X>
__device__ __forceinline__ TriRow tri_m(const X &x, const TriRow &r) {
  return TriRow{x.m(r.a), x.m(r.b), x.m(r.c), x.m(r.f)};
}
template <class X>
__device__ __forceinline__ TriRow tri_p(const X &x, const TriRow &r) {
  return TriRow{x.p(r.a), x.p(r.b), x.p(r.c), x.p(r.f)};
}

// Eliminates the couplings of row r to rows m (r-d) and p (r+d).
__device__ __forceinline__ TriRow tri_reduce(const TriRow &r, const TriRow &m, const TriRow &p) {
  double alpha = -r.a / m.b, gamma = -r.c / p.b;
  TriRow o;
  o.a = alpha * m.a;
  o.c = gamma * p.c;
  o.b = r.b + alpha * m.c + gamma * p.a;
  o.f = r.f + alpha * m.f + gamma * p.f;
  return o;
}

// Parallel cyclic reduction of NS independent 128-row systems held as
// (lo, hi) rows per lane.  Rows beyond the system are identity rows; the
// first row has a = 0 and the last c = 0.  On return x = f / b.
template <int NS, class X>
__device__ __forceinline__ void pcr_step(TriRow (&lo)[NS], TriRow (&hi)[NS], int lane, int d, const X &x) {
  const bool has_m = lane >= d, low_p = lane + d < 64;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const TriRow Lm = tri_m(x, lo[q]), Hm = tri_m(x, hi[q]);
    const TriRow Lp = X::same ? Lm : tri_p(x, lo[q]), Hp = X::same ? Hm : tri_p(x, hi[q]);
    TriRow lm = has_m ? Lm : tri_ident();
    TriRow lp = low_p ? Lp : Hp;
    TriRow hm = has_m ? Hm : Lm;
    TriRow hp = low_p ? Hp : tri_ident();
    TriRow nlo = tri_reduce(lo[q], lm, lp);
    TriRow nhi = tri_reduce(hi[q], hm, hp);
    lo[q] = nlo;
    hi[q] = nhi;
  }
}
template <int NS>
__device__ __forceinline__ void wave_pcr(TriRow (&lo)[NS], TriRow (&hi)[NS], int lane) {
  pcr_step<NS>(lo, hi, lane, 1, XchgRot{});
#pragma unroll
  for (int d = 2; d < 32; d <<= 1) pcr_step<NS>(lo, hi, lane, d, XchgPerm(lane, d));
  pcr_step<NS>(lo, hi, lane, 32, XchgXor32(lane));
#pragma unroll
  for (int q = 0; q < NS; ++q) {  // d = 64: lo and hi of the same lane
    TriRow nlo = tri_reduce(lo[q], tri_ident(), hi[q]);
    TriRow nhi = tri_reduce(hi[q], lo[q], tri_ident());
    lo[q] = nlo;
    hi[q] = nhi;
  }
}

// The same for one row per lane (systems of at most 64 rows): partners past
// either end of the wave are absent.
template <int NS, class X>
__device__ __forceinline__ void pcr1_step(TriRow (&lo)[NS], int lane, int d, const X &x) {
  const bool has_m = lane >= d, has_p = lane + d < 64;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const TriRow Lm = tri_m(x, lo[q]), Lp = X::same ? Lm : tri_p(x, lo[q]);
    lo[q] = tri_reduce(lo[q], has_m ? Lm : tri_ident(), has_p ? Lp : tri_ident());
  }
}
template <int NS>
__device__ __forceinline__ void wave_pcr1(TriRow (&lo)[NS], int lane) {
  pcr1_step<NS>(lo, lane, 1, XchgRot{});
#pragma unroll
  for (int d = 2; d < 32; d <<= 1) pcr1_step<NS>(lo, lane, d, XchgPerm(lane, d));
  pcr1_step<NS>(lo, lane, 32, XchgXor32(lane));
}

// Block rows A x(k-1) + B x(k) + C x(k+1) = F with N x N blocks, row-major
// (e = n*N + m is row n, column m).
template <int N>
struct BlkRow { double A[N * N], B[N * N], C[N * N], F[N]; };

template <int N>
__device__ __forceinline__ void blk_inv(const double *M, double *R);
template <>
__device__ __forceinline__ void blk_inv<2>(const double *M, double *R) {
  double det = M[0] * M[3] - M[1] * M[2];
  R[0] = M[3] / det; R[1] = -M[1] / det; R[2] = -M[2] / det; R[3] = M[0] / det;
}
template <>
__device__ __forceinline__ void blk_inv<3>(const double *M, double *R) {
  double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
  double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
  double r = 1. / det;
  R[0] = c00 * r; R[3] = c01 * r; R[6] = c02 * r;
  R[1] = (M[2] * M[7] - M[1] * M[8]) * r; R[4] = (M[0] * M[8] - M[2] * M[6]) * r;
  R[7] = (M[1] * M[6] - M[0] * M[7]) * r;
  R[2] = (M[1] * M[5] - M[2] * M[4]) * r; R[5] = (M[2] * M[3] - M[0] * M[5]) * r;
  R[8] = (M[0] * M[4] - M[1] * M[3]) * r;
}
template <int N>
__device__ __forceinline__ void blk_mm(const double *X, const double *Y, double *Z) {  // Z = X Y
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int m = 0; m < N; ++m) {
      double t = 0.;
#pragma unroll
      for (int l = 0; l < N; ++l) t += X[n * N + l] * Y[l * N + m];
      Z[n * N + m] = t;
    }
}
template <int N>
__device__ __forceinline__ void blk_mv(const double *X, const double *y, double *z) {  // z = X y
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double t = 0.;
#pragma unroll
    for (int l = 0; l < N; ++l) t += X[n * N + l] * y[l];
    z[n] = t;
  }
}

// A row divided by its diagonal block: Ah = B^-1 A, Ch = B^-1 C, Fh = B^-1 F.
template <int N>
struct BlkNorm { double A[N * N], C[N * N], F[N]; };

template <int N>
__device__ __forceinline__ BlkNorm<N> blk_normalize(const BlkRow<N> &r) {
  double Bi[N * N];
  blk_inv<N>(r.B, Bi);
  BlkNorm<N> o;
  blk_mm<N>(Bi, r.A, o.A);
  blk_mm<N>(Bi, r.C, o.C);
  blk_mv<N>(Bi, r.F, o.F);
  return o;
}
template <int N, bool M, class X>
__device__ __forceinline__ BlkNorm<N> blk_xchg(const X &x, const BlkNorm<N> &r) {  // M: from m, else p
  BlkNorm<N> o;
#pragma unroll
  for (int e = 0; e < N * N; ++e) {
    o.A[e] = M ? x.m(r.A[e]) : x.p(r.A[e]);
    o.C[e] = M ? x.m(r.C[e]) : x.p(r.C[e]);
  }
#pragma unroll
  for (int e = 0; e < N; ++e) o.F[e] = M ? x.m(r.F[e]) : x.p(r.F[e]);
  return o;
}
template <int N>
__device__ __forceinline__ BlkNorm<N> blk_zero() {
  BlkNorm<N> o;
#pragma unroll
  for (int e = 0; e < N * N; ++e) { o.A[e] = 0.; o.C[e] = 0.; }
#pragma unroll
  for (int e = 0; e < N; ++e) o.F[e] = 0.;
  return o;
}
template <int N>
__device__ __forceinline__ BlkNorm<N> blk_sel(bool c, const BlkNorm<N> &x, const BlkNorm<N> &y) {
  BlkNorm<N> o;
#pragma unroll
  for (int e = 0; e < N * N; ++e) { o.A[e] = c ? x.A[e] : y.A[e]; o.C[e] = c ? x.C[e] : y.C[e]; }
#pragma unroll
  for (int e = 0; e < N; ++e) o.F[e] = c ? x.F[e] : y.F[e];
  return o;
}
// Eliminates the couplings of r to the normalised rows m (r-d) and p (r+d).
template <int N>
__device__ __forceinline__ void blk_reduce(BlkRow<N> &r, const BlkNorm<N> &m, const BlkNorm<N> &p) {
  double T[N * N], nA[N * N], nC[N * N], t[N];
  blk_mm<N>(r.A, m.C, T);
#pragma unroll
  for (int e = 0; e < N * N; ++e) r.B[e] -= T[e];
  blk_mm<N>(r.C, p.A, T);
#pragma unroll
  for (int e = 0; e < N * N; ++e) r.B[e] -= T[e];
  blk_mv<N>(r.A, m.F, t);
#pragma unroll
  for (int e = 0; e < N; ++e) r.F[e] -= t[e];
  blk_mv<N>(r.C, p.F, t);
#pragma unroll
  for (int e = 0; e < N; ++e) r.F[e] -= t[e];
  blk_mm<N>(r.A, m.A, nA);
  blk_mm<N>(r.C, p.C, nC);
#pragma unroll
  for (int e = 0; e < N * N; ++e) { r.A[e] = -nA[e]; r.C[e] = -nC[e]; }
}

// Block parallel cyclic reduction over 128 rows held as (lo, hi) per lane;
// rows beyond the system are identity rows.  On return x = B^-1 F per row.
template <int N, class X>
__device__ __forceinline__ void blk_pcr_step(BlkRow<N> &lo, BlkRow<N> &hi, int lane, int d, const X &x) {
  const bool has_m = lane >= d, low_p = lane + d < 64;
  BlkNorm<N> nlo = blk_normalize<N>(lo), nhi = blk_normalize<N>(hi);
  BlkNorm<N> Lm = blk_xchg<N, true>(x, nlo), Hm = blk_xchg<N, true>(x, nhi);
  BlkNorm<N> Lp = X::same ? Lm : blk_xchg<N, false>(x, nlo), Hp = X::same ? Hm : blk_xchg<N, false>(x, nhi);
  const BlkNorm<N> Z = blk_zero<N>();
  blk_reduce<N>(lo, blk_sel<N>(has_m, Lm, Z), blk_sel<N>(low_p, Lp, Hp));
  blk_reduce<N>(hi, blk_sel<N>(has_m, Hm, Lm), blk_sel<N>(low_p, Hp, Z));
}
template <int N>
__device__ __forceinline__ void wave_block_pcr(BlkRow<N> &lo, BlkRow<N> &hi, int lane, double *xlo,
                                               double *xhi) {
  blk_pcr_step<N>(lo, hi, lane, 1, XchgRot{});
#pragma unroll 1
  for (int d = 2; d < 32; d <<= 1) blk_pcr_step<N>(lo, hi, lane, d, XchgPerm(lane, d));
  blk_pcr_step<N>(lo, hi, lane, 32, XchgXor32(lane));
  {  // d = 64
    BlkNorm<N> nlo = blk_normalize<N>(lo), nhi = blk_normalize<N>(hi);
    blk_reduce<N>(lo, blk_zero<N>(), nhi);
    blk_reduce<N>(hi, nlo, blk_zero<N>());
  }
  double Bi[N * N];
  blk_inv<N>(lo.B, Bi);
  blk_mv<N>(Bi, lo.F, xlo);
  blk_inv<N>(hi.B, Bi);
  blk_mv<N>(Bi, hi.F, xhi);
}

// The same for one row per lane (at most 64 rows): partners past either end
// of the wave are absent.
template <int N, class X>
__device__ __forceinline__ void blk_pcr1_step(BlkRow<N> &lo, int lane, int d, const X &x) {
  const bool has_m = lane >= d, has_p = lane + d < 64;
  BlkNorm<N> nlo = blk_normalize<N>(lo);
  BlkNorm<N> Lm = blk_xchg<N, true>(x, nlo), Lp = X::same ? Lm : blk_xchg<N, false>(x, nlo);
  const BlkNorm<N> Z = blk_zero<N>();
  blk_reduce<N>(lo, blk_sel<N>(has_m, Lm, Z), blk_sel<N>(has_p, Lp, Z));
}
template <int N>
__device__ __forceinline__ void wave_block_pcr1(BlkRow<N> &lo, int lane, double *xlo) {
  blk_pcr1_step<N>(lo, lane, 1, XchgRot{});
#pragma unroll 1
  for (int d = 2; d < 32; d <<= 1) blk_pcr1_step<N>(lo, lane, d, XchgPerm(lane, d));
  blk_pcr1_step<N>(lo, lane, 32, XchgXor32(lane));
  double Bi[N * N];
  blk_inv<N>(lo.B, Bi);
  blk_mv<N>(Bi, lo.F, xlo);
}

struct Nrm3 { double A[9], C[9], F[3]; };

// dst = (ok ? the source lane's (pick_a ? a : b) : 0), element by element
// (from the m partner if M, else from the p partner)
template <int K, bool M, class X>
__device__ __forceinline__ void fetchk(const X &x, const double *a, const double *b, bool pick_a, bool ok,
                                       double *dst) {
#pragma unroll
  for (int e = 0; e < K; ++e) {
    const double v = pick_a ? a[e] : b[e];
    const double t = M ? x.m(v) : x.p(v);
    dst[e] = ok ? t : 0.;
  }
}
// The reduced row of r from its neighbours m (partner lane m of x, row ma or
// mb) and p (partner lane p, row pa or pb), normalised again by its new
// diagonal block.
template <class X>
__device__ __forceinline__ void nrm3_update(const Nrm3 &r, const X &x, const Nrm3 &ma, const Nrm3 &mb,
                                            bool m_a, bool okm, const Nrm3 &pa, const Nrm3 &pb, bool p_a,
                                            bool okp, Nrm3 &o) {
  double Bp[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.}, Ap[9], Cp[9], Fp[3], T[9], t[3], nb[9], nf[3];
  // m side: B -= A Cm, A' = -A Am, F -= A Fm
  fetchk<9, true>(x, ma.C, mb.C, m_a, okm, nb);
  blk_mm<3>(r.A, nb, T);
#pragma unroll
  for (int e = 0; e < 9; ++e) Bp[e] -= T[e];
  fetchk<9, true>(x, ma.A, mb.A, m_a, okm, nb);
  blk_mm<3>(r.A, nb, T);
#pragma unroll
  for (int e = 0; e < 9; ++e) Ap[e] = -T[e];
  fetchk<3, true>(x, ma.F, mb.F, m_a, okm, nf);
  blk_mv<3>(r.A, nf, t);
#pragma unroll
  for (int e = 0; e < 3; ++e) Fp[e] = r.F[e] - t[e];
  // p side: B -= C Ap, C' = -C Cp, F -= C Fp
  fetchk<9, false>(x, pa.A, pb.A, p_a, okp, nb);
  blk_mm<3>(r.C, nb, T);
#pragma unroll
  for (int e = 0; e < 9; ++e) Bp[e] -= T[e];
  fetchk<9, false>(x, pa.C, pb.C, p_a, okp, nb);
  blk_mm<3>(r.C, nb, T);
#pragma unroll
  for (int e = 0; e < 9; ++e) Cp[e] = -T[e];
  fetchk<3, false>(x, pa.F, pb.F, p_a, okp, nf);
  blk_mv<3>(r.C, nf, t);
#pragma unroll
  for (int e = 0; e < 3; ++e) Fp[e] -= t[e];
  double Bi[9];
  blk_inv<3>(Bp, Bi);
  blk_mm<3>(Bi, Ap, o.A);
  blk_mm<3>(Bi, Cp, o.C);
  blk_mv<3>(Bi, Fp, o.F);
}
// 3x3 block cyclic reduction on rows kept normalised by their diagonal
// block (B = I): each lane holds A, C, F of its lo and hi rows.  A source
// lane offers the row its requester needs, so one neighbour block at a time
// is live.  On return lo.F / hi.F hold the solution.
template <class X>
__device__ __forceinline__ void pcr3_step(Nrm3 &lo, Nrm3 &hi, int lane, int d, const X &x) {
  bool lo_m_ok, hi_p_ok, offer_p_lo, offer_m_hi;
  if (d < 64) {
    lo_m_ok = lane >= d;       // lo's m-neighbour: the source's lo
    hi_p_ok = lane + d < 64;   // hi's p-neighbour: the source's hi
    offer_p_lo = lane >= d;    // lo's p-neighbour: source lo if source >= d, else its hi
    offer_m_hi = lane + d < 64;  // hi's m-neighbour: source hi if source + d < 64, else its lo
  } else {                     // the partner row sits in the same lane
    lo_m_ok = false;
    hi_p_ok = false;
    offer_p_lo = false;        // lo's p-neighbour: own hi
    offer_m_hi = false;        // hi's m-neighbour: own lo
  }
  Nrm3 nlo, nhi;
  nrm3_update(lo, x, lo, lo, true, lo_m_ok, lo, hi, offer_p_lo, true, nlo);
  nrm3_update(hi, x, hi, lo, offer_m_hi, true, hi, hi, true, hi_p_ok, nhi);
  lo = nlo;
  hi = nhi;
}
__device__ __forceinline__ void wave_pcr3(Nrm3 &lo, Nrm3 &hi, int lane) {
  pcr3_step(lo, hi, lane, 1, XchgRot{});
#pragma unroll 1
  for (int d = 2; d < 64; d <<= 1) pcr3_step(lo, hi, lane, d, XchgPerm(lane, d));
  pcr3_step(lo, hi, lane, 64, XchgOwn{});
}

// The same reduction for one row per lane (at most 64 rows): partners past
// either end of the wave are absent.
template <class X>
__device__ __forceinline__ void pcr3_1_step(Nrm3 &lo, int lane, int d, const X &x) {
  Nrm3 nlo;
  nrm3_update(lo, x, lo, lo, true, lane >= d, lo, lo, true, lane + d < 64, nlo);
  lo = nlo;
}
__device__ __forceinline__ void wave_pcr3_1(Nrm3 &lo, int lane) {
  pcr3_1_step(lo, lane, 1, XchgRot{});
#pragma unroll 1
  for (int d = 2; d < 64; d <<= 1) pcr3_1_step(lo, lane, d, XchgPerm(lane, d));
}

// Column index of this wave (consecutive waves take consecutive columns).
__device__ __forceinline__ bool wave_column_of(const Geo &G, int &i, int &j) {
  int nb = gridDim.x;
  int b = tgx_xcd_block(nb);
  long c = (long)b * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= (long)G.nlond * G.nlatd) return false;
  j = (int)(c / G.nlond);
  i = (int)(c - (long)j * G.nlond);
  return true;
}

// minor_ng.F90:60-237 for species SA and SB, one wavefront per column, the
// solve by parallel cyclic reduction over the nk rows.
// LIN: the NO and N4S sources formed here from the chemistry pass's linear
// coefficients and this step's N2D (a + N2D * b, as the Thomas NO/N4S kernel does).
template <int SA, int SB, int NH, bool LIN = false>
__global__ void __launch_bounds__(256) minor_wave_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                         tgx_ng_out out, MinorTerms MT) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk;
  const int Lr[2] = {lane, lane + 64};
  const double dzp = g.dz, dtx2inv = g.dtx2inv, small = 1.e-12, rdzp = 1. / g.dz;
  const size_t c2 = G.a2(i, j);
  const int SP[2] = {SA, SB};
  double *outs[2], *outs_nm[2];
  const double *fc[2], *fctm1[2], *lonb[2], *latb[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int sp = SP[q];
    outs[q] = sp == 0 ? out.no : sp == 1 ? out.n4s : sp == 2 ? out.n2d : out.ar;
    outs_nm[q] = sp == 0 ? out.no_nm : sp == 1 ? out.n4s_nm : sp == 2 ? out.n2d_nm : out.ar_nm;
    fc[q] = sp == 0 ? in.no : sp == 1 ? in.n4s : sp == 2 ? in.n2d : in.ar;
    fctm1[q] = sp == 0 ? in.no_nm : sp == 1 ? in.n4s_nm : sp == 2 ? in.n2d_nm : in.ar_nm;
    lonb[q] = sp == 0 ? in.lonb_no : sp == 1 ? in.lonb_n4s : sp == 2 ? in.lonb_n2d : in.lonb_ar;
    latb[q] = sp == 0 ? in.latb_no : sp == 1 ? in.latb_n4s : sp == 2 ? in.latb_n2d : in.latb_ar;
  }
  double v;
  if (lateral_value(G, in.lonb_no, in.latb_no, 0, i, j, v)) {
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int L = Lr[h];
      if (L >= nk) continue;
      size_t p = G.a3(L, i, j);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        lateral_value(G, lonb[q], latb[q], L, i, j, v);
        double nm = DTSMOOTH * fc[q][p] + DTSMOOTH_DIV2 * (fctm1[q][p] + v);
        outs[q][p] = v > small ? v : small;
        outs_nm[q][p] = nm > small ? nm : small;
      }
    }
    return;
  }
  const double pi = 4.0 * atan(1.0);
  double dfac;
  {  // minor_ng.F90:145-154
    const double hor = .25, dtr = pi / 180.;
    if (!(g.difhor > 0)) dfac = 1.;
    else {
      double rlat = g.glat[j] * dtr;
      if (fabs(rlat) >= pi / 4.5) dfac = hor + 1.;
      else dfac = hor + .5 * (1. + sin(pi * (fabs(rlat) - pi / 9.) / (pi / 4.5)));
    }
  }
  SpeciesConst sc[2] = {species_const(SA), species_const(SB)};
  const double *n2d_upd = out.n2d;
  // lower boundary of the major species (:66-69)
  const double o2_0 = in.o2[c2], o1_0 = in.o1[c2], he_0 = in.he[c2];
  const double bo2 = g.b[0] * o2_0 + g.b[1] * o1_0 + g.b[2] * he_0 + g.fb[0];
  const double bo1 = g.b[3] * o2_0 + g.b[4] * o1_0 + g.b[5] * he_0 + g.fb[1];
  const double bhe = g.b[6] * o2_0 + g.b[7] * o1_0 + g.b[8] * he_0 + g.fb[2];
  const double xmbari = 1. / (bo2 * RIO2 + bo1 * RIO1 + (1. - bo2 - bo1 - bhe) * RIN2 + bhe * RIHE);
  // lower-boundary data flbc(:,:,1:3) of each species (comp_*_ng.F90)
  double l1[2], l2[2], l3[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int s = SP[q];
    size_t p = c2;
    double xb = in.xnmbari[p];
    l1[q] = 0.;
    l2[q] = 1.;
    if (s == 0) l3[q] = -4.e6 * RMNO / xb;
    else if (s == 1)
      l3[q] = -(RMN4S / xb *
                (in.qtef[p] * (1. - BRN2D) / xb +
                 n2d_upd[p] * RIN2D * (BETA4 * xb * in.o1[p] * RIO1 + in.beta5[p] * in.ne[p] + BETA7) +
                 in.beta8[p] * in.no[p] * RINO) /
                (in.beta1[p] * in.o2[p] * RIO2 + in.beta3[p] * in.no[p] * RINO +
                 xb * in.beta17[p] * in.o1[p] * RIO1 * in.n2[p] * RIN2));
    else if (s == 2)
      l3[q] = -(RMN2D / xb * in.qtef[p] * BRN2D /
                (xb * (BETA2 * in.o2[p] * RIO2 + BETA4 * in.o1[p] * RIO1 + BETA6 * in.no[p] * RINO) + BETA7 +
                 in.beta5[p] * in.ne[p]));
    else l3[q] = -sqrt(g.ar_glbm1 * g.ar_glbm2);
  }
  const double fubc = 0.;  // every species sets fubc = 0 (comp_*_ng.F90)

  // own-level fields (levels beyond the column read the top level; their
  // rows become identity rows)
  struct MR { double o1, o2, he, mbar, tn, barm, w; };
  MR R[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    size_t p = G.a3(Lr[h] < nk ? Lr[h] : nk - 1, i, j);
    R[h].o1 = in.o1[p]; R[h].o2 = in.o2[p]; R[h].he = in.he[p]; R[h].mbar = in.mbar[p];
    R[h].tn = in.tn[p]; R[h].barm = in.barm[p]; R[h].w = in.w_upd[p];
  }
  MR A[2];  // level L-1
  wave_dn_h<NH>(R[0].o1, R[1].o1, lane, A[0].o1, A[1].o1);
  wave_dn_h<NH>(R[0].o2, R[1].o2, lane, A[0].o2, A[1].o2);
  wave_dn_h<NH>(R[0].he, R[1].he, lane, A[0].he, A[1].he);
  wave_dn_h<NH>(R[0].mbar, R[1].mbar, lane, A[0].mbar, A[1].mbar);
  wave_dn_h<NH>(R[0].tn, R[1].tn, lane, A[0].tn, A[1].tn);
  wave_dn_h<NH>(R[0].w, R[1].w, lane, A[0].w, A[1].w);
  // :66-126 interface quantities at level L
  MinorShared S[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    MinorShared &Q = S[h];
    if (L == 0) {
      Q.dmdz = (R[h].mbar - xmbari) * rdzp;
      Q.pso1 = .5 * (bo1 + o1_0);
      Q.pso2 = .5 * (bo2 + o2_0);
      Q.pshe = .5 * (bhe + he_0);
      Q.do1dz = (o1_0 - bo1) * rdzp;
      Q.do2dz = (o2_0 - bo2) * rdzp;
      Q.dhedz = (he_0 - bhe) * rdzp;
      Q.tni = in.tlbc[c2];
    } else {
      Q.dmdz = (R[h].mbar - A[h].mbar) * rdzp;
      Q.pso1 = .5 * (R[h].o1 + A[h].o1);
      Q.pso2 = .5 * (R[h].o2 + A[h].o2);
      Q.pshe = .5 * (R[h].he + A[h].he);
      Q.do1dz = (R[h].o1 - A[h].o1) * rdzp;
      Q.do2dz = (R[h].o2 - A[h].o2) * rdzp;
      Q.dhedz = (R[h].he - A[h].he) * rdzp;
      Q.tni = L == nk - 1 ? A[h].tn : .5 * (R[h].tn + A[h].tn);  // :89-93
    }
    Q.barm = R[h].barm;
    double df[3], al[9];
    const TPow tp = tpows(Q.tni);  // the temperature powers from one logarithm
    for (int n = 0; n < 3; ++n) df[n] = tp.df[n];
    Q.q = tp.q;
    al[0] = -(PHI(1, 4) + (PHI(1, 2) - PHI(1, 4)) * Q.pso1 + (df[0] * PHI(1, 3) - PHI(1, 4)) * Q.pshe);
    al[4] = -(PHI(2, 4) + (PHI(2, 1) - PHI(2, 4)) * Q.pso2 + (df[1] * PHI(2, 3) - PHI(2, 4)) * Q.pshe);
    al[8] = -(df[2] * PHI(3, 4) + (df[0] * PHI(3, 1) - df[2] * PHI(3, 4)) * Q.pso2 +
              (df[1] * PHI(3, 2) - df[2] * PHI(3, 4)) * Q.pso1);
    al[3] = (PHI(1, 2) - PHI(1, 4)) * Q.pso2;
    al[6] = (df[0] * PHI(1, 3) - PHI(1, 4)) * Q.pso2;
    al[1] = (PHI(2, 1) - PHI(2, 4)) * Q.pso1;
    al[7] = (df[1] * PHI(2, 3) - PHI(2, 4)) * Q.pso1;
    al[2] = (df[0] * PHI(3, 1) - df[2] * PHI(3, 4)) * Q.pshe;
    al[5] = (df[1] * PHI(3, 2) - df[2] * PHI(3, 4)) * Q.pshe;
    matinv3_r(al, Q.inv);
  }
  double tnim[2], tnip[2];
  wave_dn_h<NH>(S[0].tni, S[1].tni, lane, tnim[0], tnim[1]);
  wave_up_h<NH>(S[0].tni, S[1].tni, lane, tnip[0], tnip[1]);
  // species terms at level L
  double ax[2][2] = {}, thd[2][2] = {}, slo[2][2] = {}, s0p[2][2] = {}, rhs[2][2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    const int Lc = L < nk ? L : nk - 1;
    size_t p = G.a3(Lc, i, j);
    double dtnidz;
    if (L == 0) dtnidz = tnip[h] - S[h].tni;
    else if (L == nk - 1) dtnidz = S[h].tni - tnim[h];
    else dtnidz = (tnip[h] - tnim[h]) / 2.;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const SpeciesConst &cq = sc[q];
      const MinorShared &Q = S[h];
      double sai[3];
      for (int n = 0; n < 3; ++n) {
        double t = 0.;
        for (int m = 0; m < 3; ++m) t = t + cq.salfax[m] * Q.inv[m + 3 * n];
        sai[n] = t;
      }
      double ex = 1. - (cq.rmx + Q.dmdz) / Q.barm +
                  sai[0] * (Q.do2dz - (1. - (RMO2 + Q.dmdz) / Q.barm) * Q.pso2) +
                  sai[1] * (Q.do1dz - (1. - (RMO1 + Q.dmdz) / Q.barm) * Q.pso1) +
                  sai[2] * (Q.dhedz - (1. - (RMHE + Q.dmdz) / Q.barm) * Q.pshe);
      ax[h][q] = -(Q.barm / (TAU * RMN2) * Q.q /
                   (cq.phix[2] + cq.salfax[0] * Q.pso2 + cq.salfax[1] * Q.pso1 + cq.salfax[2] * Q.pshe));
      thd[h][q] = ex - cq.alfax * dtnidz / (dzp * Q.tni);
      rhs[h][q] = MT.rt[G.tile(Lc, i, j, SP[q], MR_NV)];
      const size_t ts = G.tile(Lc, i, j, 0, MS_NV);
      if (SP[q] == 3) { slo[h][q] = 0.; s0p[h][q] = 0.; }
      else if (LIN && SP[q] == 0) {
        const size_t tl = G.tile(Lc, i, j, 0, ML_NV);
        const double n2d = MT.n2d[p];
        slo[h][q] = MT.lt[tl + 128] + n2d * MT.lt[tl + 192];
        s0p[h][q] = MT.lt[tl + 256] + n2d * MT.lt[tl + 320];
      } else if (LIN && SP[q] == 1) {
        const size_t tl = G.tile(Lc, i, j, 0, ML_NV);
        slo[h][q] = MT.st[ts + ms_loss(1) * 64];
        s0p[h][q] = MT.lt[tl] + MT.n2d[p] * MT.lt[tl + 64];
      }
      else { slo[h][q] = MT.st[ts + ms_loss(SP[q]) * 64]; s0p[h][q] = MT.st[ts + ms_prod(SP[q]) * 64]; }
    }
  }
  // level L+1 terms from the next lane
  double axn[2][2], thdn[2][2], dmn[2], wn[2];
  {
    double dm0 = S[0].dmdz / S[0].barm, dm1 = NH == 2 ? S[1].dmdz / S[1].barm : 0.;
    wave_up_h<NH>(dm0, dm1, lane, dmn[0], dmn[1]);
    wave_up_h<NH>(R[0].w, R[1].w, lane, wn[0], wn[1]);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double u0, u1;
      wave_up_h<NH>(ax[0][q], ax[1][q], lane, u0, u1);
      axn[0][q] = u0; axn[1][q] = u1;
      wave_up_h<NH>(thd[0][q], thd[1][q], lane, u0, u1);
      thdn[0][q] = u0; thdn[1][q] = u1;
    }
  }
  // rows of the two systems (:160-198), identity beyond the column
  TriRow rlo[2] = {}, rhi[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    const double dmdzc = S[h].dmdz / S[h].barm;
    const int Lc = L < nk ? L : nk - 1;
    const double wic = L < nk - 1 ? .5 * (R[h].w + wn[h]) : 1.5 * R[h].w - .5 * A[h].w;
    const double difk0 = g.difk[Lc], difk1 = g.difk[L + 1 < nk ? L + 1 : nk - 1], expzk = g.expz[Lc];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double pc, qc, rc, fr;
      if (L <= nk - 2) {
        double molp = ax[h][q] * (rdzp + .5 * thd[h][q]);
        double molr = axn[h][q] * (rdzp - .5 * thdn[h][q]);
        double molq = ax[h][q] * (rdzp - .5 * thd[h][q]) + axn[h][q] * (rdzp + .5 * thdn[h][q]);
        double eddyp = g.expzmid_inv * difk0 * (rdzp - .5 * dmdzc);
        double eddyr = g.expzmid * difk1 * (rdzp + .5 * dmn[h]);
        double eddyq = g.expzmid_inv * difk0 * (rdzp + .5 * dmdzc) +
                       g.expzmid * difk1 * (rdzp - .5 * dmn[h]);
        pc = molp * rdzp - expzk * (eddyp * dfac + .5 * wic) * rdzp;
        rc = molr * rdzp - expzk * (eddyr * dfac - .5 * wic) * rdzp;
        qc = -molq * rdzp + expzk * (eddyq * dfac * rdzp + dtx2inv - slo[h][q]);
        fr = rhs[h][q] + s0p[h][q];
        fr = fr * expzk;
        if (L == 0) {
          qc = qc + pc * (l1[q] + .5 * l2[q] * dzp) / (l1[q] - .5 * l2[q] * dzp);
          fr = fr - pc * l3[q] * dzp / (l1[q] - .5 * l2[q] * dzp);
          pc = 0.;
        }
      } else {
        pc = 1. + .5 * dzp * thd[h][q];
        qc = pc - 2.;
        rc = 0.;
        fr = -GRAV * sc[q].rmx * fubc * dzp / (P0 * ax[h][q] * AVO);
      }
      TriRow t = L < nk ? TriRow{pc, qc, rc, fr} : tri_ident();
      if (h == 0) rlo[q] = t; else rhi[q] = t;
    }
  }
  if constexpr (NH == 2)
    wave_pcr<2>(rlo, rhi, lane);
  else
    wave_pcr1<2>(rlo, lane);
  // outputs (:234-237)
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L >= nk) continue;
    size_t p = G.a3(L, i, j);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const TriRow &t = h == 0 ? rlo[q] : rhi[q];
      double x = t.f / t.b;
      double nm = DTSMOOTH * fc[q][p] + DTSMOOTH_DIV2 * (fctm1[q][p] + x);
      outs[q][p] = x > small ? x : small;
      outs_nm[q][p] = nm > small ? nm : small;
    }
  }
}


// dt_ng.F90:54-213, one wavefront per column: rows 0..nk-2 solved by
// parallel cyclic reduction, the top level extrapolated (:167).
template <int NH>
__global__ void __launch_bounds__(256) dt_wave_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                      tgx_ng_out out,
                                                      const double *__restrict__ hadv_tn,
                                                      const double *__restrict__ smooth_tn_nm) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk;
  const int Lr[2] = {lane, lane + 64};
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
  const double t_lbc = in.t_lbc[c2], tl = in.tlbc[c2];
  if (lane == 0) {
    out.tlbc_nm[c2] = tl;
    out.tlbc[c2] = t_lbc;
  }
  double v;
  if (lateral_value(G, in.lonb_tn, in.latb_tn, 0, i, j, v)) {
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int L = Lr[h];
      if (L >= nk) continue;
      lateral_value(G, in.lonb_tn, in.latb_tn, L, i, j, v);
      size_t p = G.a3(L, i, j);
      out.tn_nm[p] = DTSMOOTH_DIV2 * (in.tn_nm[p] + v) + DTSMOOTH * in.tn[p];
      out.tn[p] = v > 100. ? v : 100.;
    }
    return;
  }
  struct DL { double tn, w, cp, km, un, vn, qtotal, schti, kt; };
  DL R[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    size_t p = G.a3(Lr[h] < nk ? Lr[h] : nk - 1, i, j);
    R[h].tn = in.tn[p]; R[h].w = in.w_upd[p]; R[h].cp = in.cp[p]; R[h].km = in.km[p];
    R[h].un = in.un[p]; R[h].vn = in.vn[p]; R[h].qtotal = in.qtotal[p]; R[h].schti = in.schti[p];
    R[h].kt = in.kt[p];
  }
  double tnm[2], tnp[2], unm[2], unp[2], vnm[2], vnp[2], wp[2], cpp[2], kmp[2], qtp[2];
  wave_dn_h<NH>(R[0].tn, R[1].tn, lane, tnm[0], tnm[1]);
  wave_up_h<NH>(R[0].tn, R[1].tn, lane, tnp[0], tnp[1]);
  wave_dn_h<NH>(R[0].un, R[1].un, lane, unm[0], unm[1]);
  wave_up_h<NH>(R[0].un, R[1].un, lane, unp[0], unp[1]);
  wave_dn_h<NH>(R[0].vn, R[1].vn, lane, vnm[0], vnm[1]);
  wave_up_h<NH>(R[0].vn, R[1].vn, lane, vnp[0], vnp[1]);
  wave_up_h<NH>(R[0].w, R[1].w, lane, wp[0], wp[1]);
  wave_up_h<NH>(R[0].cp, R[1].cp, lane, cpp[0], cpp[1]);
  wave_up_h<NH>(R[0].km, R[1].km, lane, kmp[0], kmp[1]);
  wave_up_h<NH>(R[0].qtotal, R[1].qtotal, lane, qtp[0], qtp[1]);
  // g(L), f(L) of dt_ng.F90:121-139 at this lane's levels
  double gg[2] = {}, ff[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h], Lc = L < nk ? L : nk - 1;
    double tni = L == 0 ? tl : .5 * (tnm[h] + R[h].tn);
    double hh = R[h].schti;
    double rho = P0 * g.expzmid_inv / (hh * GRAV);
    rho = rho * g.expz[Lc];
    double gpart = hh * hh * rho * R[h].cp;
    double gpart_a = gpart * g.dift[Lc];
    gg[h] = GRAV * (R[h].kt + gpart_a) / (P0 * hh * (dz * dz));
    double f = GRAV * GRAV * (hh * hh) * rho / (tni * P0 * 2. * dz);
    ff[h] = f * g.dift[Lc];
  }
  double ggp[2], ffp[2];
  wave_up_h<NH>(gg[0], gg[1], lane, ggp[0], ggp[1]);
  wave_up_h<NH>(ff[0], ff[1], lane, ffp[0], ffp[1]);
  // dudz, dvdz (:85-94); the top row takes a third of the row below's
  double dudz[2] = {}, dvdz[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L == 0) {
      dudz[h] = (R[h].un + 1. / 3. * unp[h] - 4. / 3. * in.ulbc[c2]) / dz;
      dvdz[h] = (R[h].vn + 1. / 3. * vnp[h] - 4. / 3. * in.vlbc[c2]) / dz;
    } else {
      dudz[h] = (unp[h] - unm[h]) / (2. * dz);
      dvdz[h] = (vnp[h] - vnm[h]) / (2. * dz);
    }
  }
  double dudzm[2], dvdzm[2];
  wave_dn_h<NH>(dudz[0], dudz[1], lane, dudzm[0], dudzm[1]);
  wave_dn_h<NH>(dvdz[0], dvdz[1], lane, dvdzm[0], dvdzm[1]);
  TriRow rows[2][1] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L > nk - 2) { rows[h][0] = tri_ident(); continue; }
    size_t p = G.a3(L, i, j);
    const double expzk = g.expz[L];
    double du = dudz[h], dv = dvdz[h];
    if (L == nk - 2) { du = dudzm[h] / 3.; dv = dvdzm[h] / 3.; }
    // veradv (:57-65)
    double raw_k = L == 0 ? (R[h].tn - t_lbc) * R[h].w * 2. : (R[h].tn - tnm[h]) * R[h].w;
    double raw_k1 = (L + 1 <= nk - 2) ? (tnp[h] - R[h].tn) * wp[h] : 0.;
    double veradv = (raw_k + raw_k1) * .5;
    veradv = veradv / dz;
    double cpi = .5 * (R[h].cp + cpp[h]);
    double kmi = .5 * (R[h].km + kmp[h]);
    double wi = .5 * (R[h].w + wp[h]);
    double advec_tn = hadv_tn[p] + veradv;
    double cptn = cpi * (advec_tn - dtx2inv * smooth_tn_nm[p]);
    cptn = cptn * expzk;
    double qsolar = .5 * (R[h].qtotal + qtp[h]);
    double t = in.xnmbar[p] * in.o1[p] * RIO1;
    const double mbar = in.mbar[p];
    double qrecomb = TSURPLUS * in.rkm12[p] * (t * t) * AVO / mbar;
    double qm = GRAV * kmi / (P0 * in.scht[p]) * (du * du + dv * dv);
    qm = qm / expzk;
    double total_heat = qsolar + in.hdt[p] + qrecomb + in.qji_tn[p] + qm;
    double total_heat_a = total_heat * expzk;
    cptn = cptn - total_heat_a;
    double pc = gg[h] - ff[h], qc, rc;
    if (L <= nk - 3) {
      qc = -(ggp[h] - ffp[h] + gg[h] + ff[h]);
      rc = ggp[h] + ffp[h];
    } else {
      qc = -gg[h] - ff[h];
      rc = 0.;
    }
    double rhs = cptn;
    double qpart = cpi * (dtx2inv + in.cool_imp[p]) + wi * GASK / mbar;
    rhs = rhs + in.cool_exp[p];
    double qpart_a = qpart * expzk;
    qc = qc - qpart_a;
    if (L == 0) {
      qc = qc - pc;
      rhs = rhs - 2. * pc * t_lbc;
      pc = 0.;
    }
    rows[h][0] = TriRow{pc, qc, rc, rhs};
  }
  TriRow lo[1] = {rows[0][0]}, hi[1] = {rows[1][0]};
  if constexpr (NH == 2)
    wave_pcr<1>(lo, hi, lane);
  else
    wave_pcr1<1>(lo, lane);
  double x[2] = {lo[0].f / lo[0].b, NH == 2 ? hi[0].f / hi[0].b : 0.}, xm1[2], xm2[2];
  wave_dn_h<NH>(x[0], x[1], lane, xm1[0], xm1[1]);
  wave_dn_h<NH>(xm1[0], xm1[1], lane, xm2[0], xm2[1]);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L >= nk) continue;
    double xv = L == nk - 1 ? xm1[h] * xm1[h] / xm2[h] : x[h];  // :167
    size_t p = G.a3(L, i, j);
    out.tn_nm[p] = DTSMOOTH_DIV2 * (in.tn_nm[p] + xv) + DTSMOOTH * in.tn[p];
    out.tn[p] = xv > 100. ? xv : 100.;
  }
}


// =============================================================================
// comp_o2o_ng.F90:44-81 (fused) + comp_ng.F90:71-457
struct CompTerms {
  // tiled comp inputs (CI_NV values): the prelude's smooth(*_nm)*dtx2inv -
  // (advec - fs(:,:,0)) of O2, O, He, then fs(:,1,1), fs(:,1,2), fs(:,2,1),
  // fs(:,2,2) of comp_o2o_ng
  double *ci;
  double *fs0[2];  // fs(:,1,0), fs(:,2,0) planes (read by the prelude)
};

// fs(k, isp, m) of comp_o2o_ng.F90:44-81 at the levels comp_ng uses (k <
// nk-1), nonzero planes only (the He row and column are zero), level-parallel.
__global__ void __launch_bounds__(256) o2o_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                  CompTerms CT) {
  int nb = gridDim.x;
  long t = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  long npt = (long)G.nlond * G.nlatd * (G.nk - 1);
  if (t >= npt) return;
  int i = (int)(t % G.nlond);
  long r = t / G.nlond;
  int j = (int)(r % G.nlatd);
  int k = (int)(r / G.nlatd);
  const size_t p = G.a3(k, i, j);
  double rji = 0.5 * (in.rj[p] + in.rj[p + G.P]);
  double qo2pi = 0.5 * (in.qo2p[p] + in.qo2p[p + G.P]);
  double qopi = 0.5 * (in.qop[p] + in.qop[p + G.P]);
  double beta8i = 0.5 * (in.beta8[p] + in.beta8[p + G.P]);
  double nei = sqrt(in.ne[p] * in.ne[p + G.P]);
  double xn = in.xnmbar[p], no = in.no[p], n4s = in.n4s[p], n2d = in.n2d[p], o2p = in.o2p[p];
  double op = in.op[p], nplus = in.nplus[p];
  double pox1 = xn * xn * (in.beta3[p] * n4s / RMN4S * no / RMNO + BETA6 * n2d / RMN2D * no / RMNO) +
                beta8i * no / RMNO * xn + xn * (RK4 * o2p * n4s / RMN4S + RK10 * op * n2d / RMN2D) +
                (in.ra1[p] * in.nop[p] + 2. * in.ra2[p] * o2p) * nei;
  double pox2 = xn * (in.beta1[p] * n4s / RMN4S + BETA2 * n2d / RMN2D) + in.rk1[p] * op + RK7 * nplus + 2. * rji;
  double lox1 = 2. * in.rkm12[p] * xn / in.mbar[p];
  double lox2 = in.rk3[p] * in.n2p[p] + RK8 * nplus;
  double lox3 = qopi;
  double po21 = in.rkm12[p] * xn / in.mbar[p];
  double po23 = RK5 * no / RMNO * o2p * xn;
  double lo21 = xn * (in.beta1[p] * n4s / RMN4S + BETA2 * n2d / RMN2D) + in.rk1[p] * op + (RK6 + RK7) * nplus +
                RK9 * in.n2p[p] + rji;
  double lo22 = qo2pi;
  double o1 = in.o1[p];
  CT.ci[G.tile(k, i, j, 3, CI_NV)] = -lo21;                                  // fs(k,1,1)
  CT.ci[G.tile(k, i, j, 4, CI_NV)] = xn * po21 * o1 * (RIO1 * RIO1) * RMO2;  // fs(k,1,2)
  CT.ci[G.tile(k, i, j, 5, CI_NV)] = pox2 * RMO1 * RIO2;                     // fs(k,2,1)
  CT.ci[G.tile(k, i, j, 6, CI_NV)] = -lox2 - lox1 * o1 * RIO1 * xn;          // fs(k,2,2)
  CT.fs0[0][p] = (po23 - lo22) * RMO2 / xn;                                  // fs(k,1,0)
  CT.fs0[1][p] = (pox1 - lox3) * RMO1 / xn;                                  // fs(k,2,0)
}

// comp_ng.F90 on a lateral-boundary column: the Dirichlet values of O2, O
// and He (:391-413), then the post pass :415-457 on them.
__device__ void comp_lateral_column(const tgx_ng_grid &g, const Geo &G, const tgx_ng_fields &in,
                                    const tgx_ng_out &out, int i, int j) {
  const int nk = G.nk;
  const double small = 1.e-9, rdz = 1. / g.dz;
  double v;
  for (int k = 0; k < nk; ++k) {
    size_t p = G.a3(k, i, j);
    lateral_value(G, in.lonb_o2, in.latb_o2, k, i, j, v); out.o2[p] = v;
    lateral_value(G, in.lonb_o1, in.latb_o1, k, i, j, v); out.o1[p] = v;
    lateral_value(G, in.lonb_he, in.latb_he, k, i, j, v); out.he[p] = v;
  }
  bool found = false, extrap = false;
  double n2top = 0., dn2dz = 0., zmtop = 0., n2prev = 0., xnprev = 0.;
  for (int k = 0; k < nk; ++k) {
    size_t p = G.a3(k, i, j);
    const double o2u = out.o2[p], o1u = out.o1[p], heu = out.he[p], xn = in.xnmbar[p];
    double o2nm = DTSMOOTH_DIV2 * (in.o2_nm[p] + o2u) + DTSMOOTH * in.o2[p];
    double o1nm = DTSMOOTH_DIV2 * (in.o1_nm[p] + o1u) + DTSMOOTH * in.o1[p];
    double henm = DTSMOOTH_DIV2 * (in.he_nm[p] + heu) + DTSMOOTH * in.he[p];
    double n2 = 1. - o2u - o1u - heu;
    double n2nm = 1. - o2nm - o1nm - henm;
    if (!found) {
      if (n2 < 0.01) {
        found = true;
        if (k + 1 < nk && k >= 1) {
          extrap = true;
          zmtop = g.zpmid[k];
          n2top = log(n2 * xn);
          double n2top1 = log(n2prev * xnprev);
          dn2dz = (n2top - n2top1) * rdz;
        }
      }
    } else if (extrap) {
      n2 = exp(n2top + dn2dz * (g.zpmid[k] - zmtop)) / xn;
    }
    n2prev = n2;
    xnprev = xn;
    double a = o2u > small ? o2u : small, b = o1u > small ? o1u : small;
    double c = heu > small ? heu : small, d = n2 > small ? n2 : small;
    double norm = a + b + c + d;
    out.o2[p] = a / norm;
    out.o1[p] = b / norm;
    out.he[p] = c / norm;
    double an = o2nm > small ? o2nm : small, bn = o1nm > small ? o1nm : small;
    double cn = henm > small ? henm : small, dn = n2nm > small ? n2nm : small;
    double normn = an + bn + cn + dn;
    out.o2_nm[p] = an / normn;
    out.o1_nm[p] = bn / normn;
    out.he_nm[p] = cn / normn;
  }
}

template <bool CAP>
__global__ void __launch_bounds__(256) comp_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                   tgx_ng_out out, double *__restrict__ gs,
                                                   CompTerms CT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv, thdiffalpha = -0.38, small = 1.e-9, rdz = 1. / g.dz;
  const size_t c2 = G.a2(i, j);
  const double fhe = g.calc_helium == 0 ? 0. : in.flx_he[c2];
  out.flx_he[c2] = fhe;
#define F3(f, k) (f)[G.a3((k), i, j)]
  double v;
  const bool bnd = lateral_value(G, in.lonb_o2, in.latb_o2, 0, i, j, v);
  if (bnd) {
    comp_lateral_column(g, G, in, out, i, j);
    return;
  }
  {
    const double dfac = dfactor_at(g, j), racs = g.racs[j];
    const double tl = in.t_lbc[c2];  // flds%tlbc after dt_ng of this step
    const double rmass[3] = {RMO2, RMO1, RMHE};
    const double o2_0 = F3(in.o2, 0), o1_0 = F3(in.o1, 0), he_0 = F3(in.he, 0);
    double ps0[3];
    for (int l = 0; l < 3; ++l)
      ps0[l] = g.b[3 * l + 0] * o2_0 + g.b[3 * l + 1] * o1_0 + g.b[3 * l + 2] * he_0 + g.fb[l];
    const double embar0 =
        1. / (ps0[0] * RIO2 + ps0[1] * RIO1 + ps0[2] * RIHE + (1. - ps0[0] - ps0[1] - ps0[2]) * RIN2);
    struct Lev { double o2i, o1i, hei, embari, tni, dembardz, dtndz, df[3], q; };
    // Fields read at two neighbouring levels (interface means); one level of
    // them is carried in registers.
    struct LR { double o2, o1, he, mbar, tn, w; };
    // Fields and level constants read at the level being eliminated only.
    struct LS {
      double fs[2][3];
      double rhs[3], hdo2, hdo1, hdhe, expz, difk0, difk1;
    };
    auto load_lr = [&](int L, LR &R) {
      size_t p = G.a3(L, i, j);
      R.o2 = in.o2[p]; R.o1 = in.o1[p]; R.he = in.he[p]; R.mbar = in.mbar[p]; R.tn = in.tn[p];
      R.w = in.w[p];
    };
    auto load_ls = [&](int L, LS &S) {
      size_t p = G.a3(L, i, j);
      const size_t t = G.tile(L, i, j, 0, CI_NV);
      S.rhs[0] = CT.ci[t]; S.rhs[1] = CT.ci[t + 64]; S.rhs[2] = CT.ci[t + 128];
      S.fs[0][1] = CT.ci[t + 192]; S.fs[0][2] = CT.ci[t + 256]; S.fs[1][1] = CT.ci[t + 320];
      S.fs[1][2] = CT.ci[t + 384];
      S.hdo2 = in.hdo2[p]; S.hdo1 = in.hdo1[p]; S.hdhe = in.hdhe[p];
      S.expz = g.expz[L]; S.difk0 = g.difk[L]; S.difk1 = g.difk[L + 1];
    };
    auto lev_df = [&](Lev &Q) {  // comp_ng.F90:124-126
      Q.dembardz = Q.dembardz * rdz;
      Q.dtndz = Q.dtndz * rdz;
      const TPow tp = tpows(Q.tni);
      for (int n = 0; n < 3; ++n) Q.df[n] = tp.df[n];
      Q.q = tp.q;
    };
    auto lev0 = [&](const LR &R, Lev &Q) {  // comp_ng.F90:80-97 at the bottom interface
      Q.o2i = .5 * (ps0[0] + o2_0);
      Q.o1i = .5 * (ps0[1] + o1_0);
      Q.hei = .5 * (ps0[2] + he_0);
      Q.embari = .5 * (embar0 + R.mbar);
      Q.tni = tl;
      Q.dembardz = R.mbar - embar0;
      Q.dtndz = (R.tn - tl) * 2;
      lev_df(Q);
    };
    // interface between A (L-1) and B (L); the global comp.F takes the top
    // interface's temperature from the level below it (:174-175, :446-449)
    auto levk = [&](const LR &A, const LR &B, Lev &Q, bool top) {
      Q.o2i = .5 * (A.o2 + B.o2);
      Q.o1i = .5 * (A.o1 + B.o1);
      Q.hei = .5 * (A.he + B.he);
      Q.embari = .5 * (A.mbar + B.mbar);
      Q.tni = (G.glb && top) ? A.tn : .5 * (A.tn + B.tn);
      Q.dembardz = B.mbar - A.mbar;
      Q.dtndz = B.tn - A.tn;
      lev_df(Q);
    };
    auto alpha_of = [&](const Lev &Q, double *al) {  // :139-151
      const double *d_ = Q.df;
      al[0] = -(PHI(1, 4) + (PHI(1, 2) - PHI(1, 4)) * Q.o1i + (d_[0] * PHI(1, 3) - PHI(1, 4)) * Q.hei);
      al[4] = -(PHI(2, 4) + (PHI(2, 1) - PHI(2, 4)) * Q.o2i + (d_[1] * PHI(2, 3) - PHI(2, 4)) * Q.hei);
      al[8] = -(d_[2] * PHI(3, 4) + (d_[0] * PHI(3, 1) - d_[2] * PHI(3, 4)) * Q.o2i +
                (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * Q.o1i);
      al[3] = (PHI(1, 2) - PHI(1, 4)) * Q.o2i;
      al[6] = (d_[0] * PHI(1, 3) - PHI(1, 4)) * Q.o2i;
      al[1] = (PHI(2, 1) - PHI(2, 4)) * Q.o1i;
      al[7] = (d_[1] * PHI(2, 3) - PHI(2, 4)) * Q.o1i;
      al[2] = (d_[0] * PHI(3, 1) - d_[2] * PHI(3, 4)) * Q.hei;
      al[5] = (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * Q.hei;
    };
    LR r1, r2;
    load_lr(0, r1);
    Lev Qc, Qn;
    lev0(r1, Qc);
    double epm[3], epp[3], akm[9], akp[9], alpha[9];  // levels k (km) and k+1 (kp)
    double wks3, wks4 = Qc.dembardz / (Qc.embari * 2.);
    for (int l = 0; l < 3; ++l) epp[l] = 1. - 1. / Qc.embari * (rmass[l] + Qc.dembardz);
    epp[2] = epp[2] - thdiffalpha * Qc.dtndz / Qc.tni;
    alpha_of(Qc, alpha);
    matinv3_r(alpha, akp);
    double wks1 = Qc.embari * RIN2 * Qc.q * (1. / TAU);
    for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
    double gama[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, zz[3] = {0, 0, 0};
    double alpha22 = 0, alpha33 = 0, alpha23 = 0, alpha32 = 0, flx00 = 0;
    // The elimination's gamma (9) and forward solution z (3) of level k, in
    // a tiled region of 12 values.
    auto tile = [&](int k, int n) { return G.tile(k, i, j, n, 12); };
    // The loads of level k+1 are issued before the arithmetic of level k, so
    // the memory round trip overlaps the elimination.
    LR r2n;
    LS lsn;
    load_lr(1, r2n);
    load_ls(0, lsn);
    for (int k = 0; k < nk - 1; ++k) {
      const LS ls = lsn;
      r2 = r2n;
      if (k + 1 < nk - 1) {
        load_lr(k + 2, r2n);
        load_ls(k + 1, lsn);
      }
      asm volatile("" ::: "memory");
      for (int n = 0; n < 9; ++n) akm[n] = akp[n];
      for (int n = 0; n < 3; ++n) epm[n] = epp[n];
      levk(r1, r2, Qn, k == nk - 2);
      const double rem = 1. / Qn.embari;
      for (int l = 0; l < 3; ++l) epp[l] = 1. - rem * (rmass[l] + Qn.dembardz);
      if (k != nk - 2) epp[2] = epp[2] - thdiffalpha * Qn.dtndz / Qn.tni;
      alpha_of(Qn, alpha);
      matinv3_r(alpha, akp);
      wks1 = Qn.embari * RIN2 * Qn.q * (1. / TAU);
      wks3 = wks4;
      wks4 = Qn.dembardz / (Qn.embari * 2.);
      for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
      double molp[9], molr[9], molq[9], pk[9], rk[9], qk[9], fk[3];
      for (int isp = 0; isp < 3; ++isp)
        for (int m = 0; m < 3; ++m) {
          int e = isp + 3 * m;
          molp[e] = akm[e] * (rdz + epm[m] / 2.);
          molr[e] = akp[e] * (rdz - epp[m] / 2.);
          molq[e] = akm[e] * (rdz - epm[m] / 2.) + akp[e] * (rdz + epp[m] / 2.);
        }
      double eddyp = g.expzmid_inv * ls.difk0 * (rdz - wks3);
      double eddyr = g.expzmid * ls.difk1 * (rdz + wks4);
      double eddyq = g.expzmid_inv * ls.difk0 * (rdz + wks3) + g.expzmid * ls.difk1 * (rdz - wks4);
      const double(&fs)[2][3] = ls.fs;
      const double expzk = ls.expz;
      const double wik = .5 * (r1.w + r2.w);
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) {
          int e = isp + 3 * m;
          double delta = isp == m ? 1. : 0.;
          double fsv = (isp < 2 && m < 2) ? fs[isp][m + 1] : 0.;
          pk[e] = molp[e] * rdz - expzk * delta * rdz * (eddyp * dfac + .5 * wik);
          rk[e] = molr[e] * rdz - expzk * delta * rdz * (eddyr * dfac - .5 * wik);
          qk[e] = -molq[e] * rdz + expzk * delta * (eddyq * dfac * rdz + dtx2inv) - expzk * fsv;
        }
      fk[0] = expzk * (ls.rhs[0] + ls.hdo2);
      fk[1] = expzk * (ls.rhs[1] + ls.hdo1);
      fk[2] = expzk * (ls.rhs[2] + ls.hdhe);
      if (k == 0) {
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n)
            for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] + pk[isp + 3 * n] * g.b[3 * n + m];
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp) fk[isp] = fk[isp] - pk[isp + 3 * m] * g.fb[m];
        for (int n = 0; n < 9; ++n) pk[n] = 0.;
      } else if (k == nk - 2) {
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp)
            qk[isp + 3 * m] = qk[isp + 3 * m] + (1. + .5 * epp[m] * dz) / (1. - .5 * epp[m] * dz) * rk[isp + 3 * m];
        const double *d_ = Qn.df;
        alpha22 = -(PHI(2, 4) + (PHI(2, 1) - PHI(2, 4)) * Qn.o2i + (d_[1] * PHI(2, 3) - PHI(2, 4)) * Qn.hei);
        alpha33 = -(d_[2] * PHI(3, 4) + (d_[0] * PHI(3, 1) - d_[2] * PHI(3, 4)) * Qn.o2i +
                    (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * Qn.o1i);
        alpha23 = (d_[1] * PHI(2, 3) - PHI(2, 4)) * Qn.o1i;
        alpha32 = (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * Qn.hei;
        // the mean mass of the interface below (comp_ng.F90:272) or of the top
        // interface (the global comp.F:632-633)
        flx00 = (G.glb ? Qn.embari : Qc.embari) * RIN2 * P0 * pow(T00 / r1.tn, 0.25) / (TAU * GRAV);
        for (int isp = 0; isp < 3; ++isp)
          fk[isp] = fk[isp] - rk[isp + 3 * 1] * (alpha23 - alpha22) * fhe / (flx00 * (rdz - 0.5 * epp[1])) -
                    rk[isp + 3 * 2] * (alpha33 - alpha32) * fhe / (flx00 * (rdz - 0.5 * epp[2]));
        for (int n = 0; n < 9; ++n) rk[n] = 0.;
      }
      if (CAP) {
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) {
            cap_put(g, G, CAP_CO + 0 + 3 * r + c, i, j, k, pk[r + 3 * c]);
            cap_put(g, G, CAP_CO + 9 + 3 * r + c, i, j, k, qk[r + 3 * c]);
            cap_put(g, G, CAP_CO + 18 + 3 * r + c, i, j, k, rk[r + 3 * c]);
          }
          cap_put(g, G, CAP_CO + 27 + r, i, j, k, fk[r]);
        }
      }
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] - pk[isp + 3 * n] * gama[n + 3 * m];
      double wkm1[9], wkv1[3], ng[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, nz[3] = {0, 0, 0};
      matinv3_r(qk, wkm1);
      for (int isp = 0; isp < 3; ++isp) wkv1[isp] = fk[isp];
      for (int m = 0; m < 3; ++m) {
        for (int isp = 0; isp < 3; ++isp) wkv1[isp] = wkv1[isp] - pk[isp + 3 * m] * zz[m];
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp) ng[isp + 3 * m] = ng[isp + 3 * m] + wkm1[isp + 3 * n] * rk[n + 3 * m];
      }
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) nz[isp] = nz[isp] + wkm1[isp + 3 * m] * wkv1[m];
      for (int n = 0; n < 9; ++n) gama[n] = ng[n];
      for (int n = 0; n < 9; ++n) gs[tile(k, n)] = ng[n];
      for (int isp = 0; isp < 3; ++isp) zz[isp] = nz[isp];
      for (int isp = 0; isp < 3; ++isp) gs[tile(k, 9 + isp)] = nz[isp];
      Qc = Qn;
      r1 = r2;
    }
    // :317-332 back substitution (top down), the closed-form top level and,
    // as each level's solution is known, the post pass :386-457 (time
    // smoothing, N2 above the O-dominated top, clamps, normalisation).  The
    // lowest level with N2 < 0.01 is known only at the bottom; the levels
    // above it are then normalised again with the extrapolated N2, their
    // solution recomputed by the same back substitution.
    struct BS { double z[3], gm[9], o2nm, o1nm, henm, o2, o1, he, xn; };
    auto load_bs = [&](int k, BS &B) {
      size_t p = G.a3(k, i, j);
      for (int isp = 0; isp < 3; ++isp) B.z[isp] = gs[tile(k, 9 + isp)];
      for (int n = 0; n < 9; ++n) B.gm[n] = gs[tile(k, n)];
      B.o2nm = in.o2_nm[p]; B.o1nm = in.o1_nm[p]; B.henm = in.he_nm[p];
      B.o2 = in.o2[p]; B.o1 = in.o1[p]; B.he = in.he[p];
      B.xn = in.xnmbar[p];
    };
    auto back = [&](const BS &B, const double *u1, double *u) {
      for (int isp = 0; isp < 3; ++isp) {
        double uu = B.z[isp];
        for (int m = 0; m < 3; ++m) uu = uu - B.gm[isp + 3 * m] * u1[m];
        u[isp] = uu;
      }
    };
    auto top_of = [&](const double *u, double *t) {  // :317-332 from the solution at nk-2
      for (int isp = 0; isp < 3; ++isp) t[isp] = (1. + .5 * epp[isp] * dz) / (1. - .5 * epp[isp] * dz) * u[isp];
      t[1] = t[1] + (alpha23 - alpha22) * fhe / (flx00 * (rdz - 0.5 * epp[1]));
      t[2] = t[2] + (alpha33 - alpha32) * fhe / (flx00 * (rdz - 0.5 * epp[2]));
    };
    auto put_norm = [&](size_t p, double o2u, double o1u, double heu, double n2) {
      double a = o2u > small ? o2u : small, b = o1u > small ? o1u : small;
      double c = heu > small ? heu : small, d = n2 > small ? n2 : small;
      double norm = a + b + c + d;
      out.o2[p] = a / norm;
      out.o1[p] = b / norm;
      out.he[p] = c / norm;
    };
    int kf = -1;  // lowest level so far with N2 < 0.01
    double n2kf = 0., xnkf = 0., n2b = 0., xnb = 0.;
    auto post = [&](int k, const double *u, const BS &B) {
      size_t p = G.a3(k, i, j);
      const double o2u = u[0], o1u = u[1], heu = g.calc_helium == 0 ? g.pshelb : u[2];
      if (G.raw_row(j)) {  // the raw solution; filter_o2o and the post pass follow (comp.F:925-1022)
        out.o2[p] = o2u;
        out.o1[p] = o1u;
        out.he[p] = heu;
        return;
      }
      double o2nm = DTSMOOTH_DIV2 * (B.o2nm + o2u) + DTSMOOTH * B.o2;
      double o1nm = DTSMOOTH_DIV2 * (B.o1nm + o1u) + DTSMOOTH * B.o1;
      double henm = DTSMOOTH_DIV2 * (B.henm + heu) + DTSMOOTH * B.he;
      double n2 = 1. - o2u - o1u - heu;
      double n2nm = 1. - o2nm - o1nm - henm;
      const bool lo = n2 < 0.01, below = !lo && kf == k + 1;  // (selects keep these in registers)
      n2kf = lo ? n2 : n2kf;
      xnkf = lo ? B.xn : xnkf;
      kf = lo ? k : kf;
      n2b = below ? n2 : n2b;
      xnb = below ? B.xn : xnb;
      put_norm(p, o2u, o1u, heu, n2);
      double an = o2nm > small ? o2nm : small, bn = o1nm > small ? o1nm : small;
      double cn = henm > small ? henm : small, dn = n2nm > small ? n2nm : small;
      double normn = an + bn + cn + dn;
      out.o2_nm[p] = an / normn;
      out.o1_nm[p] = bn / normn;
      out.he_nm[p] = cn / normn;
    };
    {
      BS bc, bn, bt;
      load_bs(nk - 2, bc);
      load_bs(nk - 1, bt);
      load_bs(nk - 3, bn);
      asm volatile("" ::: "memory");
      const double zero[3] = {0., 0., 0.};
      double u1[3], u[3];
      back(bc, zero, u1);  // level nk-2
      double ut[3];
      top_of(u1, ut);
      post(nk - 1, ut, bt);
      post(nk - 2, u1, bc);
      // Levels nk-3 .. 0, each level's loads issued three levels ahead: four
      // register slots rotate through a loop unrolled by four, so no slot
      // is indexed at run time.
      BS s0 = bn, s1, s2, s3;
      int k = nk - 3;
      if (k - 1 >= 0) load_bs(k - 1, s1);
      if (k - 2 >= 0) load_bs(k - 2, s2);
      auto it = [&](int kk, const BS &cur, BS &pf) {
        if (kk - 3 >= 0) load_bs(kk - 3, pf);
        asm volatile("" ::: "memory");
        back(cur, u1, u);
        post(kk, u, cur);
        for (int isp = 0; isp < 3; ++isp) u1[isp] = u[isp];
      };
      for (; k >= 3; k -= 4) {
        it(k, s0, s3);
        it(k - 1, s1, s0);
        it(k - 2, s2, s1);
        it(k - 3, s3, s2);
      }
      if (k >= 0) it(k, s0, s3);
      if (k - 1 >= 0) it(k - 1, s1, s0);
      if (k - 2 >= 0) it(k - 2, s2, s1);
    }
    if (!G.raw_row(j) && kf >= 1 && kf + 1 < nk) {  // :419-431 extrapolated N2 above kf
      const double zmtop = g.zpmid[kf];
      const double n2top = log(n2kf * xnkf);
      const double n2top1 = log(n2b * xnb);
      const double dn2dz = (n2top - n2top1) * rdz;
      BS bc;
      const double zero[3] = {0., 0., 0.};
      double u1[3], u[3];
      load_bs(nk - 2, bc);
      back(bc, zero, u1);
      double t[3];
      top_of(u1, t);
      for (int k = nk - 1; k > kf; --k) {
        if (k < nk - 2) {
          load_bs(k, bc);
          back(bc, u1, u);
          for (int isp = 0; isp < 3; ++isp) u1[isp] = u[isp];
        }
        double uk[3];
        for (int isp = 0; isp < 3; ++isp) uk[isp] = k == nk - 1 ? t[isp] : u1[isp];
        const size_t p = G.a3(k, i, j);
        const double xn = in.xnmbar[p];
        const double n2 = exp(n2top + dn2dz * (g.zpmid[k] - zmtop)) / xn;
        put_norm(p, uk[0], uk[1], g.calc_helium == 0 ? g.pshelb : uk[2], n2);
      }
    }
    return;
  }
#undef F3
}

// =============================================================================
// The precision form of comp (comp_ng.F90:71-457): the block rows assembled
// and eliminated in FP32, the 3x3 block pivots inverted in FP64, then
// iterative refinement: each sweep forms the residual r = f - A u with the
// rows assembled again in FP64 and solves A d = r with the FP32 elimination
// of those rows, u += d.  The last sweep applies the post pass.  u (nk levels
// of O2, O, He before the post pass, the top level from the closed form) is
// kept in a tiled region (CM_NV values), the FP32 gamma and forward solution
// in the elimination's tiled region (12 floats where the FP64 kernel keeps 12
// doubles).

template <class T> __device__ __forceinline__ T t_log(T x);
template <> __device__ __forceinline__ float t_log<float>(float x) { return logf(x); }
template <> __device__ __forceinline__ double t_log<double>(double x) { return log(x); }
template <class T> __device__ __forceinline__ T t_exp(T x);
template <> __device__ __forceinline__ float t_exp<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ double t_exp<double>(double x) { return exp(x); }

template <class T>
__device__ __forceinline__ void matinv3_t(const T *A, T *Bm) {
#define M3(r, c) A[((r) - 1) + 3 * ((c) - 1)]
  T d = M3(1,1)*M3(2,2)*M3(3,3) - M3(1,1)*M3(2,3)*M3(3,2)
      - M3(1,2)*M3(2,1)*M3(3,3) + M3(1,2)*M3(2,3)*M3(3,1)
      + M3(1,3)*M3(2,1)*M3(3,2) - M3(1,3)*M3(2,2)*M3(3,1);
  const T r = T(1) / d;
  Bm[0] =  (M3(2,2)*M3(3,3) - M3(2,3)*M3(3,2)) * r;
  Bm[1] = -(M3(2,1)*M3(3,3) - M3(2,3)*M3(3,1)) * r;
  Bm[2] =  (M3(2,1)*M3(3,2) - M3(2,2)*M3(3,1)) * r;
  Bm[3] = -(M3(1,2)*M3(3,3) - M3(1,3)*M3(3,2)) * r;
  Bm[4] =  (M3(1,1)*M3(3,3) - M3(1,3)*M3(3,1)) * r;
  Bm[5] = -(M3(1,1)*M3(3,2) - M3(1,2)*M3(3,1)) * r;
  Bm[6] =  (M3(1,2)*M3(2,3) - M3(1,3)*M3(2,2)) * r;
  Bm[7] = -(M3(1,1)*M3(2,3) - M3(1,3)*M3(2,1)) * r;
  Bm[8] =  (M3(1,1)*M3(2,2) - M3(1,2)*M3(2,1)) * r;
#undef M3
}

constexpr int CM_NV = 3;  // u: O2, O, He

// The block rows of comp_ng.F90:168-326 for one column, level by level
// (k = 0 .. nk-2 in order), in arithmetic type T: the same terms as
// comp_kernel's forward sweep.
template <class T>
struct CompRows {
  const tgx_ng_grid &g;
  const Geo &G;
  const tgx_ng_fields &in;
  const CompTerms &CT;
  int i, j;
  T rdz, dz, dtx2inv, dfac, fhe, tl, thdiffalpha;
  T ps0[3], embar0, o2_0, o1_0, he_0;
  T r1[6];  // o2, o1, he, mbar, tn, w at the level being eliminated
  T epp[3], akp[9], wks4, embari_c;
  T alpha22 = 0, alpha33 = 0, alpha23 = 0, alpha32 = 0, flx00 = 0;
  struct Lev { T o2i, o1i, hei, embari, tni, dembardz, dtndz, df[3], q; };

  __device__ CompRows(const tgx_ng_grid &g_, const Geo &G_, const tgx_ng_fields &in_, const CompTerms &CT_,
                      int i_, int j_)
      : g(g_), G(G_), in(in_), CT(CT_), i(i_), j(j_) {}
  __device__ T rmass(int l) const { return T(l == 0 ? RMO2 : (l == 1 ? RMO1 : RMHE)); }
  __device__ void load_r(int L, T *r) const {
    const size_t p = G.a3(L, i, j);
    r[0] = T(in.o2[p]); r[1] = T(in.o1[p]); r[2] = T(in.he[p]);
    r[3] = T(in.mbar[p]); r[4] = T(in.tn[p]); r[5] = T(in.w[p]);
  }
  __device__ void lev_df(Lev &Q) const {
    Q.dembardz = Q.dembardz * rdz;
    Q.dtndz = Q.dtndz * rdz;
    const T L = t_log<T>(Q.tni / T(T00));
    for (int n = 0; n < 3; ++n) Q.df[n] = t_exp<T>(T(1.75 - ssn(n)) * L);
    Q.q = t_exp<T>(T(-0.25) * L);
  }
  __device__ void alpha_of(const Lev &Q, T *al) const {
    const T *d_ = Q.df;
#define PH(r, c) T(PHI(r, c))
    al[0] = -(PH(1, 4) + (PH(1, 2) - PH(1, 4)) * Q.o1i + (d_[0] * PH(1, 3) - PH(1, 4)) * Q.hei);
    al[4] = -(PH(2, 4) + (PH(2, 1) - PH(2, 4)) * Q.o2i + (d_[1] * PH(2, 3) - PH(2, 4)) * Q.hei);
    al[8] = -(d_[2] * PH(3, 4) + (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * Q.o2i +
              (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * Q.o1i);
    al[3] = (PH(1, 2) - PH(1, 4)) * Q.o2i;
    al[6] = (d_[0] * PH(1, 3) - PH(1, 4)) * Q.o2i;
    al[1] = (PH(2, 1) - PH(2, 4)) * Q.o1i;
    al[7] = (d_[1] * PH(2, 3) - PH(2, 4)) * Q.o1i;
    al[2] = (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * Q.hei;
    al[5] = (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * Q.hei;
#undef PH
  }
  __device__ void init() {
    rdz = T(1. / g.dz);
    dz = T(g.dz);
    dtx2inv = T(g.dtx2inv);
    thdiffalpha = T(-0.38);
    dfac = T(dfactor_at(g, j));
    const size_t c2 = G.a2(i, j);
    fhe = g.calc_helium == 0 ? T(0) : T(in.flx_he[c2]);
    tl = T(in.t_lbc[c2]);  // flds%tlbc after dt_ng of this step
    o2_0 = T(in.o2[G.a3(0, i, j)]); o1_0 = T(in.o1[G.a3(0, i, j)]); he_0 = T(in.he[G.a3(0, i, j)]);
    for (int l = 0; l < 3; ++l)
      ps0[l] = T(g.b[3 * l + 0]) * o2_0 + T(g.b[3 * l + 1]) * o1_0 + T(g.b[3 * l + 2]) * he_0 + T(g.fb[l]);
    embar0 = T(1) / (ps0[0] * T(RIO2) + ps0[1] * T(RIO1) + ps0[2] * T(RIHE) +
                     (T(1) - ps0[0] - ps0[1] - ps0[2]) * T(RIN2));
    load_r(0, r1);
    Lev Q;
    Q.o2i = T(.5) * (ps0[0] + o2_0);
    Q.o1i = T(.5) * (ps0[1] + o1_0);
    Q.hei = T(.5) * (ps0[2] + he_0);
    Q.embari = T(.5) * (embar0 + r1[3]);
    Q.tni = tl;
    Q.dembardz = r1[3] - embar0;
    Q.dtndz = (r1[4] - tl) * T(2);
    lev_df(Q);
    wks4 = Q.dembardz / (Q.embari * T(2));
    for (int l = 0; l < 3; ++l) epp[l] = T(1) - T(1) / Q.embari * (rmass(l) + Q.dembardz);
    epp[2] = epp[2] - thdiffalpha * Q.dtndz / Q.tni;
    T alpha[9];
    alpha_of(Q, alpha);
    matinv3_t<T>(alpha, akp);
    const T wks1 = Q.embari * T(RIN2) * Q.q * T(1. / TAU);
    for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
    embari_c = Q.embari;
  }
  // rows of level k; the boundary conditions of the bottom (k = 0) and top
  // (k = nk-2) rows are folded in as comp_kernel does
  __device__ void rows(int k, T *pk, T *qk, T *rk, T *fk) {
    const int nk = G.nk;
    T r2[6];
    load_r(k + 1, r2);
    const size_t p = G.a3(k, i, j);
    const size_t t = G.tile(k, i, j, 0, CI_NV);
    const T rhs0 = T(CT.ci[t]), rhs1 = T(CT.ci[t + 64]), rhs2 = T(CT.ci[t + 128]);
    const T fs01 = T(CT.ci[t + 192]), fs02 = T(CT.ci[t + 256]), fs11 = T(CT.ci[t + 320]), fs12 = T(CT.ci[t + 384]);
    const T hdo2 = T(in.hdo2[p]), hdo1 = T(in.hdo1[p]), hdhe = T(in.hdhe[p]);
    const T expzk = T(g.expz[k]), difk0 = T(g.difk[k]), difk1 = T(g.difk[k + 1]);
    T akm[9], epm[3];
    for (int n = 0; n < 9; ++n) akm[n] = akp[n];
    for (int n = 0; n < 3; ++n) epm[n] = epp[n];
    Lev Q;
    Q.o2i = T(.5) * (r1[0] + r2[0]);
    Q.o1i = T(.5) * (r1[1] + r2[1]);
    Q.hei = T(.5) * (r1[2] + r2[2]);
    Q.embari = T(.5) * (r1[3] + r2[3]);
    Q.tni = T(.5) * (r1[4] + r2[4]);
    Q.dembardz = r2[3] - r1[3];
    Q.dtndz = r2[4] - r1[4];
    lev_df(Q);
    const T rem = T(1) / Q.embari;
    for (int l = 0; l < 3; ++l) epp[l] = T(1) - rem * (rmass(l) + Q.dembardz);
    if (k != nk - 2) epp[2] = epp[2] - thdiffalpha * Q.dtndz / Q.tni;
    {
      T alpha[9];
      alpha_of(Q, alpha);
      matinv3_t<T>(alpha, akp);
    }
    const T wks1 = Q.embari * T(RIN2) * Q.q * T(1. / TAU);
    const T wks3 = wks4;
    wks4 = Q.dembardz / (Q.embari * T(2));
    for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
    const T eddyp = T(g.expzmid_inv) * difk0 * (rdz - wks3);
    const T eddyr = T(g.expzmid) * difk1 * (rdz + wks4);
    const T eddyq = T(g.expzmid_inv) * difk0 * (rdz + wks3) + T(g.expzmid) * difk1 * (rdz - wks4);
    const T wik = T(.5) * (r1[5] + r2[5]);
    const T fs[2][3] = {{T(0), fs01, fs02}, {T(0), fs11, fs12}};
    for (int m = 0; m < 3; ++m)
      for (int isp = 0; isp < 3; ++isp) {
        const int e = isp + 3 * m;
        const T molp = akm[e] * (rdz + epm[m] / T(2));
        const T molr = akp[e] * (rdz - epp[m] / T(2));
        const T molq = akm[e] * (rdz - epm[m] / T(2)) + akp[e] * (rdz + epp[m] / T(2));
        const T delta = isp == m ? T(1) : T(0);
        const T fsv = (isp < 2 && m < 2) ? fs[isp][m + 1] : T(0);
        pk[e] = molp * rdz - expzk * delta * rdz * (eddyp * dfac + T(.5) * wik);
        rk[e] = molr * rdz - expzk * delta * rdz * (eddyr * dfac - T(.5) * wik);
        qk[e] = -molq * rdz + expzk * delta * (eddyq * dfac * rdz + dtx2inv) - expzk * fsv;
      }
    fk[0] = expzk * (rhs0 + hdo2);
    fk[1] = expzk * (rhs1 + hdo1);
    fk[2] = expzk * (rhs2 + hdhe);
    if (k == 0) {
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp)
            qk[isp + 3 * m] = qk[isp + 3 * m] + pk[isp + 3 * n] * T(g.b[3 * n + m]);
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) fk[isp] = fk[isp] - pk[isp + 3 * m] * T(g.fb[m]);
      for (int n = 0; n < 9; ++n) pk[n] = T(0);
    } else if (k == nk - 2) {
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp)
          qk[isp + 3 * m] = qk[isp + 3 * m] + (T(1) + T(.5) * epp[m] * dz) / (T(1) - T(.5) * epp[m] * dz) *
                                                  rk[isp + 3 * m];
      const T *d_ = Q.df;
#define PH(r, c) T(PHI(r, c))
      alpha22 = -(PH(2, 4) + (PH(2, 1) - PH(2, 4)) * Q.o2i + (d_[1] * PH(2, 3) - PH(2, 4)) * Q.hei);
      alpha33 = -(d_[2] * PH(3, 4) + (d_[0] * PH(3, 1) - d_[2] * PH(3, 4)) * Q.o2i +
                  (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * Q.o1i);
      alpha23 = (d_[1] * PH(2, 3) - PH(2, 4)) * Q.o1i;
      alpha32 = (d_[1] * PH(3, 2) - d_[2] * PH(3, 4)) * Q.hei;
#undef PH
      flx00 = embari_c * T(RIN2) * T(P0) * t_exp<T>(T(0.25) * t_log<T>(T(T00) / r1[4])) / T(TAU * GRAV);
      for (int isp = 0; isp < 3; ++isp)
        fk[isp] = fk[isp] - rk[isp + 3 * 1] * (alpha23 - alpha22) * fhe / (flx00 * (rdz - T(0.5) * epp[1])) -
                  rk[isp + 3 * 2] * (alpha33 - alpha32) * fhe / (flx00 * (rdz - T(0.5) * epp[2]));
      for (int n = 0; n < 9; ++n) rk[n] = T(0);
    }
    embari_c = Q.embari;
    for (int n = 0; n < 6; ++n) r1[n] = r2[n];
  }
  // :317-332, the top level from the solution at nk-2
  __device__ void top_of(const double *u, double *t) const {
    for (int isp = 0; isp < 3; ++isp)
      t[isp] = double((T(1) + T(.5) * epp[isp] * dz) / (T(1) - T(.5) * epp[isp] * dz)) * u[isp];
    t[1] = t[1] + double((alpha23 - alpha22) * fhe / (flx00 * (rdz - T(0.5) * epp[1])));
    t[2] = t[2] + double((alpha33 - alpha32) * fhe / (flx00 * (rdz - T(0.5) * epp[2])));
  }
};

// One level of the FP32 block elimination with an FP64 pivot inverse:
// gamma(k) = W r_k, z(k) = W (f_k - p_k z(k-1)), W = (q_k - p_k gamma(k-1))^-1.
__device__ __forceinline__ void comp_elim32(const float *pk, float *qk, const float *rk, const float *fk,
                                            float *gama, float *zz) {
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n)
      for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] - pk[isp + 3 * n] * gama[n + 3 * m];
  double q64[9], w[9];
  for (int n = 0; n < 9; ++n) q64[n] = qk[n];
  matinv3_t<double>(q64, w);
  float wkv[3];
  for (int isp = 0; isp < 3; ++isp) wkv[isp] = fk[isp];
  for (int m = 0; m < 3; ++m)
    for (int isp = 0; isp < 3; ++isp) wkv[isp] = wkv[isp] - pk[isp + 3 * m] * zz[m];
  for (int m = 0; m < 3; ++m)
    for (int isp = 0; isp < 3; ++isp) {
      double a = 0.;
      for (int n = 0; n < 3; ++n) a = a + w[isp + 3 * n] * (double)rk[n + 3 * m];
      gama[isp + 3 * m] = (float)a;
    }
  for (int isp = 0; isp < 3; ++isp) {
    double a = 0.;
    for (int m = 0; m < 3; ++m) a = a + w[isp + 3 * m] * (double)wkv[m];
    zz[isp] = (float)a;
  }
}

// Pass 0: FP32 assembly and elimination, u to the tiled region (FP64).
__global__ void __launch_bounds__(256) comp_mixed_solve(tgx_ng_grid g, Geo G, tgx_ng_fields in, float *gs,
                                                        double *U, CompTerms CT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  double v;
  if (lateral_value(G, in.lonb_o2, in.latb_o2, 0, i, j, v)) return;
  const int nk = G.nk;
  CompRows<float> R(g, G, in, CT, i, j);
  R.init();
  float gama[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, zz[3] = {0, 0, 0};
  for (int k = 0; k < nk - 1; ++k) {
    float pk[9], qk[9], rk[9], fk[3];
    R.rows(k, pk, qk, rk, fk);
    comp_elim32(pk, qk, rk, fk, gama, zz);
    for (int n = 0; n < 9; ++n) gs[G.tile(k, i, j, n, 12)] = gama[n];
    for (int n = 0; n < 3; ++n) gs[G.tile(k, i, j, 9 + n, 12)] = zz[n];
  }
  float u1[3] = {zz[0], zz[1], zz[2]};
  double ud[3] = {u1[0], u1[1], u1[2]}, t[3];
  R.top_of(ud, t);
  for (int n = 0; n < 3; ++n) {
    U[G.tile(nk - 1, i, j, n, CM_NV)] = t[n];
    U[G.tile(nk - 2, i, j, n, CM_NV)] = ud[n];
  }
  for (int k = nk - 3; k >= 0; --k) {
    float gm[9], z[3], u[3];
    for (int n = 0; n < 9; ++n) gm[n] = gs[G.tile(k, i, j, n, 12)];
    for (int n = 0; n < 3; ++n) z[n] = gs[G.tile(k, i, j, 9 + n, 12)];
    for (int isp = 0; isp < 3; ++isp) {
      float uu = z[isp];
      for (int m = 0; m < 3; ++m) uu = uu - gm[isp + 3 * m] * u1[m];
      u[isp] = uu;
    }
    for (int n = 0; n < 3; ++n) {
      U[G.tile(k, i, j, n, CM_NV)] = (double)u[n];
      u1[n] = u[n];
    }
  }
}

// A refinement sweep (REFINE) and/or the post pass :386-457 (FINAL).
template <bool REFINE, bool FINAL>
__global__ void __launch_bounds__(256) comp_mixed_refine(tgx_ng_grid g, Geo G, tgx_ng_fields in, tgx_ng_out out,
                                                         float *gs, double *U, CompTerms CT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const size_t c2 = G.a2(i, j);
  if (FINAL) out.flx_he[c2] = g.calc_helium == 0 ? 0. : in.flx_he[c2];
  double v;
  if (lateral_value(G, in.lonb_o2, in.latb_o2, 0, i, j, v)) {
    if (FINAL) comp_lateral_column(g, G, in, out, i, j);
    return;
  }
  auto uat = [&](int k, int n) -> double & { return U[G.tile(k, i, j, n, CM_NV)]; };
  if (REFINE) {
    CompRows<double> R(g, G, in, CT, i, j);
    R.init();
    float gama[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, zz[3] = {0, 0, 0};
    double um[3] = {0, 0, 0}, u0[3], up[3];
    for (int n = 0; n < 3; ++n) u0[n] = uat(0, n);
    for (int k = 0; k < nk - 1; ++k) {
      for (int n = 0; n < 3; ++n) up[n] = k + 1 <= nk - 2 ? uat(k + 1, n) : 0.;
      double pk[9], qk[9], rk[9], fk[3];
      R.rows(k, pk, qk, rk, fk);
      // r = f - (p u(k-1) + q u(k) + r u(k+1)) in FP64
      float p32[9], q32[9], r32[9], f32[3];
      for (int isp = 0; isp < 3; ++isp) {
        double a = fk[isp];
        for (int m = 0; m < 3; ++m)
          a = a - (pk[isp + 3 * m] * um[m] + qk[isp + 3 * m] * u0[m] + rk[isp + 3 * m] * up[m]);
        f32[isp] = (float)a;
      }
      for (int n = 0; n < 9; ++n) {
        p32[n] = (float)pk[n];
        q32[n] = (float)qk[n];
        r32[n] = (float)rk[n];
      }
      comp_elim32(p32, q32, r32, f32, gama, zz);
      for (int n = 0; n < 9; ++n) gs[G.tile(k, i, j, n, 12)] = gama[n];
      for (int n = 0; n < 3; ++n) gs[G.tile(k, i, j, 9 + n, 12)] = zz[n];
      for (int n = 0; n < 3; ++n) {
        um[n] = u0[n];
        u0[n] = up[n];
      }
    }
    // correction, top down, u += d; the top level again from the closed form
    float d1[3] = {zz[0], zz[1], zz[2]};
    double un[3], t[3];
    for (int n = 0; n < 3; ++n) un[n] = uat(nk - 2, n) + (double)d1[n];
    R.top_of(un, t);
    for (int n = 0; n < 3; ++n) {
      uat(nk - 2, n) = un[n];
      uat(nk - 1, n) = t[n];
    }
    for (int k = nk - 3; k >= 0; --k) {
      float gm[9], z[3], d[3];
      for (int n = 0; n < 9; ++n) gm[n] = gs[G.tile(k, i, j, n, 12)];
      for (int n = 0; n < 3; ++n) z[n] = gs[G.tile(k, i, j, 9 + n, 12)];
      for (int isp = 0; isp < 3; ++isp) {
        float dd = z[isp];
        for (int m = 0; m < 3; ++m) dd = dd - gm[isp + 3 * m] * d1[m];
        d[isp] = dd;
      }
      for (int n = 0; n < 3; ++n) {
        uat(k, n) = uat(k, n) + (double)d[n];
        d1[n] = d[n];
      }
    }
  }
  if (!FINAL) return;
  // :386-457 on u: time smoothing, N2 above the O-dominated top, clamps and
  // normalisation, as comp_kernel's post pass
  const double small = 1.e-9, rdz = 1. / g.dz;
  auto put_norm = [&](size_t p, double o2u, double o1u, double heu, double n2) {
    double a = o2u > small ? o2u : small, b = o1u > small ? o1u : small;
    double c = heu > small ? heu : small, d = n2 > small ? n2 : small;
    double norm = a + b + c + d;
    out.o2[p] = a / norm;
    out.o1[p] = b / norm;
    out.he[p] = c / norm;
  };
  int kf = -1;
  double n2kf = 0., xnkf = 0., n2b = 0., xnb = 0.;
  for (int k = nk - 1; k >= 0; --k) {
    const size_t p = G.a3(k, i, j);
    const double o2u = uat(k, 0), o1u = uat(k, 1), heu = g.calc_helium == 0 ? g.pshelb : uat(k, 2);
    double o2nm = DTSMOOTH_DIV2 * (in.o2_nm[p] + o2u) + DTSMOOTH * in.o2[p];
    double o1nm = DTSMOOTH_DIV2 * (in.o1_nm[p] + o1u) + DTSMOOTH * in.o1[p];
    double henm = DTSMOOTH_DIV2 * (in.he_nm[p] + heu) + DTSMOOTH * in.he[p];
    double n2 = 1. - o2u - o1u - heu;
    double n2nm = 1. - o2nm - o1nm - henm;
    const double xn = in.xnmbar[p];
    const bool lo = n2 < 0.01, below = !lo && kf == k + 1;
    n2kf = lo ? n2 : n2kf;
    xnkf = lo ? xn : xnkf;
    kf = lo ? k : kf;
    n2b = below ? n2 : n2b;
    xnb = below ? xn : xnb;
    put_norm(p, o2u, o1u, heu, n2);
    double an = o2nm > small ? o2nm : small, bn = o1nm > small ? o1nm : small;
    double cn = henm > small ? henm : small, dn = n2nm > small ? n2nm : small;
    double normn = an + bn + cn + dn;
    out.o2_nm[p] = an / normn;
    out.o1_nm[p] = bn / normn;
    out.he_nm[p] = cn / normn;
  }
  if (kf >= 1 && kf + 1 < nk) {  // :419-431 extrapolated N2 above kf
    const double zmtop = g.zpmid[kf];
    const double n2top = log(n2kf * xnkf);
    const double n2top1 = log(n2b * xnb);
    const double dn2dz = (n2top - n2top1) * rdz;
    for (int k = nk - 1; k > kf; --k) {
      const size_t p = G.a3(k, i, j);
      const double n2 = exp(n2top + dn2dz * (g.zpmid[k] - zmtop)) / in.xnmbar[p];
      put_norm(p, uat(k, 0), uat(k, 1), g.calc_helium == 0 ? g.pshelb : uat(k, 2), n2);
    }
  }
}

// comp_ng.F90:71-457 with one wavefront per column: the 3x3 block rows
// 0..nk-2 by block cyclic reduction, the closed-form top level (:317-332),
// then the post pass (:415-457) with the N2 extrapolation level found by a
// wave ballot.
template <int NH>
__global__ void __launch_bounds__(256) comp_wave_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                        tgx_ng_out out, CompTerms CT) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk;
  const int Lr[2] = {lane, lane + 64};
  const double dz = g.dz, dtx2inv = g.dtx2inv, thdiffalpha = -0.38, small = 1.e-9, rdz = 1. / g.dz;
  const size_t c2 = G.a2(i, j);
  const double fhe = g.calc_helium == 0 ? 0. : in.flx_he[c2];
  if (lane == 0) out.flx_he[c2] = fhe;
  double x[2][3] = {};  // O2, O, He at this lane's levels before the post pass
  double v;
  const bool bnd = lateral_value(G, in.lonb_o2, in.latb_o2, 0, i, j, v);
  if (bnd) {
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int Lc = Lr[h] < nk ? Lr[h] : nk - 1;
      lateral_value(G, in.lonb_o2, in.latb_o2, Lc, i, j, x[h][0]);
      lateral_value(G, in.lonb_o1, in.latb_o1, Lc, i, j, x[h][1]);
      lateral_value(G, in.lonb_he, in.latb_he, Lc, i, j, x[h][2]);
    }
  } else {
    const double dfac = dfactor_at(g, j);
    const double tl = in.t_lbc[c2];
    const double rmass[3] = {RMO2, RMO1, RMHE};
    const double o2_0 = in.o2[c2], o1_0 = in.o1[c2], he_0 = in.he[c2];
    double ps0[3];
    for (int l = 0; l < 3; ++l)
      ps0[l] = g.b[3 * l + 0] * o2_0 + g.b[3 * l + 1] * o1_0 + g.b[3 * l + 2] * he_0 + g.fb[l];
    const double embar0 =
        1. / (ps0[0] * RIO2 + ps0[1] * RIO1 + ps0[2] * RIHE + (1. - ps0[0] - ps0[1] - ps0[2]) * RIN2);
    struct LR { double o2, o1, he, mbar, tn, w; };
    LR R[2] = {}, A[2] = {};
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      size_t p = G.a3(Lr[h] < nk ? Lr[h] : nk - 1, i, j);
      R[h].o2 = in.o2[p]; R[h].o1 = in.o1[p]; R[h].he = in.he[p]; R[h].mbar = in.mbar[p];
      R[h].tn = in.tn[p]; R[h].w = in.w[p];
    }
    wave_dn_h<NH>(R[0].o2, R[1].o2, lane, A[0].o2, A[1].o2);
    wave_dn_h<NH>(R[0].o1, R[1].o1, lane, A[0].o1, A[1].o1);
    wave_dn_h<NH>(R[0].he, R[1].he, lane, A[0].he, A[1].he);
    wave_dn_h<NH>(R[0].mbar, R[1].mbar, lane, A[0].mbar, A[1].mbar);
    wave_dn_h<NH>(R[0].tn, R[1].tn, lane, A[0].tn, A[1].tn);
    // interface L (between levels L-1 and L): ak = matinv3(alpha)*wks1, ep, wks
    double ak[2][9] = {}, ep[2][3] = {}, wks[2] = {}, a22[2] = {}, a33[2] = {}, a23[2] = {}, a32[2] = {},
           flx[2] = {};
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int L = Lr[h];
      double o2i, o1i, hei, embari, tni, dembardz, dtndz, df[3];
      if (L == 0) {  // :80-97
        o2i = .5 * (ps0[0] + o2_0);
        o1i = .5 * (ps0[1] + o1_0);
        hei = .5 * (ps0[2] + he_0);
        embari = .5 * (embar0 + R[h].mbar);
        tni = tl;
        dembardz = R[h].mbar - embar0;
        dtndz = (R[h].tn - tl) * 2;
      } else {
        o2i = .5 * (A[h].o2 + R[h].o2);
        o1i = .5 * (A[h].o1 + R[h].o1);
        hei = .5 * (A[h].he + R[h].he);
        embari = .5 * (A[h].mbar + R[h].mbar);
        tni = .5 * (A[h].tn + R[h].tn);
        dembardz = R[h].mbar - A[h].mbar;
        dtndz = R[h].tn - A[h].tn;
      }
      dembardz = dembardz * rdz;
      dtndz = dtndz * rdz;
      const TPow tp = tpows(tni);  // the temperature powers from one logarithm
      for (int n = 0; n < 3; ++n) df[n] = tp.df[n];
      for (int l = 0; l < 3; ++l) ep[h][l] = 1. - 1. / embari * (rmass[l] + dembardz);
      if (L <= nk - 2) ep[h][2] = ep[h][2] - thdiffalpha * dtndz / tni;
      double al[9];
      const double *d_ = df;
      al[0] = -(PHI(1, 4) + (PHI(1, 2) - PHI(1, 4)) * o1i + (d_[0] * PHI(1, 3) - PHI(1, 4)) * hei);
      al[4] = -(PHI(2, 4) + (PHI(2, 1) - PHI(2, 4)) * o2i + (d_[1] * PHI(2, 3) - PHI(2, 4)) * hei);
      al[8] = -(d_[2] * PHI(3, 4) + (d_[0] * PHI(3, 1) - d_[2] * PHI(3, 4)) * o2i +
                (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * o1i);
      al[3] = (PHI(1, 2) - PHI(1, 4)) * o2i;
      al[6] = (d_[0] * PHI(1, 3) - PHI(1, 4)) * o2i;
      al[1] = (PHI(2, 1) - PHI(2, 4)) * o1i;
      al[7] = (d_[1] * PHI(2, 3) - PHI(2, 4)) * o1i;
      al[2] = (d_[0] * PHI(3, 1) - d_[2] * PHI(3, 4)) * hei;
      al[5] = (d_[1] * PHI(3, 2) - d_[2] * PHI(3, 4)) * hei;
      matinv3_r(al, ak[h]);
      double wks1 = embari * RIN2 * tp.q / TAU;
      for (int n = 0; n < 9; ++n) ak[h][n] = ak[h][n] * wks1;
      wks[h] = dembardz / (embari * 2.);
      // upper boundary terms of the top interface (:601-652)
      a22[h] = al[4];
      a33[h] = al[8];
      a23[h] = al[7];
      a32[h] = al[5];
      flx[h] = 0.;
      if (L == nk - 2) flx[h] = embari * RIN2 * P0 * pow(T00 / R[h].tn, 0.25) / (TAU * GRAV);
    }
    // the interface above (L+1) from the next lane
    double akp[2][9], epp[2][3], wksp[2], wp[2], a22p[2], a33p[2], a23p[2], a32p[2];
    for (int n = 0; n < 9; ++n) wave_up_h<NH>(ak[0][n], ak[1][n], lane, akp[0][n], akp[1][n]);
    for (int n = 0; n < 3; ++n) wave_up_h<NH>(ep[0][n], ep[1][n], lane, epp[0][n], epp[1][n]);
    wave_up_h<NH>(wks[0], wks[1], lane, wksp[0], wksp[1]);
    wave_up_h<NH>(R[0].w, R[1].w, lane, wp[0], wp[1]);
    wave_up_h<NH>(a22[0], a22[1], lane, a22p[0], a22p[1]);
    wave_up_h<NH>(a33[0], a33[1], lane, a33p[0], a33p[1]);
    wave_up_h<NH>(a23[0], a23[1], lane, a23p[0], a23p[1]);
    wave_up_h<NH>(a32[0], a32[1], lane, a32p[0], a32p[1]);
    Nrm3 rows[2] = {};
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int k = Lr[h];
      Nrm3 &N3 = rows[h];
      if (k > nk - 2) {
        for (int e = 0; e < 9; ++e) { N3.A[e] = 0.; N3.C[e] = 0.; }
        for (int e = 0; e < 3; ++e) N3.F[e] = 0.;
        continue;
      }
      size_t p = G.a3(k, i, j);
      const double expzk = g.expz[k];
      double fs[2][3];
      const size_t t = G.tile(k, i, j, 0, CI_NV);
      fs[0][1] = CT.ci[t + 192]; fs[0][2] = CT.ci[t + 256]; fs[1][1] = CT.ci[t + 320];
      fs[1][2] = CT.ci[t + 384];
      const double *akm = ak[h], *epm = ep[h], *akpp = akp[h], *eppp = epp[h];
      const double wks3 = wks[h], wks4 = wksp[h];
      double pk[9], rk[9], qk[9], fk[3];
      double eddyp = g.expzmid_inv * g.difk[k] * (rdz - wks3);
      double eddyr = g.expzmid * g.difk[k + 1] * (rdz + wks4);
      double eddyq = g.expzmid_inv * g.difk[k] * (rdz + wks3) + g.expzmid * g.difk[k + 1] * (rdz - wks4);
      const double wik = .5 * (R[h].w + wp[h]);
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) {
          int e = isp + 3 * m;
          double molp = akm[e] * (rdz + epm[m] / 2.);
          double molr = akpp[e] * (rdz - eppp[m] / 2.);
          double molq = akm[e] * (rdz - epm[m] / 2.) + akpp[e] * (rdz + eppp[m] / 2.);
          double delta = isp == m ? 1. : 0.;
          double fsv = (isp < 2 && m < 2) ? fs[isp][m + 1] : 0.;
          pk[e] = molp * rdz - expzk * delta * rdz * (eddyp * dfac + .5 * wik);
          rk[e] = molr * rdz - expzk * delta * rdz * (eddyr * dfac - .5 * wik);
          qk[e] = -molq * rdz + expzk * delta * (eddyq * dfac * rdz + dtx2inv) - expzk * fsv;
        }
      fk[0] = expzk * (CT.ci[t] + in.hdo2[p]);
      fk[1] = expzk * (CT.ci[t + 64] + in.hdo1[p]);
      fk[2] = expzk * (CT.ci[t + 128] + in.hdhe[p]);
      if (k == 0) {
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n)
            for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] + pk[isp + 3 * n] * g.b[3 * n + m];
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp) fk[isp] = fk[isp] - pk[isp + 3 * m] * g.fb[m];
        for (int n = 0; n < 9; ++n) pk[n] = 0.;
      }
      if (k == nk - 2) {
        for (int m = 0; m < 3; ++m)
          for (int isp = 0; isp < 3; ++isp)
            qk[isp + 3 * m] = qk[isp + 3 * m] + (1. + .5 * eppp[m] * dz) / (1. - .5 * eppp[m] * dz) * rk[isp + 3 * m];
        const double flx00 = flx[h];
        const double alpha22 = a22p[h], alpha33 = a33p[h], alpha23 = a23p[h], alpha32 = a32p[h];
        for (int isp = 0; isp < 3; ++isp)
          fk[isp] = fk[isp] - rk[isp + 3 * 1] * (alpha23 - alpha22) * fhe / (flx00 * (rdz - 0.5 * eppp[1])) -
                    rk[isp + 3 * 2] * (alpha33 - alpha32) * fhe / (flx00 * (rdz - 0.5 * eppp[2]));
        for (int n = 0; n < 9; ++n) rk[n] = 0.;
      }
      // to row-major blocks, normalised by the diagonal block
      double Bq[9], Ap[9], Cr[9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          Bq[3 * r + c] = qk[r + 3 * c];
          Ap[3 * r + c] = pk[r + 3 * c];
          Cr[3 * r + c] = rk[r + 3 * c];
        }
      double Bi[9];
      blk_inv<3>(Bq, Bi);
      blk_mm<3>(Bi, Ap, N3.A);
      blk_mm<3>(Bi, Cr, N3.C);
      blk_mv<3>(Bi, fk, N3.F);
    }
    if constexpr (NH == 2)
      wave_pcr3(rows[0], rows[1], lane);
    else
      wave_pcr3_1(rows[0], lane);
    // the closed-form top level (:317-332) from level nk-2
    double xm[2][3], flxm[2];
#pragma unroll
    for (int n = 0; n < 3; ++n) wave_dn_h<NH>(rows[0].F[n], rows[1].F[n], lane, xm[0][n], xm[1][n]);
    wave_dn_h<NH>(flx[0], flx[1], lane, flxm[0], flxm[1]);
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      for (int n = 0; n < 3; ++n) x[h][n] = rows[h].F[n];
      if (Lr[h] == nk - 1) {
        const double *e_ = ep[h];  // ep of the top interface
        double top[3];
        for (int isp = 0; isp < 3; ++isp) top[isp] = (1. + .5 * e_[isp] * dz) / (1. - .5 * e_[isp] * dz) * xm[h][isp];
        const double flx00 = flxm[h];
        top[1] = top[1] + (a23[h] - a22[h]) * fhe / (flx00 * (rdz - 0.5 * e_[1]));
        top[2] = top[2] + (a33[h] - a32[h]) * fhe / (flx00 * (rdz - 0.5 * e_[2]));
        for (int n = 0; n < 3; ++n) x[h][n] = top[n];
      }
      if (g.calc_helium == 0) x[h][2] = g.pshelb;
    }
  }
  // :415-457 time smoothing, N2 above the O-dominated top, clamps, normalisation
  double n2[2] = {}, xn[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    size_t p = G.a3(Lr[h] < nk ? Lr[h] : nk - 1, i, j);
    xn[h] = in.xnmbar[p];
    n2[h] = 1. - x[h][0] - x[h][1] - x[h][2];
  }
  // first level with n2 < 0.01 over the column
  const unsigned long long blo = __ballot(Lr[0] < nk && n2[0] < 0.01);
  const unsigned long long bhi = __ballot(Lr[1] < nk && n2[1] < 0.01);
  int kf = -1;
  if (blo) kf = __ffsll((long long)blo) - 1;
  else if (bhi) kf = 64 + __ffsll((long long)bhi) - 1;
  bool extrap = false;
  double n2top = 0., dn2dz = 0., zmtop = 0.;
  {
    // n2 and xnmbar at kf and kf-1 from the lanes holding them (uniform)
    int kk = kf >= 0 ? kf : 0, km1 = kf >= 1 ? kf - 1 : 0;
    double n2a = __shfl(kk < 64 ? n2[0] : n2[1], kk & 63, 64);
    double xna = __shfl(kk < 64 ? xn[0] : xn[1], kk & 63, 64);
    double n2b = __shfl(km1 < 64 ? n2[0] : n2[1], km1 & 63, 64);
    double xnb = __shfl(km1 < 64 ? xn[0] : xn[1], km1 & 63, 64);
    if (kf >= 1 && kf + 1 < nk) {
      extrap = true;
      zmtop = g.zpmid[kf];
      n2top = log(n2a * xna);
      double n2top1 = log(n2b * xnb);
      dn2dz = (n2top - n2top1) * rdz;
    }
  }
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int k = Lr[h];
    if (k >= nk) continue;
    size_t p = G.a3(k, i, j);
    double o2u = x[h][0], o1u = x[h][1], heu = x[h][2];
    double o2nm = DTSMOOTH_DIV2 * (in.o2_nm[p] + o2u) + DTSMOOTH * in.o2[p];
    double o1nm = DTSMOOTH_DIV2 * (in.o1_nm[p] + o1u) + DTSMOOTH * in.o1[p];
    double henm = DTSMOOTH_DIV2 * (in.he_nm[p] + heu) + DTSMOOTH * in.he[p];
    double n2v = n2[h];
    double n2nm = 1. - o2nm - o1nm - henm;
    if (extrap && k > kf) n2v = exp(n2top + dn2dz * (g.zpmid[k] - zmtop)) / xn[h];
    double a = o2u > small ? o2u : small, b = o1u > small ? o1u : small;
    double c = heu > small ? heu : small, d = n2v > small ? n2v : small;
    double norm = a + b + c + d;
    out.o2[p] = a / norm;
    out.o1[p] = b / norm;
    out.he[p] = c / norm;
    double an = o2nm > small ? o2nm : small, bn = o1nm > small ? o1nm : small;
    double cn = henm > small ? henm : small, dn = n2nm > small ? n2nm : small;
    double normn = an + bn + cn + dn;
    out.o2_nm[p] = an / normn;
    out.o1_nm[p] = bn / normn;
    out.he_nm[p] = cn / normn;
  }
}


// =============================================================================
// oplus_ng.F90, sub-step stage 1 (every column, ghosts included): the
// pointwise and vertical quantities whose horizontal derivatives are needed
// (:106-170).  Writes dj, bdotu, bvel, tp_op, diffj, djbz, op/bmod2**2 to the
// workspace and Fe, Fn to the outputs.  A thread walks one of nrun runs of
// levels of one column (the vertical derivative of tp_op carries three
// levels): the whole column on large grids, short runs on small nest
// domains so that they still fill the chip (oplus_runs).
struct OpWork {
  double *dj, *bdotu, *bvel, *tpop, *diffj, *djbz, *opbm2, *bdopj, *w2;
  double *fac;  // the factorisation of the sub-cycles' common matrix (OF_NV values per level)
};
struct OpTerms {  // prelude planes of one sub-step
  const double *bd_diff, *bd_diffj, *bd_bvel, *vdotn, *bd_djbz, *opts2;
};

// FULL = false (the middle sub-steps of a sub-cycled step): only the values
// that depend on O+ (bvel, tp_op, diffj, op/bmod2**2); dj, bdotu and dj*bz
// stay those of the first sub-step, and Fe, Fn are left to the last one.
template <bool RUNS, bool FULL>  // RUNS false: nrun == 1, the whole column (a fixed loop)
__global__ void __launch_bounds__(256) oplus_stage1(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                    const double *__restrict__ op, OpWork W,
                                                    double *__restrict__ Fe, double *__restrict__ Fn,
                                                    int nrun) {
  const int nb = gridDim.x;
  const long tid = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  const int nk = G.nk;
  if (!RUNS) nrun = 1;
  if (tid >= (long)G.nlond * G.nlatd * nrun) return;
  const int i = (int)(tid % G.nlond);
  const long r = tid / G.nlond;
  const int j = (int)(r % G.nlatd);
  const int run = RUNS ? (int)(r / G.nlatd) : 0;
  const int k0 = RUNS ? (int)((long)run * nk / nrun) : 0, k1 = RUNS ? (int)((long)(run + 1) * nk / nrun) : nk;
  const double dz = g.dz;
  const size_t c2 = G.a2(i, j);
  const double bx = in.bx[c2], by = in.by[c2], bz = in.bz[c2], bm = in.bmod2[c2];
  const double sncsdip = sin(in.dipmag[c2]) * cos(in.dipmag[c2]);
  const double sndec = in.sndec[c2], csdec = in.csdec[c2];
#define F3(f, k) (f)[G.a3((k), i, j)]
  auto tp_at = [&](int L) { return (F3(in.te, L) + F3(in.ti, L)) * F3(op, L); };
  double tpm = k0 > 0 ? tp_at(k0 - 1) : 0., tpc = tp_at(k0), tpn = k0 + 1 < nk ? tp_at(k0 + 1) : tpc;
  for (int k = k0; k < k1; ++k) {
    size_t p = G.a3(k, i, j);
    double hj = in.scht[p];
    double dtp;
    if (k == 0) dtp = tpn - tpc;
    else if (k <= nk - 3) dtp = (tpn - tpm) / 2.;
    else dtp = tpc - tpm;
    double diffj = 1. / (hj * dz) * dtp + RMOP * GRAV / GASK * op[p];
    if (!FULL) {
      W.bvel[p] = W.bdotu[p] * op[p];
      W.tpop[p] = tpc;
      W.diffj[p] = diffj;
      W.opbm2[p] = op[p] / (bm * bm);
      tpm = tpc;
      tpc = tpn;
      if (k + 2 < nk) tpn = tp_at(k + 2);
      continue;
    }
    double wni = k < nk - 1 ? .5 * (in.w[p] + in.w[p + G.P]) : 1.5 * in.w[p] - .5 * in.w[p - G.P];
    double tr = 0.5 * (in.tn[p] + in.ti[p]);
    double t = 1. - 0.064 * log10(tr);
    double v = 18.1 * in.o2[p] * RIO2 + in.o1[p] * RIO1 * sqrt(tr) * (t * t) * g.colfac +
               3.6 * in.he[p] * RIHE + 18.6 * in.n2[p] * RIN2;
    double dj = 1.42E17 / (in.xnmbar[p] * v);
    double vni = 16 * 3.53E-11 * v;
    if (g.opdiffcap != 0.) {
      double cap = g.opdiffcap / (1 + pow(2., g.zpint[k] - 6));
      dj = dj < cap ? dj : cap;
    }
    double bdotu = bx * in.un[p] + by * in.vn[p] + hj * bz * wni;
    double wd = vni * diffj * dj;
    Fe[p] = wd * sncsdip * sndec;
    Fn[p] = wd * sncsdip * csdec;
    W.dj[p] = dj;
    W.bdotu[p] = bdotu;
    W.bvel[p] = bdotu * op[p];
    W.tpop[p] = tpc;
    W.diffj[p] = diffj;
    W.djbz[p] = dj * bz;
    W.opbm2[p] = op[p] / (bm * bm);
    tpm = tpc;
    tpc = tpn;
    if (k + 2 < nk) tpn = tp_at(k + 2);
  }
#undef F3
}

// stage 2 (every column, nrun runs of levels as stage 1): bdotdh_opj =
// bdotdh(tp_op) * dj (:175-176)
template <bool RUNS>  // false: nrun == 1, the whole column (a fixed loop)
__global__ void __launch_bounds__(256) oplus_stage2(tgx_ng_grid g, Geo G, tgx_ng_fields in, OpWork W,
                                                    int nrun) {
  const int nb = gridDim.x;
  const long tid = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  if (!RUNS) nrun = 1;
  if (tid >= (long)G.nlond * G.nlatd * nrun) return;
  const int i = (int)(tid % G.nlond);
  const long r = tid / G.nlond;
  const int j = (int)(r % G.nlatd);
  const int run = RUNS ? (int)(r / G.nlatd) : 0;
  const int k0 = RUNS ? (int)((long)run * G.nk / nrun) : 0, k1 = RUNS ? (int)((long)(run + 1) * G.nk / nrun) : G.nk;
  const size_t c2 = G.a2(i, j);
  const double bx = in.bx[c2], by = in.by[c2], cs = g.cs[j];
  for (int k = k0; k < k1; ++k) {
    const size_t off = (size_t)k * G.P, p = G.a3(k, i, j);
    W.bdopj[p] = bdotdh(W.tpop + off, G, i, j, bx, by, cs, g.dlamda, g.dphi) * W.dj[p];
  }
}
// Grids of fewer than 65,536 columns do not fill the chip with one thread
// per column: there the chemistry pass and O+ stage 1 take runs of two
// levels, O+ stage 2 single levels, and duv's stage 1 one wavefront per
// column (nk <= 64).  Every form gives bitwise the same values;
// TGX_LARGE_GRID_FORMS=1 in the environment selects the large-grid forms on
// every grid (used by the tests to check the two against each other).
__host__ inline bool small_grid(long ncol) {
  const char *force = getenv("TGX_LARGE_GRID_FORMS");
  return ncol < 65536 && !(force && force[0] == '1' && force[1] == 0);
}
// Runs of levels per column for the O+ stages: the whole column on large
// grids (one column loop was measured faster there), else as above.
__host__ inline int oplus_runs(long ncol, int nk, int stage) {
  if (!small_grid(ncol)) return 1;
  return stage == 1 ? (nk + 1) / 2 : nk;
}

// stage 3: assembly, solve and outputs of one sub-step (:172-340)
// O+ chemistry of oplus_ng.F90:288-299 at every column and level: the
// xiop2p/xiop2d outputs, the loss and the production.  None of it depends on
// O+, so it is computed once per step for all sub-cycles (level-parallel).
__global__ void __launch_bounds__(256) oplus_chem_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                         double *__restrict__ oploss_o,
                                                         double *__restrict__ opprod_o,
                                                         double *__restrict__ xiop2p,
                                                         double *__restrict__ xiop2d) {
  int nb = gridDim.x;
  long t = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  long npt = (long)G.nlond * G.nlatd * G.nk;
  if (t >= npt) return;
  int i = (int)(t % G.nlond);
  long r = t / G.nlond;
  int j = (int)(r % G.nlatd);
  int L = (int)(r / G.nlatd);
  const int nk = G.nk;
  size_t p = G.a3(L, i, j);
  double xn = in.xnmbar[p], ne = in.ne[p];
  double qop2pi = L < nk - 1 ? .5 * (in.qop2p[p] + in.qop2p[p + G.P]) : 1.5 * in.qop2p[p] - .5 * in.qop2p[p - G.P];
  double qop2di = L < nk - 1 ? .5 * (in.qop2d[p] + in.qop2d[p + G.P]) : 1.5 * in.qop2d[p] - .5 * in.qop2d[p - G.P];
  double x2p = qop2pi / (xn * ((RK16 + RK17) * in.n2[p] * RIN2 + RK18 * in.o1[p] * RIO1) +
                         (in.rk19[p] + in.rk20[p]) * ne + RK21 + RK22);
  double x2d = (qop2di + (in.rk20[p] * ne + RK22) * x2p) /
               (xn * (RK23 * in.n2[p] * RIN2 + RK24 * in.o1[p] * RIO1 + RK26 * in.o2[p] * RIO2) + in.rk25[p] * ne + RK27);
  double oploss = xn * (in.rk1[p] * in.o2[p] * RIO2 + in.rk2[p] * in.n2[p] * RIN2 + RK10 * in.n2d[p] * RIN2D);
  xiop2p[p] = x2p;
  xiop2d[p] = x2d;
  oploss_o[p] = oploss;
  if (L < nk - 1) {  // the production enters rows 1..nk-1 only
    double qopi = .5 * (in.qop[p] + in.qop[p + G.P]);
    opprod_o[p] = qopi + (in.rk19[p] * ne + RK21) * x2p + (in.rk25[p] * ne + RK27) * x2d +
                  (RK18 * x2p + RK24 * x2d) * in.o1[p] * RIO1 * xn;
  }
}

// oplus_ng for one sub-step in one pass per column tile: the horizontal
// stencils of the O+ prelude (b.grad of bdotdh_opj, diffj, bvel and dj*bz,
// the vdotn term, the Shapiro smoothing of O+ at n-1) and the assembly and
// elimination of the O+ column (:172-340), level by level.  A block covers a TILE_X x
// TILE_Y tile; at level k the tile and its 2-point halo of the six stencil
// inputs are staged through double-buffered LDS.  Every thread of the block
// takes part in the staging and the barriers; lateral-boundary and
// out-of-domain columns do not solve.
// MODE 0: one sub-step on its own.  MODE 1: the first sub-step of a
// sub-cycled step, which also keeps the factorisation of the matrix (it does
// not depend on O+, oplus_ng.F90:252-303) in W.fac.  MODE 2: a later
// sub-step, which rebuilds only the right-hand side and runs the two sweeps
// with the kept factorisation: the same arithmetic, so bitwise the values a
// full sub-step gives.
template <bool CAP, int MODE>
__global__ void __launch_bounds__(256) oplus_tile_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                         const double *__restrict__ op,
                                                         const double *__restrict__ optm1,
                                                         const double *__restrict__ lonb,
                                                         const double *__restrict__ latb, OpWork W,
                                                         const double *__restrict__ oploss_p,
                                                         const double *__restrict__ opprod_p,
                                                         double *__restrict__ opout,
                                                         double *__restrict__ optm1out) {
  constexpr int TS = TILE_R * TILE_W, NF = 6, TL = TILE_Y * TILE_W;
  enum { T_BDOPJ = 0, T_DIFFJ, T_BVEL, T_OPBM2, T_DJBZ, T_OPTM1 };
  __shared__ double tb[2][NF][TS];
  __shared__ double tls[TL];  // latitude pass of O+ at n-1
  const int ntx = (G.nlond + TILE_X - 1) / TILE_X;
  const int bb = tgx_xcd_block(gridDim.x);
  const int i0 = (bb % ntx) * TILE_X, j0 = (bb / ntx) * TILE_Y;
  const int i = i0 + (int)threadIdx.x % TILE_X, j = j0 + (int)threadIdx.x / TILE_X;
  const bool active = i < G.nlond && j < G.nlatd;
  const bool lat = active && !G.glb && (j == 0 || j == G.nlatd - 1 || i == 0 || i == G.nlond - 1);
  const bool solve = active && !lat;
  const int nk = G.nk, nsub = g.nstep_sub;
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  int te[2];
  bool tok[2];
  size_t toff[2];
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const int e = (int)threadIdx.x + sl * 256;
    const int ii = i0 - TILE_H + e % TILE_W, jj = j0 - TILE_H + e / TILE_W;
    te[sl] = e;
    tok[sl] = e < TS && ii >= 0 && ii < G.nlond && jj >= 0 && jj < G.nlatd;
    toff[sl] = tok[sl] ? (size_t)jj * G.ld + ii : 0;
  }
  const double *tsrc[NF] = {W.bdopj, W.diffj, W.bvel, W.opbm2, W.djbz, optm1};
  double stage[NF][2];
  auto tile_load = [&](int k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int sl = 0; sl < 2; ++sl)
        if (tok[sl]) stage[f][sl] = tsrc[f][(size_t)k * G.P + toff[sl]];
  };
  auto tile_store = [&](int buf) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int sl = 0; sl < 2; ++sl)
        if (tok[sl]) tb[buf][f][te[sl]] = stage[f][sl];
  };
  const int ic = active ? i : 0, jc = active ? j : 0;
  const size_t c2 = G.a2(ic, jc);
#define F3(f, k) (f)[G.a3((k), ic, jc)]
  if (lat) {
    double v;
    for (int k = 0; k < nk; ++k) {
      size_t p = G.a3(k, i, j);
      lateral_value(G, lonb, latb, k, i, j, v);
      double nm = DTSMOOTH * op[p] + DTSMOOTH_DIV2 * (optm1[p] + v);
      opout[p] = v > 0. ? v : 0.;
      optm1out[p] = nm > 0. ? nm : 0.;
    }
  }
  const double bz = in.bz[c2], cs = g.cs[jc];
  const double bxc = in.bx[c2], byc = in.by[c2], bm = in.bmod2[c2];
  const double ssm = g.shapiro / g.nstep_sub;
  double opflux = 0., dvb = 1.;
  if (solve) {
    const double pi = 4.0 * atan(1.0), rtd = 180. / pi;
    if (MODE != 2) {  // :77-91
      const double phid = 2.0e8, phin = -2.0e8, ppolar = 0.;
      double rl = in.rlatm[c2], a = 0.;
      if (fabs(rl) >= pi / 4.5) a = 1.;
      if (fabs(rl) < pi / 4.5 && fabs(rl) > pi / 18) a = .5 * (1. + cos(fabs(rl) * 6. - pi * 4. / 3.));
      if (fabs(rl) <= pi / 18) a = 0.;
      double fed = phid * a, fen = phin * a, chi = in.chi[c2];
      if (chi >= 0.5 * pi) opflux = fen; else opflux = fed;
      if ((chi * rtd - 80.) * (chi * rtd - 100.) < 0.)
        opflux = .5 * (fed + fen) + .5 * (fed - fen) * cos(pi * (chi * rtd - 80.) / 20.);
      if (fabs(rl) >= pi / 3.) opflux = opflux + ppolar;
    }
    {  // :94-104 (df_2d of bx and of cs*by)
      double dbxdx = dflon1(in.bx, G, i, j);
      auto csby = [&](int jj) { return g.cs[jj] * in.by[G.a2(i, jj)]; };
      int J1 = G.nlatd - 1;
      double dcsbydy;
      if (j == 0) dcsbydy = csby(1) - csby(0);
      else if (j == J1) dcsbydy = csby(J1) - csby(J1 - 1);
      else dcsbydy = (csby(j + 1) - csby(j - 1)) / 2.;
      double vv = dbxdx / g.dlamda + dcsbydy / g.dphi;
      vv = vv / cs;
      dvb = (vv + 2. * bz) / RE;
    }
  }
  const double dvbz = dvb / bz;
  const double gmr = GRAV * RMOP / (2. * GASK);
  struct OL { double scht, hdz, tp1, dj, bdotu, bdopj, wi; };
  auto load_ol = [&](int L, OL &Q) {
    size_t p = G.a3(L, ic, jc);
    Q.scht = in.scht[p];
    Q.bdopj = W.bdopj[p];
    if (MODE == 2) return;  // the right-hand side needs no more
    Q.hdz = 1. / (Q.scht * dz);
    Q.tp1 = 0.5 * (in.ti[p] + in.te[p]);
    Q.dj = W.dj[p]; Q.bdotu = W.bdotu[p]; Q.wi = in.wi[p];
  };
  OL qa = {0, 0, 0, 0, 0, 0, 0}, qb = {}, qc = {};
  double ol1 = 0., qop0 = 0.;
  if (solve) {
    load_ol(0, qb);
    load_ol(1, qc);
    if (MODE != 2) {
      ol1 = F3(oploss_p, 1);
      qop0 = F3(in.qop, 0);  // lower boundary (:301-303)
    }
  }
  double w1 = 0., w3 = 0., rprev = 0.;
  tile_load(0);
  tile_store(0);
  for (int k = 0; k <= nk - 2; ++k) {
    size_t p = G.a3(k, ic, jc);
    double oploss = 0., opprod = 0., ui0 = 0., ui1 = 0., vi0 = 0., vi1 = 0.;
    if (k + 1 <= nk - 2) tile_load(k + 1);
    if (solve) {
      if (k > 0) load_ol(k + 1, qc);
      if (MODE != 2) oploss = oploss_p[p];
      opprod = opprod_p[p];
      ui0 = in.ui[p]; ui1 = in.ui[p + G.P]; vi0 = in.vi[p]; vi1 = in.vi[p + G.P];
    }
    asm volatile("" ::: "memory");
    __syncthreads();  // the level-k tiles are in buffer k&1
    const int bk = k & 1;
    const LdsPlane Tbdopj{tb[bk][T_BDOPJ], i0 - TILE_H, j0 - TILE_H},
        Tdiffj{tb[bk][T_DIFFJ], i0 - TILE_H, j0 - TILE_H}, Tbvel{tb[bk][T_BVEL], i0 - TILE_H, j0 - TILE_H},
        Topbm2{tb[bk][T_OPBM2], i0 - TILE_H, j0 - TILE_H}, Tdjbz{tb[bk][T_DJBZ], i0 - TILE_H, j0 - TILE_H},
        Toptm1{tb[bk][T_OPTM1], i0 - TILE_H, j0 - TILE_H};
    // smooth_ng.F90:283-294: latitude pass of O+ at n-1 on the tile rows and their longitude halo
    for (int e = threadIdx.x; e < TL; e += blockDim.x) {
      int ii = i0 - TILE_H + e % TILE_W, jj = j0 + e / TILE_W;
      if (ii >= 0 && ii < G.nlond && jj < G.nlatd) tls[e] = smooth_lat(Toptm1, G, ii, jj, ssm);
    }
    __syncthreads();
    if (solve) {
      // the prelude's terms at (i, j, k) (oplus_ng.F90:172-211)
      const double bd_diff = bdotdh(Tbdopj, G, i, j, bxc, byc, cs, g.dlamda, g.dphi);
      const double bd_diffj = bdotdh(Tdiffj, G, i, j, bxc, byc, cs, g.dlamda, g.dphi);
      const double bd_bvel = bdotdh(Tbvel, G, i, j, bxc, byc, cs, g.dlamda, g.dphi);
      const double bd_djbz = MODE == 2 ? 0. : bdotdh(Tdjbz, G, i, j, bxc, byc, cs, g.dlamda, g.dphi);
      const double djbz = pv(Tdjbz, G, i, j);
      double vdotn;
      {  // oplus_ng.F90:196-209
        double uii = .5 * (ui0 + ui1);
        double vii = .5 * (vi0 + vi1);
        double vx = uii * dflon1(Topbm2, G, i, j), vy = vii * dflat1(Topbm2, G, i, j);
        vx = vx / cs;
        double vh = 1. / RE * (vx / g.dlamda + vy / g.dphi);
        vdotn = vh * (bm * bm);
      }
      double opts2;
      {  // smooth_ng.F90:296-307
        const double *row = tls + (j - j0) * TILE_W + TILE_H - i0;
        int L1 = G.nlond - 1;
        bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
        int i1p = i < L1 ? i + 1 : L1, i1m = i > 0 ? i - 1 : 0;
        int i2p = i + 2 <= L1 ? i + 2 : L1, i2m = i >= 2 ? i - 2 : 0;
        double t0 = row[i];
        double fm1 = (row[i1p] + row[i1m]) * 0.5;
        double fm2 = (row[i2p] + row[i2m]) * 0.5;
        double gen = t0 - ssm * (2. * fm2 - 8. * fm1 + 6. * t0);
        double nr = t0 - ssm * (-8. * fm1 + 8. * t0);
        opts2 = edge ? t0 : (near ? nr : gen);
      }
      const bool top = k == nk - 2;
      // explicit terms (:181-211, :250)
      double dbd;
      if (k == 0) dbd = qc.bdopj - qb.bdopj;
      else if (k <= nk - 3) dbd = (qc.bdopj - qa.bdopj) / 2.;
      else dbd = qb.bdopj - qa.bdopj;
      double bdzdvb = bz / (qb.scht * dz) * dbd + dvb * qb.bdopj;
      double bd_op = djbz * bd_diffj;
      double diffexp = bdzdvb + bd_diff + bd_op;
      double expl = -1. * diffexp;
      expl = expl + vdotn + bd_bvel;
      expl = expl - opts2 * dtx2inv * nsub;
      if (MODE == 2) {  // the kept factorisation and boundary terms
        const double *F = W.fac;
        if (top) expl = expl - F[G.wtile(k, 2, OF_NV)];
        expl = expl - opprod;
        if (k == 0) expl = expl - F[G.wtile(0, 2, OF_NV)];
        if (k == 0) {
          w1 = F[G.wtile(0, 1, OF_NV)];
          w3 = expl / w1;
        } else {
          const double pcf = F[G.wtile(k, 0, OF_NV)];
          w1 = F[G.wtile(k, 1, OF_NV)];
          w3 = (expl - pcf * w3) / w1;
        }
        W.w2[G.wtile(k, 1, OW_NV)] = w3;  // forward solution
        qa = qb;
        qb = qc;
      } else {
      // diffusion, drift and wind coefficients (:239-278)
      double divbz = dvbz + bd_djbz / (qb.dj * (bz * bz));
      double hm = qb.hdz - 0.5 * divbz, hp = qb.hdz + 0.5 * divbz;
      hm = hm * (bz * bz);
      hp = hp * (bz * bz);
      double hdzi_k = k == 0 ? 1.5 * qb.hdz - 0.5 * qc.hdz : 0.5 * (qa.hdz + qb.hdz);
      double hdzi_k1 = top ? 1.5 * qb.hdz - 0.5 * qa.hdz : 0.5 * (qb.hdz + qc.hdz);
      double dji = k == 0 ? 1.5 * qb.dj - 0.5 * qc.dj : 0.5 * (qa.dj + qb.dj);
      double dji1 = top ? 1.5 * qb.dj - 0.5 * qa.dj : 0.5 * (qb.dj + qc.dj);
      double t0 = k == 0 ? 2. * (2. * qb.tp1 - qc.tp1) * hdzi_k - gmr : 2. * qa.tp1 * hdzi_k - gmr;
      double t01 = 2. * qb.tp1 * hdzi_k1 - gmr;
      double t1 = 2. * qb.tp1 * hdzi_k + gmr;
      double t11 = top ? 2. * (2. * qb.tp1 - qa.tp1) * hdzi_k1 + gmr : 2. * qc.tp1 * hdzi_k1 + gmr;
      double diffp = hm * dji * t0;
      double diffq = -(hp * dji1 * t01 + hm * dji * t1);
      double diffr = hp * dji1 * t11;
      double hdzk = qb.hdz;
      double wii = .5 * (qb.wi + qc.wi);
      double driftp = wii * 0.5 * hdzk;
      double driftq = -wii * 6. / RE;
      double driftr = -wii * 0.5 * hdzk;
      double windp = k == 0 ? bz * (2. * qb.bdotu - qc.bdotu) * 0.5 * hdzk : bz * qa.bdotu * 0.5 * hdzk;
      double windq = -qb.bdotu * dvbz * bz;
      double windr = !top ? -bz * qc.bdotu * 0.5 * hdzk : bz * (2. * qb.bdotu - qa.bdotu) * 0.5 * hdzk;
      double pcf = diffp + driftp + windp;
      double qcf = diffq + driftq + windq - dtx2inv * nsub;
      double rcf = diffr + driftr + windr;
      if (top) {  // upper boundary (:280-286)
        double djn = dji1;
        double ubcb = -(bz * bz) * djn * t01;
        double ubca = -(bz * bz) * djn * t11;
        qcf = qcf + ubcb / ubca * rcf;
        const double ub = opflux * rcf / ubca;
        if (MODE == 1) W.fac[G.wtile(k, 2, OF_NV)] = ub;
        expl = expl - ub;
        rcf = 0.;
      }
      qcf = qcf - oploss;
      expl = expl - opprod;
      if (k == 0) {  // lower boundary (:301-303)
        qcf = qcf - pcf;
        const double lb = 2. * pcf * qop0 / (1.5 * oploss - 0.5 * ol1);
        if (MODE == 1) W.fac[G.wtile(0, 2, OF_NV)] = lb;
        expl = expl - lb;
        pcf = 0.;
      }
      if (CAP) {
        cap_put(g, G, CAP_OP + 0, i, j, k, pcf); cap_put(g, G, CAP_OP + 1, i, j, k, qcf);
        cap_put(g, G, CAP_OP + 2, i, j, k, rcf); cap_put(g, G, CAP_OP + 3, i, j, k, expl);
      }
      // Thomas forward (trsolv_ng, rows 1..nk-1)
      if (k == 0) {
        w1 = qcf;
        w3 = expl / w1;
      } else {
        double w2 = rprev / w1;
        W.w2[G.wtile(k - 1, 0, OW_NV)] = w2;
        w1 = qcf - pcf * w2;
        w3 = (expl - pcf * w3) / w1;
      }
      if (MODE == 1) {
        W.fac[G.wtile(k, 0, OF_NV)] = pcf;
        W.fac[G.wtile(k, 1, OF_NV)] = w1;
      }
      W.w2[G.wtile(k, 1, OW_NV)] = w3;  // forward solution
      rprev = rcf;
      qa = qb;
      qb = qc;
      }
    }
    if (k + 1 <= nk - 2) tile_store((k + 1) & 1);
  }
  if (!solve) return;
  // back substitution, top extrapolation (:307) and outputs (:337-340); the
  // next level's loads are in flight while this level is stored
  struct OB { double x, w2, op, optm1; };
  auto load_ob = [&](int k, OB &B) {
    size_t p = G.a3(k, ic, jc);
    const size_t t = G.wtile(k, 0, OW_NV);
    B.x = W.w2[t + 64]; B.w2 = k < nk - 2 ? W.w2[t] : 0.; B.op = op[p]; B.optm1 = optm1[p];
  };
  const double op_top = F3(op, nk - 1), optm1_top = F3(optm1, nk - 1);
  double xn = w3;
  OB bc, bn;
  load_ob(nk - 2, bc);
  for (int k = nk - 2; k >= 0; --k) {
    if (k > 0) load_ob(k - 1, bn);
    asm volatile("" ::: "memory");
    double x = k == nk - 2 ? xn : bc.x - bc.w2 * xn;
    if (G.raw_row(jc)) {  // the raw solution; filter_op and the post pass follow (oplus.F:896-941)
      opout[G.a3(k, ic, jc)] = x;
      xn = x;
      bc = bn;
      continue;
    }
    if (G.glb && k == nk - 2) {  // zero at the top level (oplus.F:922-923)
      size_t pt = G.a3(nk - 1, ic, jc);
      opout[pt] = 0.;
      optm1out[pt] = 0.;
    }
    if (!G.glb && k == nk - 3) {
      double xt = xn * xn / x;
      size_t pt = G.a3(nk - 1, ic, jc);
      double nm = DTSMOOTH * op_top + DTSMOOTH_DIV2 * (optm1_top + xt);
      opout[pt] = xt > 0. ? xt : 0.;
      optm1out[pt] = nm > 0. ? nm : 0.;
    }
    size_t p = G.a3(k, ic, jc);
    double nm = DTSMOOTH * bc.op + DTSMOOTH_DIV2 * (bc.optm1 + x);
    opout[p] = x > 0. ? x : 0.;
    optm1out[p] = nm > 0. ? nm : 0.;
    xn = x;
    bc = bn;
  }
#undef F3
}

// The O+ column (:172-340) with one wavefront per column: rows 0..nk-2 by parallel
// cyclic reduction, the top level extrapolated (:307).
template <int NH>
__global__ void __launch_bounds__(256) oplus_wave_stage3(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                         const double *__restrict__ op,
                                                         const double *__restrict__ optm1,
                                                         const double *__restrict__ lonb,
                                                         const double *__restrict__ latb, OpWork W,
                                                         OpTerms OT,
                                                         const double *__restrict__ oploss_p,
                                                         const double *__restrict__ opprod_p,
                                                         double *__restrict__ opout,
                                                         double *__restrict__ optm1out) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk, nsub = g.nstep_sub;
  const int Lr[2] = {lane, lane + 64};
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
  double v;
  if (lateral_value(G, lonb, latb, 0, i, j, v)) {
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int L = Lr[h];
      if (L >= nk) continue;
      size_t p = G.a3(L, i, j);
      lateral_value(G, lonb, latb, L, i, j, v);
      double nm = DTSMOOTH * op[p] + DTSMOOTH_DIV2 * (optm1[p] + v);
      opout[p] = v > 0. ? v : 0.;
      optm1out[p] = nm > 0. ? nm : 0.;
    }
    return;
  }
  const double pi = 4.0 * atan(1.0), rtd = 180. / pi;
  const double bz = in.bz[c2], cs = g.cs[j];
  double opflux;
  {  // :77-91
    const double phid = 2.0e8, phin = -2.0e8, ppolar = 0.;
    double rl = in.rlatm[c2], a = 0.;
    if (fabs(rl) >= pi / 4.5) a = 1.;
    if (fabs(rl) < pi / 4.5 && fabs(rl) > pi / 18) a = .5 * (1. + cos(fabs(rl) * 6. - pi * 4. / 3.));
    if (fabs(rl) <= pi / 18) a = 0.;
    double fed = phid * a, fen = phin * a, chi = in.chi[c2];
    if (chi >= 0.5 * pi) opflux = fen; else opflux = fed;
    if ((chi * rtd - 80.) * (chi * rtd - 100.) < 0.)
      opflux = .5 * (fed + fen) + .5 * (fed - fen) * cos(pi * (chi * rtd - 80.) / 20.);
    if (fabs(rl) >= pi / 3.) opflux = opflux + ppolar;
  }
  double dvb;
  {  // :94-104 (df_2d of bx and of cs*by)
    double dbxdx = dflon1(in.bx, G, i, j);
    auto csby = [&](int jj) { return g.cs[jj] * in.by[G.a2(i, jj)]; };
    int J1 = G.nlatd - 1;
    double dcsbydy;
    if (j == 0) dcsbydy = csby(1) - csby(0);
    else if (j == J1) dcsbydy = csby(J1) - csby(J1 - 1);
    else dcsbydy = (csby(j + 1) - csby(j - 1)) / 2.;
    double vv = dbxdx / g.dlamda + dcsbydy / g.dphi;
    vv = vv / cs;
    dvb = (vv + 2. * bz) / RE;
  }
  const double dvbz = dvb / bz;
  const double gmr = GRAV * RMOP / (2. * GASK);
  struct OL { double scht, hdz, tp1, dj, bdotu, bdopj, wi; };
  OL B[2] = {}, A[2] = {}, Cn[2] = {};  // this lane's level, the one below and the one above
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    size_t p = G.a3(Lr[h] < nk ? Lr[h] : nk - 1, i, j);
    B[h].scht = in.scht[p];
    B[h].hdz = 1. / (B[h].scht * dz);
    B[h].tp1 = 0.5 * (in.ti[p] + in.te[p]);
    B[h].dj = W.dj[p]; B[h].bdotu = W.bdotu[p]; B[h].bdopj = W.bdopj[p]; B[h].wi = in.wi[p];
  }
#define TGX_NBR(fld)                                                          \
  wave_dn_h<NH>(B[0].fld, B[1].fld, lane, A[0].fld, A[1].fld);                      \
  wave_up_h<NH>(B[0].fld, B[1].fld, lane, Cn[0].fld, Cn[1].fld);
  TGX_NBR(hdz) TGX_NBR(tp1) TGX_NBR(dj) TGX_NBR(bdotu) TGX_NBR(bdopj) TGX_NBR(wi)
#undef TGX_NBR
  const double ol1 = oploss_p[G.a3(1, i, j)], qop0 = in.qop[c2];  // lower boundary (:301-303)
  TriRow lo[1] = {}, hi[1] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int k = Lr[h];
    TriRow t = tri_ident();
    if (k <= nk - 2) {
      size_t p = G.a3(k, i, j);
      const OL &qa = A[h], &qb = B[h], &qc = Cn[h];
      const bool top = k == nk - 2;
      double dbd;
      if (k == 0) dbd = qc.bdopj - qb.bdopj;
      else if (k <= nk - 3) dbd = (qc.bdopj - qa.bdopj) / 2.;
      else dbd = qb.bdopj - qa.bdopj;
      double bdzdvb = bz / (qb.scht * dz) * dbd + dvb * qb.bdopj;
      double bd_op = W.djbz[p] * OT.bd_diffj[p];
      double diffexp = bdzdvb + OT.bd_diff[p] + bd_op;
      double expl = -1. * diffexp;
      expl = expl + OT.vdotn[p] + OT.bd_bvel[p];
      expl = expl - OT.opts2[p] * dtx2inv * nsub;
      double divbz = dvbz + OT.bd_djbz[p] / (qb.dj * (bz * bz));
      double hm = qb.hdz - 0.5 * divbz, hp = qb.hdz + 0.5 * divbz;
      hm = hm * (bz * bz);
      hp = hp * (bz * bz);
      double hdzi_k = k == 0 ? 1.5 * qb.hdz - 0.5 * qc.hdz : 0.5 * (qa.hdz + qb.hdz);
      double hdzi_k1 = top ? 1.5 * qb.hdz - 0.5 * qa.hdz : 0.5 * (qb.hdz + qc.hdz);
      double dji = k == 0 ? 1.5 * qb.dj - 0.5 * qc.dj : 0.5 * (qa.dj + qb.dj);
      double dji1 = top ? 1.5 * qb.dj - 0.5 * qa.dj : 0.5 * (qb.dj + qc.dj);
      double t0 = k == 0 ? 2. * (2. * qb.tp1 - qc.tp1) * hdzi_k - gmr : 2. * qa.tp1 * hdzi_k - gmr;
      double t01 = 2. * qb.tp1 * hdzi_k1 - gmr;
      double t1 = 2. * qb.tp1 * hdzi_k + gmr;
      double t11 = top ? 2. * (2. * qb.tp1 - qa.tp1) * hdzi_k1 + gmr : 2. * qc.tp1 * hdzi_k1 + gmr;
      double diffp = hm * dji * t0;
      double diffq = -(hp * dji1 * t01 + hm * dji * t1);
      double diffr = hp * dji1 * t11;
      double hdzk = qb.hdz;
      double wii = .5 * (qb.wi + qc.wi);
      double driftp = wii * 0.5 * hdzk;
      double driftq = -wii * 6. / RE;
      double driftr = -wii * 0.5 * hdzk;
      double windp = k == 0 ? bz * (2. * qb.bdotu - qc.bdotu) * 0.5 * hdzk : bz * qa.bdotu * 0.5 * hdzk;
      double windq = -qb.bdotu * dvbz * bz;
      double windr = !top ? -bz * qc.bdotu * 0.5 * hdzk : bz * (2. * qb.bdotu - qa.bdotu) * 0.5 * hdzk;
      double pcf = diffp + driftp + windp;
      double qcf = diffq + driftq + windq - dtx2inv * nsub;
      double rcf = diffr + driftr + windr;
      if (top) {  // upper boundary (:280-286)
        double djn = dji1;
        double ubcb = -(bz * bz) * djn * t01;
        double ubca = -(bz * bz) * djn * t11;
        qcf = qcf + ubcb / ubca * rcf;
        expl = expl - opflux * rcf / ubca;
        rcf = 0.;
      }
      const double oploss = oploss_p[p];
      qcf = qcf - oploss;
      expl = expl - opprod_p[p];
      if (k == 0) {  // lower boundary (:301-303)
        qcf = qcf - pcf;
        expl = expl - 2. * pcf * qop0 / (1.5 * oploss - 0.5 * ol1);
        pcf = 0.;
      }
      t = TriRow{pcf, qcf, rcf, expl};
    }
    if (h == 0) lo[0] = t; else hi[0] = t;
  }
  if constexpr (NH == 2)
    wave_pcr<1>(lo, hi, lane);
  else
    wave_pcr1<1>(lo, lane);
  double x[2] = {lo[0].f / lo[0].b, NH == 2 ? hi[0].f / hi[0].b : 0.}, xm1[2], xm2[2];
  wave_dn_h<NH>(x[0], x[1], lane, xm1[0], xm1[1]);
  wave_dn_h<NH>(xm1[0], xm1[1], lane, xm2[0], xm2[1]);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L >= nk) continue;
    double xv = L == nk - 1 ? xm1[h] * xm1[h] / xm2[h] : x[h];
    size_t p = G.a3(L, i, j);
    double nm = DTSMOOTH * op[p] + DTSMOOTH_DIV2 * (optm1[p] + xv);
    opout[p] = xv > 0. ? xv : 0.;
    optm1out[p] = nm > 0. ? nm : 0.;
  }
}

// =============================================================================
// duv_ng.F90 stage 1 (every column): the vertical integral ztmp (:73-80)
__global__ void __launch_bounds__(256) duv_stage1(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                  const double *__restrict__ tn_upd,
                                                  double *__restrict__ ztmp) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const double dz = g.dz, wt = 0.225;
  size_t p = G.a3(0, i, j);
  double zt = in.z[p];
  ztmp[p] = zt;
  for (int k = 0; k < G.nk - 1; ++k) {
    p = G.a3(k, i, j);
    double tbar = (in.tn[p] * (-2.) + in.tn_nm[p] + tn_upd[p]) * wt + in.tn[p];
    double dztbar = dz / DZGRAV * tbar / in.mbar[p];
    zt = dztbar + zt;
    ztmp[p + G.P] = zt;
  }
}

// The same with one wavefront per column, for columns of at most 64 levels
// on small grids: lane k forms level k's increment, and the running sum is
// carried through the lanes bottom to top in the order of the loop above, so
// every partial sum is bitwise the same.
__global__ void __launch_bounds__(256) duv_stage1_wave(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                       const double *__restrict__ tn_upd,
                                                       double *__restrict__ ztmp) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk;
  const double dz = g.dz, wt = 0.225;
  double dzt = 0.;
  if (lane < nk - 1) {
    const size_t p = G.a3(lane, i, j);
    double tbar = (in.tn[p] * (-2.) + in.tn_nm[p] + tn_upd[p]) * wt + in.tn[p];
    dzt = dz / DZGRAV * tbar / in.mbar[p];
  }
  double zt = in.z[G.a2(i, j)];  // level 0
  double mine = zt;
  for (int k = 0; k < nk - 1; ++k) {
    zt = __shfl(dzt, k, 64) + zt;
    mine = lane == k + 1 ? zt : mine;
  }
  if (lane < nk) ztmp[G.a3(lane, i, j)] = mine;
}

// stage 2: assembly of the 2x2 (UN,VN) system, block Thomas, outputs (:83-321)
struct UvTerms {  // prelude planes
  const double *divx_ztmp, *divy_z, *unms, *vnms, *advu, *advv;
};

// duv_ng in one pass per column tile: the vertical integral ztmp of stage 1
// (:73-80), the horizontal stencils (divergence of ztmp and z for the
// pressure gradient, Shapiro smoothing of UN_NM/VN_NM, advection of UN/VN)
// and the 2x2 block elimination of stage 2, level by level.  A block covers a
// TILE_X x TILE_Y tile; at level k the tile and its 2-point halo of UN, VN,
// UN_NM, VN_NM (level k) and of z and ztmp (level k+1) are staged through
// double-buffered LDS.  ztmp is integrated upward for every tile and halo
// point by the thread that stages it (the halo points are integrated by each
// block that needs them, in the same arithmetic as stage 1).  Every thread of
// the block takes part in the staging and the barriers; lateral-boundary and
// out-of-domain columns do not solve.
template <bool CAP>
__global__ void __launch_bounds__(256) duv_tile_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                       tgx_ng_out out, double *__restrict__ gs) {
  constexpr int TS = TILE_R * TILE_W, NF = 6, TL = TILE_Y * TILE_W;
  enum { T_UN = 0, T_VN, T_UNM, T_VNM, T_Z, T_ZT };
  __shared__ double tb[2][NF][TS];
  __shared__ double tls[2][TL];  // latitude passes of UN_NM, VN_NM
  const int ntx = (G.nlond + TILE_X - 1) / TILE_X;
  const int bb = tgx_xcd_block(gridDim.x);
  const int i0 = (bb % ntx) * TILE_X, j0 = (bb / ntx) * TILE_Y;
  const int i = i0 + (int)threadIdx.x % TILE_X, j = j0 + (int)threadIdx.x / TILE_X;
  const bool active = i < G.nlond && j < G.nlatd;
  const bool lat = active && !G.glb && (j == 0 || j == G.nlatd - 1 || i == 0 || i == G.nlond - 1);
  const bool solve = active && !lat;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv, wt = 0.225;
  // tile slots of this thread: e = tid and tid + 256
  int te[2];
  bool tok[2];
  size_t toff[2];
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const int e = (int)threadIdx.x + sl * 256;
    const int ii = i0 - TILE_H + e % TILE_W, jj = j0 - TILE_H + e / TILE_W;
    te[sl] = e;
    tok[sl] = e < TS && ii >= 0 && ii < G.nlond && jj >= 0 && jj < G.nlatd;
    toff[sl] = tok[sl] ? (size_t)jj * G.ld + ii : 0;
  }
  // staged values of level k: UN, VN, UN_NM, VN_NM at k, z at k+1, and the
  // stage-1 inputs at k from which ztmp(k+1) follows
  double st[2][NF], zt[2] = {0., 0.};
  auto tile_load = [&](int k) {
    const size_t o = (size_t)k * G.P, o1 = o + G.P;
#pragma unroll
    for (int sl = 0; sl < 2; ++sl)
      if (tok[sl]) {
        const size_t q = o + toff[sl];
        st[sl][T_UN] = in.un[q]; st[sl][T_VN] = in.vn[q];
        st[sl][T_UNM] = in.un_nm[q]; st[sl][T_VNM] = in.vn_nm[q];
        st[sl][T_Z] = in.z[o1 + toff[sl]];
        // duv_stage1 at this point: ztmp(k+1) = ztmp(k) + dz/dzgrav*tbar/mbar
        const double tn = in.tn[q];
        const double tbar = (tn * (-2.) + in.tn_nm[q] + out.tn[q]) * wt + tn;
        const double dztbar = dz / DZGRAV * tbar / in.mbar[q];
        st[sl][T_ZT] = dztbar + zt[sl];
      }
  };
  auto tile_store = [&](int buf) {
#pragma unroll
    for (int sl = 0; sl < 2; ++sl)
      if (tok[sl]) {
#pragma unroll
        for (int f = 0; f < NF; ++f) tb[buf][f][te[sl]] = st[sl][f];
        zt[sl] = st[sl][T_ZT];
      }
  };
  const int ic = active ? i : 0, jc = active ? j : 0;
  const size_t c2 = G.a2(ic, jc);
  if (active) {
    out.ulbc_nm[c2] = in.ulbc[c2];
    out.vlbc_nm[c2] = in.vlbc[c2];
    out.ulbc[c2] = in.u_lbc[c2];
    out.vlbc[c2] = in.v_lbc[c2];
  }
  if (lat) {
    double vu, vv;
    for (int k = 0; k < nk; ++k) {
      size_t p = G.a3(k, i, j);
      lateral_value(G, in.lonb_un, in.latb_un, k, i, j, vu);
      lateral_value(G, in.lonb_vn, in.latb_vn, k, i, j, vv);
      out.un[p] = vu;
      out.vn[p] = vv;
      out.un_nm[p] = DTSMOOTH_DIV2 * (in.un_nm[p] + vu) + DTSMOOTH * in.un[p];
      out.vn_nm[p] = DTSMOOTH_DIV2 * (in.vn_nm[p] + vv) + DTSMOOTH * in.vn[p];
    }
  }
#define F3(f, k) (f)[G.a3((k), ic, jc)]
  const double racs = g.racs[jc], cor = g.cor[jc], tanphi = g.tanphi[jc];
  const LdsPlane P0z{tb[1][T_Z], i0 - TILE_H, j0 - TILE_H};
  // level 0 of z (= ztmp(0)) for the divergences at level 0
#pragma unroll
  for (int sl = 0; sl < 2; ++sl)
    if (tok[sl]) {
      const double z0 = in.z[toff[sl]];
      tb[1][T_Z][te[sl]] = z0;
      zt[sl] = z0;
    }
  tile_load(0);
  __syncthreads();
  double divx_c = 0., divy_c = 0.;
  if (solve) {
    divx_c = divx(P0z, G, i, j);
    divy_c = divy(P0z, G, i, j);
  }
  tile_store(0);  // buffer 0; buffer 1 is read above before the next barrier
  // fields of the column at one level (the divergences come from the tiles)
  struct UL { double w, lxx, lyy, lxy, lyx, ui, vi, km, schti, scht; };
  struct US { double Fe, Fn, hdu, hdv; };
  auto load_ul = [&](int L, UL &U) {
    size_t p = G.a3(L, ic, jc);
    U.w = in.w_upd[p]; U.lxx = in.lxx[p]; U.lyy = in.lyy[p]; U.lxy = in.lxy[p]; U.lyx = in.lyx[p];
    U.ui = in.ui[p]; U.vi = in.vi[p]; U.km = in.km[p]; U.schti = in.schti[p]; U.scht = in.scht[p];
  };
  auto load_us = [&](int L, US &S) {
    size_t p = G.a3(L, ic, jc);
    S.Fe = out.Fe[p]; S.Fn = out.Fn[p]; S.hdu = in.hdu[p]; S.hdv = in.hdv[p];
  };
  auto e1b = [&](double scht, int L) -> double {
    double e = P0 * scht / GRAV;
    return e * g.expz[L] * g.xmue[L];
  };
  auto gg = [&](const UL &U, double evf) { return GRAV * (U.km + evf) / (P0 * U.schti * (dz * dz)); };
  const double ulbc = solve ? in.u_lbc[c2] : 0., vlbc = solve ? in.v_lbc[c2] : 0.;
  UL ua = {}, ub = {};
  double e1_m = 0., e1_c = 0., gk = 0.;
  if (solve) {
    load_ul(0, ua);
    load_ul(1, ub);
    e1_c = e1b(ua.scht, 0);
    const double e1_1 = e1b(ub.scht, 1);
    gk = gg(ua, e1_c * e1_c / sqrt(e1_c * e1_1));
  }
  double be11 = 0, be12 = 0, be21 = 0, be22 = 0, y1 = 0, y2 = 0;
  double r11p = 0, r12p = 0, r21p = 0, r22p = 0;
  for (int k = 0; k <= nk - 2; ++k) {
    US us = {};
    if (k + 1 <= nk - 2) tile_load(k + 1);
    if (solve) {
      if (k > 0) load_ul(k + 1, ub);
      load_us(k, us);
    }
    asm volatile("" ::: "memory");
    __syncthreads();  // the level-k tiles are in buffer k&1
    const int bk = k & 1;
    const LdsPlane Tun{tb[bk][T_UN], i0 - TILE_H, j0 - TILE_H}, Tvn{tb[bk][T_VN], i0 - TILE_H, j0 - TILE_H},
        Tunm{tb[bk][T_UNM], i0 - TILE_H, j0 - TILE_H}, Tvnm{tb[bk][T_VNM], i0 - TILE_H, j0 - TILE_H},
        Tz{tb[bk][T_Z], i0 - TILE_H, j0 - TILE_H}, Tzt{tb[bk][T_ZT], i0 - TILE_H, j0 - TILE_H};
    // smooth_ng.F90:283-294: latitude passes on the tile rows and their longitude halo
    for (int e = threadIdx.x; e < TL; e += blockDim.x) {
      int ii = i0 - TILE_H + e % TILE_W, jj = j0 + e / TILE_W;
      if (ii >= 0 && ii < G.nlond && jj < G.nlatd) {
        tls[0][e] = smooth_lat(Tunm, G, ii, jj, g.shapiro);
        tls[1][e] = smooth_lat(Tvnm, G, ii, jj, g.shapiro);
      }
    }
    __syncthreads();
    if (solve) {
      // :296-307 longitude passes, advec_ng of UN and VN, and the divergences at k+1
      double sm[2];
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const double *row = tls[f] + (j - j0) * TILE_W + TILE_H - i0;
        const double s_ = g.shapiro;
        int L1 = G.nlond - 1;
        bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
        int i1p = i < L1 ? i + 1 : L1, i1m = i > 0 ? i - 1 : 0;
        int i2p = i + 2 <= L1 ? i + 2 : L1, i2m = i >= 2 ? i - 2 : 0;
        double t0 = row[i];
        double fm1 = (row[i1p] + row[i1m]) * 0.5;
        double fm2 = (row[i2p] + row[i2m]) * 0.5;
        double gen = t0 - s_ * (2. * fm2 - 8. * fm1 + 6. * t0);
        double nr = t0 - s_ * (-8. * fm1 + 8. * t0);
        sm[f] = edge ? t0 : (near ? nr : gen);
      }
      const double unms = sm[0], vnms = sm[1];
      const double advu = advec(Tun, Tun, Tvn, G, i, j, racs, g.dlamda, g.dphi);
      const double advv = advec(Tvn, Tun, Tvn, G, i, j, racs, g.dlamda, g.dphi);
      const double divx_n = divx(Tzt, G, i, j), divy_n = divy(Tz, G, i, j);
      const double un_k = pv(Tun, G, i, j);
      size_t p = G.a3(k, i, j);
      double e1_n = k + 1 <= nk - 2 ? e1b(ub.scht, k + 1) : sqrt(e1_c * e1_c * e1_c / e1_m);
      double gk1 = gg(ub, sqrt(e1_c * e1_n));
      double g_1 = gk1;
      const double expzk = g.expz[k];
      double wi = .5 * (ua.w + ub.w);
      double dwdz = wi / (2. * dz);
      dwdz = dwdz * expzk;
      double lxxi = .5 * (ua.lxx + ub.lxx), lyyi = .5 * (ua.lyy + ub.lyy);
      double lxyi = .5 * (ua.lxy + ub.lxy), lyxi = .5 * (ua.lyx + ub.lyx);
      double uii = .5 * (ua.ui + ub.ui), vii = .5 * (ua.vi + ub.vi);
      double udc = us.Fe + lxxi * uii + lxyi * vii;
      double vdc = us.Fn + lyyi * vii - lyxi * uii;
      double zl = (divx_c + divx_n) * .5;
      zl = zl * racs;
      zl = zl * GRAV / g.dlamda;
      double zp = (divy_c + divy_n) * .5;
      zp = zp * GRAV * RE_INV / g.dphi;
      double s1 = dtx2inv * unms + us.hdu + udc - advu - zl;
      double s2 = dtx2inv * vnms + us.hdv + vdc - advv - zp;
      s1 = s1 * expzk;
      s2 = s2 * expzk;
      double p11 = -(gk + dwdz), p12 = 0., p21 = 0., p22 = p11;
      double r11 = -(g_1 - dwdz), r12 = 0., r21 = 0., r22 = r11;
      double q11 = gk + g_1;
      double qa = dtx2inv + lxxi;
      qa = qa * expzk;
      double qq11 = q11 + qa;
      qa = dtx2inv + lyyi;
      qa = qa * expzk;
      double qq22 = q11 + qa;
      double q12 = cor + un_k / RE * tanphi;
      double q21 = q12;
      qa = q12 - lxyi;
      qa = qa * expzk;
      double qq12 = -qa;
      qa = q21 - lyxi;
      qa = qa * expzk;
      double qq21 = qa;
      if (k == 0) {  // :207-212
        qq11 = qq11 - p11; qq12 = qq12 - p12; qq21 = qq21 - p21; qq22 = qq22 - p22;
        s1 = s1 - 2. * (p11 * ulbc + p12 * vlbc);
        s2 = s2 - 2. * (p21 * ulbc + p22 * vlbc);
        p11 = 0.; p12 = 0.; p21 = 0.; p22 = 0.;
      }
      if (k == nk - 2) {  // :214-215
        qq11 = qq11 + r11; qq12 = qq12 + r12; qq21 = qq21 + r21; qq22 = qq22 + r22;
        r11 = 0.; r12 = 0.; r21 = 0.; r22 = 0.;
      }
      if (CAP) {
        const double Ab[4] = {p11, p12, p21, p22}, Bb[4] = {qq11, qq12, qq21, qq22}, Cb[4] = {r11, r12, r21, r22};
        for (int e = 0; e < 4; ++e) {
          cap_put(g, G, CAP_UV + e, i, j, k, Ab[e]);
          cap_put(g, G, CAP_UV + 4 + e, i, j, k, Bb[e]);
          cap_put(g, G, CAP_UV + 8 + e, i, j, k, Cb[e]);
        }
        cap_put(g, G, CAP_UV + 12, i, j, k, s1);
        cap_put(g, G, CAP_UV + 13, i, j, k, s2);
      }
      if (k == 0) {
        double det = qq11 * qq22 - qq12 * qq21;
        be11 = qq22 / det; be12 = -qq12 / det; be21 = -qq21 / det; be22 = qq11 / det;
        y1 = be11 * s1 + be12 * s2;
        y2 = be21 * s1 + be22 * s2;
      } else {
        double g11 = be11 * r11p + be12 * r21p;
        double g12 = be11 * r12p + be12 * r22p;
        double g21 = be21 * r11p + be22 * r21p;
        double g22 = be21 * r12p + be22 * r22p;
        const size_t tm = G.wtile(k - 1, 0, UW_NV);
        gs[tm] = g11; gs[tm + 64] = g21; gs[tm + 128] = g12; gs[tm + 192] = g22;
        double h11 = qq11 - p11 * g11 - p12 * g21;
        double h12 = qq12 - p11 * g12 - p12 * g22;
        double h21 = qq21 - p21 * g11 - p22 * g21;
        double h22 = qq22 - p21 * g12 - p22 * g22;
        double dk = h11 * h22 - h12 * h21;
        be11 = h22 / dk; be12 = -h12 / dk; be21 = -h21 / dk; be22 = h11 / dk;
        double x1 = s1 - p11 * y1 - p12 * y2;
        double x2 = s2 - p21 * y1 - p22 * y2;
        y1 = be11 * x1 + be12 * x2;
        y2 = be21 * x1 + be22 * x2;
      }
      const size_t ty = G.wtile(k, 4, UW_NV);  // forward solution
      gs[ty] = y1;
      gs[ty + 64] = y2;
      r11p = r11; r12p = r12; r21p = r21; r22p = r22;
      gk = gk1;
      e1_m = e1_c;
      e1_c = e1_n;
      ua = ub;
      divx_c = divx_n;
      divy_c = divy_n;
    }
    if (k + 1 <= nk - 2) tile_store((k + 1) & 1);
  }
  if (!solve) return;
  // back substitution (:244-250) and outputs (:320-321)
  struct UB { double x1, x2, g11, g21, g12, g22, unm, un, vnm, vn; };
  auto load_ub = [&](int k, UB &B) {
    size_t p = G.a3(k, i, j);
    const size_t t = G.wtile(k, 0, UW_NV);
    B.x1 = gs[t + 256]; B.x2 = gs[t + 320];
    if (k < nk - 2) { B.g11 = gs[t]; B.g21 = gs[t + 64]; B.g12 = gs[t + 128]; B.g22 = gs[t + 192]; }
    else { B.g11 = B.g21 = B.g12 = B.g22 = 0.; }
    B.unm = in.un_nm[p]; B.un = in.un[p]; B.vnm = in.vn_nm[p]; B.vn = in.vn[p];
  };
  const double unm_top = F3(in.un_nm, nk - 1), un_top = F3(in.un, nk - 1);
  const double vnm_top = F3(in.vn_nm, nk - 1), vn_top = F3(in.vn, nk - 1);
  double xn1 = y1, xn2 = y2;
  UB bc, bn;
  load_ub(nk - 2, bc);
  for (int k = nk - 2; k >= 0; --k) {
    if (k > 0) load_ub(k - 1, bn);
    asm volatile("" ::: "memory");
    double x1, x2;
    if (k == nk - 2) { x1 = xn1; x2 = xn2; }
    else {
      x1 = bc.x1 - bc.g11 * xn1 - bc.g12 * xn2;
      x2 = bc.x2 - bc.g21 * xn1 - bc.g22 * xn2;
    }
    if (G.raw_row(j)) {  // the raw solution; filter_uv and the post pass follow (duv.F:490-514)
      const size_t p = G.a3(k, i, j);
      out.un[p] = x1;
      out.vn[p] = x2;
      xn1 = x1;
      xn2 = x2;
      bc = bn;
      continue;
    }
    if (G.glb && k == nk - 2) {  // spval at the top level (duv.F:385-386, :512-513)
      size_t pt = G.a3(nk - 1, i, j);
      out.un[pt] = SPVAL;
      out.vn[pt] = SPVAL;
      out.un_nm[pt] = SPVAL;
      out.vn_nm[pt] = SPVAL;
    }
    if (!G.glb && k == nk - 3) {
      size_t pt = G.a3(nk - 1, i, j);
      double u = 2 * xn1 - x1, vv = 2 * xn2 - x2;
      out.un[pt] = u;
      out.vn[pt] = vv;
      out.un_nm[pt] = DTSMOOTH_DIV2 * (unm_top + u) + DTSMOOTH * un_top;
      out.vn_nm[pt] = DTSMOOTH_DIV2 * (vnm_top + vv) + DTSMOOTH * vn_top;
    }
    size_t p = G.a3(k, i, j);
    out.un[p] = x1;
    out.vn[p] = x2;
    out.un_nm[p] = DTSMOOTH_DIV2 * (bc.unm + x1) + DTSMOOTH * bc.un;
    out.vn_nm[p] = DTSMOOTH_DIV2 * (bc.vnm + x2) + DTSMOOTH * bc.vn;
    xn1 = x1; xn2 = x2;
    bc = bn;
  }
#undef F3
}

// duv_ng stage 2 with one wavefront per column: the 2x2 block rows 0..nk-2 by
// block cyclic reduction, the top level extrapolated (:250).
template <int NH>
__global__ void __launch_bounds__(256) duv_wave_stage2(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                       tgx_ng_out out, UvTerms UT) {
  int i, j;
  if (!wave_column_of(G, i, j)) return;
  const int lane = threadIdx.x & 63, nk = G.nk;
  const int Lr[2] = {lane, lane + 64};
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
  if (lane == 0) {
    out.ulbc_nm[c2] = in.ulbc[c2];
    out.vlbc_nm[c2] = in.vlbc[c2];
    out.ulbc[c2] = in.u_lbc[c2];
    out.vlbc[c2] = in.v_lbc[c2];
  }
  double v;
  if (lateral_value(G, in.lonb_un, in.latb_un, 0, i, j, v)) {
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int L = Lr[h];
      if (L >= nk) continue;
      size_t p = G.a3(L, i, j);
      double vu, vv;
      lateral_value(G, in.lonb_un, in.latb_un, L, i, j, vu);
      lateral_value(G, in.lonb_vn, in.latb_vn, L, i, j, vv);
      out.un[p] = vu;
      out.vn[p] = vv;
      out.un_nm[p] = DTSMOOTH_DIV2 * (in.un_nm[p] + vu) + DTSMOOTH * in.un[p];
      out.vn_nm[p] = DTSMOOTH_DIV2 * (in.vn_nm[p] + vv) + DTSMOOTH * in.vn[p];
    }
    return;
  }
  const double racs = g.racs[j], cor = g.cor[j], tanphi = g.tanphi[j];
  struct UL { double w, lxx, lyy, lxy, lyx, ui, vi, divx, divy, km, schti; };
  UL U[2] = {}, Un[2] = {};
  double e1r[2];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int Lc = Lr[h] < nk ? Lr[h] : nk - 1;
    size_t p = G.a3(Lc, i, j);
    U[h].w = in.w_upd[p]; U[h].lxx = in.lxx[p]; U[h].lyy = in.lyy[p]; U[h].lxy = in.lxy[p];
    U[h].lyx = in.lyx[p]; U[h].ui = in.ui[p]; U[h].vi = in.vi[p]; U[h].divx = UT.divx_ztmp[p];
    U[h].divy = UT.divy_z[p]; U[h].km = in.km[p]; U[h].schti = in.schti[p];
    double e = P0 * in.scht[p] / GRAV;  // eddy viscosity (:143-152)
    e1r[h] = e * g.expz[Lc] * g.xmue[Lc];
  }
#define TGX_UP(fld) wave_up_h<NH>(U[0].fld, U[1].fld, lane, Un[0].fld, Un[1].fld);
  TGX_UP(w) TGX_UP(lxx) TGX_UP(lyy) TGX_UP(lxy) TGX_UP(lyx) TGX_UP(ui) TGX_UP(vi) TGX_UP(divx) TGX_UP(divy)
#undef TGX_UP
  // e1 at the top level from the two below (:148-150)
  double e1m[2], e1mm[2], e1[2];
  wave_dn_h<NH>(e1r[0], e1r[1], lane, e1m[0], e1m[1]);
  wave_dn_h<NH>(e1m[0], e1m[1], lane, e1mm[0], e1mm[1]);
#pragma unroll
  for (int h = 0; h < NH; ++h) e1[h] = Lr[h] == nk - 1 ? sqrt(e1m[h] * e1m[h] * e1m[h] / e1mm[h]) : e1r[h];
  double e1d[2], e1u[2];
  wave_dn_h<NH>(e1[0], e1[1], lane, e1d[0], e1d[1]);
  wave_up_h<NH>(e1[0], e1[1], lane, e1u[0], e1u[1]);
  double gg[2], ggp[2];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    double evf = Lr[h] >= 1 ? sqrt(e1d[h] * e1[h]) : e1[h] * e1[h] / sqrt(e1[h] * e1u[h]);
    gg[h] = GRAV * (U[h].km + evf) / (P0 * U[h].schti * (dz * dz));
  }
  wave_up_h<NH>(gg[0], gg[1], lane, ggp[0], ggp[1]);
  const double ulbc = in.u_lbc[c2], vlbc = in.v_lbc[c2];
  BlkRow<2> rw[2] = {};
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int k = Lr[h];
    BlkRow<2> &R = rw[h];
    if (k > nk - 2) {
      R.A[0] = R.A[1] = R.A[2] = R.A[3] = 0.;
      R.C[0] = R.C[1] = R.C[2] = R.C[3] = 0.;
      R.B[0] = 1.; R.B[1] = 0.; R.B[2] = 0.; R.B[3] = 1.;
      R.F[0] = R.F[1] = 0.;
      continue;
    }
    size_t p = G.a3(k, i, j);
    const double expzk = g.expz[k];
    const double gk = gg[h], g_1 = ggp[h];
    double wi = .5 * (U[h].w + Un[h].w);
    double dwdz = wi / (2. * dz);
    dwdz = dwdz * expzk;
    double lxxi = .5 * (U[h].lxx + Un[h].lxx), lyyi = .5 * (U[h].lyy + Un[h].lyy);
    double lxyi = .5 * (U[h].lxy + Un[h].lxy), lyxi = .5 * (U[h].lyx + Un[h].lyx);
    double uii = .5 * (U[h].ui + Un[h].ui), vii = .5 * (U[h].vi + Un[h].vi);
    double udc = out.Fe[p] + lxxi * uii + lxyi * vii;
    double vdc = out.Fn[p] + lyyi * vii - lyxi * uii;
    double zl = (U[h].divx + Un[h].divx) * .5;
    zl = zl * racs;
    zl = zl * GRAV / g.dlamda;
    double zp = (U[h].divy + Un[h].divy) * .5;
    zp = zp * GRAV * RE_INV / g.dphi;
    double s1 = dtx2inv * UT.unms[p] + in.hdu[p] + udc - UT.advu[p] - zl;
    double s2 = dtx2inv * UT.vnms[p] + in.hdv[p] + vdc - UT.advv[p] - zp;
    s1 = s1 * expzk;
    s2 = s2 * expzk;
    double p11 = -(gk + dwdz), p12 = 0., p21 = 0., p22 = p11;
    double r11 = -(g_1 - dwdz), r12 = 0., r21 = 0., r22 = r11;
    double q11 = gk + g_1;
    double qa = dtx2inv + lxxi;
    qa = qa * expzk;
    double qq11 = q11 + qa;
    qa = dtx2inv + lyyi;
    qa = qa * expzk;
    double qq22 = q11 + qa;
    double q12 = cor + in.un[p] / RE * tanphi;
    double q21 = q12;
    qa = q12 - lxyi;
    qa = qa * expzk;
    double qq12 = -qa;
    qa = q21 - lyxi;
    qa = qa * expzk;
    double qq21 = qa;
    if (k == 0) {  // :207-212
      qq11 = qq11 - p11; qq12 = qq12 - p12; qq21 = qq21 - p21; qq22 = qq22 - p22;
      s1 = s1 - 2. * (p11 * ulbc + p12 * vlbc);
      s2 = s2 - 2. * (p21 * ulbc + p22 * vlbc);
      p11 = 0.; p12 = 0.; p21 = 0.; p22 = 0.;
    }
    if (k == nk - 2) {  // :214-215
      qq11 = qq11 + r11; qq12 = qq12 + r12; qq21 = qq21 + r21; qq22 = qq22 + r22;
      r11 = 0.; r12 = 0.; r21 = 0.; r22 = 0.;
    }
    R.A[0] = p11; R.A[1] = p12; R.A[2] = p21; R.A[3] = p22;
    R.B[0] = qq11; R.B[1] = qq12; R.B[2] = qq21; R.B[3] = qq22;
    R.C[0] = r11; R.C[1] = r12; R.C[2] = r21; R.C[3] = r22;
    R.F[0] = s1; R.F[1] = s2;
  }
  double xlo[2] = {}, xhi[2] = {};
  if constexpr (NH == 2)
    wave_block_pcr<2>(rw[0], rw[1], lane, xlo, xhi);
  else
    wave_block_pcr1<2>(rw[0], lane, xlo);
  double x1[2] = {xlo[0], xhi[0]}, x2[2] = {xlo[1], xhi[1]};
  double x1m[2], x1mm[2], x2m[2], x2mm[2];
  wave_dn_h<NH>(x1[0], x1[1], lane, x1m[0], x1m[1]);
  wave_dn_h<NH>(x1m[0], x1m[1], lane, x1mm[0], x1mm[1]);
  wave_dn_h<NH>(x2[0], x2[1], lane, x2m[0], x2m[1]);
  wave_dn_h<NH>(x2m[0], x2m[1], lane, x2mm[0], x2mm[1]);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int L = Lr[h];
    if (L >= nk) continue;
    double u = x1[h], vv = x2[h];
    if (L == nk - 1) { u = 2 * x1m[h] - x1mm[h]; vv = 2 * x2m[h] - x2mm[h]; }  // :250
    size_t p = G.a3(L, i, j);
    out.un[p] = u;
    out.vn[p] = vv;
    out.un_nm[p] = DTSMOOTH_DIV2 * (in.un_nm[p] + u) + DTSMOOTH * in.un[p];
    out.vn_nm[p] = DTSMOOTH_DIV2 * (in.vn_nm[p] + vv) + DTSMOOTH * in.vn[p];
  }
}

}  // namespace ng
}  // namespace tgx

// ---------------------------------------------------------------------------
// Launcher used by step_api.cpp.  bufs holds the workspace planes (and for
// O+ the input/output state of the sub-step) in the order documented there.
// Stencil prelude launch: a run of levels of one tile per block.
template <class KS, int COMB = tgx::ng::STC_NONE>
static void launch_stencil_run(const tgx_ng_grid &g, const tgx::ng::Geo &G, const tgx_ng_fields &in,
                               const tgx::ng::StencilJobs &J, hipStream_t s) {
  using namespace tgx::ng;
  const int ntx = (G.nlond + TILE_X - 1) / TILE_X, nty = (G.nlatd + TILE_Y - 1) / TILE_Y;
  const int nrun = (G.nk + STENCIL_RUN - 1) / STENCIL_RUN;
  hipLaunchKernelGGL((stencil_run_kernel<KS, COMB>), dim3((unsigned)(ntx * nty * nrun)), dim3(256), KS::lds_bytes(),
                     s, g, G, in, J);
}


extern "C" int tgx_ng_launch(int which, const tgx_ng_grid *gp, const tgx_ng_fields *inp,
                             const tgx_ng_out *outp, void *const *bufs, int sub, hipStream_t s) {
  using namespace tgx::ng;
  const tgx_ng_grid g = *gp;
  const tgx_ng_fields in = *inp;
  const tgx_ng_out out = *outp;
  Geo G = make_geo(g);
  long ncol = (long)G.nlond * G.nlatd;
  int nb = (int)((ncol + 255) / 256);
  int nbp = (int)((ncol * G.nk + 255) / 256);
  // thread per column (grid), per column-level (gridp), wavefront per column
  // (gridw)
  const int wcols = t_wave_cols;
  int nbw = (int)((ncol + wcols - 1) / wcols);
  dim3 grid(nb), gridp(nbp), gridw(nbw), block(256), blockw(64 * wcols);
  auto D = [&](int n) { return (double *)bufs[n]; };
  auto job = [](int kind, const double *f, double *o, double sh) {
    StencilJob jb;
    jb.kind = kind; jb.f = f; jb.out = o; jb.s = sh;
    return jb;
  };
  switch (which) {
    case 20:
      if (G.nk <= 64)
        hipLaunchKernelGGL(dt_wave_kernel<1>, gridw, blockw, 0, s, g, G, in, out, (const double *)D(1),
                           (const double *)D(2));
      else
        hipLaunchKernelGGL(dt_wave_kernel<2>, gridw, blockw, 0, s, g, G, in, out, (const double *)D(1),
                           (const double *)D(2));
      break;
    case 25: {  // bufs = {comp input tiles, fs(:,1,0), fs(:,2,0), (unused x3), minor source tiles,
                //         (unused x2), linear-coefficient tiles, (unused x5), oploss, opprod}
      ChemOut C;
      C.ci = D(0);
      C.fs0[0] = D(1);
      C.fs0[1] = D(2);
      C.st = D(6);
      C.lt = D(9);
      C.oploss = D(15); C.opprod = D(16);
      if ((unsigned long long)G.nk * G.P * sizeof(double) >= (1ull << 32))
        return tgx_set_error(1, "chem_kernel: a field of %d levels x %ld points exceeds 4 GiB", G.nk, (long)G.P);
      const int nrun = small_grid(ncol) ? (G.nk + 1) / 2 : chem_runs(G.nk);
      long nt = (long)G.nlond * G.nlatd * nrun;
      dim3 gridc((unsigned)((nt + 255) / 256));
      hipLaunchKernelGGL(chem_kernel, gridc, block, 0, s, g, G, in, out, C, nrun);
      break;
    }
    case 28: {  // duv with stage 1 and the stencils fused: bufs = {ztmp (unused), Thomas tiles (UW_NV)}
      dim3 gridt((unsigned)(((G.nlond + TILE_X - 1) / TILE_X) * ((G.nlatd + TILE_Y - 1) / TILE_Y)));
      if (g.capture)
        hipLaunchKernelGGL(duv_tile_kernel<true>, gridt, block, 0, s, g, G, in, out, D(1));
      else
        hipLaunchKernelGGL(duv_tile_kernel<false>, gridt, block, 0, s, g, G, in, out, D(1));
      break;
    }
    case 24: {  // dt with the stencils fused, one block per column tile: bufs = {Thomas tiles (DW_NV)}
      dim3 gridt((unsigned)(((G.nlond + TILE_X - 1) / TILE_X) * ((G.nlatd + TILE_Y - 1) / TILE_Y)));
      // the tiles by direct-to-LDS loads where their fields are 16-byte aligned
      // and the row pitch is even
      const bool dma = g.ld % 2 == 0 && ((uintptr_t)in.tn | (uintptr_t)in.un | (uintptr_t)in.vn |
                                         (uintptr_t)in.tn_nm) % 16 == 0 && t_tile_dma;
      if (g.capture)
        hipLaunchKernelGGL((dt_tile_kernel<true, false>), gridt, block, 0, s, g, G, in, out, D(0));
      else if (dma)
        hipLaunchKernelGGL((dt_tile_kernel<false, true>), gridt, block, 0, s, g, G, in, out, D(0));
      else
        hipLaunchKernelGGL((dt_tile_kernel<false, false>), gridt, block, 0, s, g, G, in, out, D(0));
      break;
    }
    case 1:
    case 2:
    case 15:
    case 16:
    case 18:
    case 19:
    case 27:
    case 30: {  // minors: bufs = {ratio tiles (MW_NV), right-hand-side tiles (MR_NV), (unused x8),
                // source tiles (MS_NV), (unused x3)}; N2D's source values are reused by NO after
                // N2D's solve
      MinorTerms MT{};
      MT.rt = D(1);
      MT.st = D(10);
      if (which == 1 && g.capture)
        hipLaunchKernelGGL((minor_kernel<2, 3, true>), grid, block, 0, s, g, G, in, out, D(0), MT);
      else if (which == 1)
        hipLaunchKernelGGL((minor_kernel<2, 3, false>), grid, block, 0, s, g, G, in, out, D(0), MT);
      else if (which == 27) {  // buf 14: chem_kernel's linear-coefficient tiles
        MT.lt = D(14);
        MT.n2d = out.n2d;
        if (g.capture)
          hipLaunchKernelGGL((minor_kernel<1, 0, true, true>), grid, block, 0, s, g, G, in, out, D(0), MT);
        else
          hipLaunchKernelGGL((minor_kernel<1, 0, false, true>), grid, block, 0, s, g, G, in, out, D(0), MT);
      } else if (which == 2 && g.capture)
        hipLaunchKernelGGL((minor_kernel<1, 0, true>), grid, block, 0, s, g, G, in, out, D(0), MT);
      else if (which == 2)
        hipLaunchKernelGGL((minor_kernel<1, 0, false>), grid, block, 0, s, g, G, in, out, D(0), MT);
      else if (which == 18 && G.nk <= 64)
        hipLaunchKernelGGL((minor_wave_kernel<2, 3, 1>), gridw, blockw, 0, s, g, G, in, out, MT);
      else if (which == 18)
        hipLaunchKernelGGL((minor_wave_kernel<2, 3, 2>), gridw, blockw, 0, s, g, G, in, out, MT);
      else if (which == 30) {  // buf 14: chem_kernel's linear-coefficient tiles
        MT.lt = D(14);
        MT.n2d = out.n2d;
        if (G.nk <= 64)
          hipLaunchKernelGGL((minor_wave_kernel<1, 0, 1, true>), gridw, blockw, 0, s, g, G, in, out, MT);
        else
          hipLaunchKernelGGL((minor_wave_kernel<1, 0, 2, true>), gridw, blockw, 0, s, g, G, in, out, MT);
      } else if (which == 19 && G.nk <= 64)
        hipLaunchKernelGGL((minor_wave_kernel<1, 0, 1>), gridw, blockw, 0, s, g, G, in, out, MT);
      else if (which == 19)
        hipLaunchKernelGGL((minor_wave_kernel<1, 0, 2>), gridw, blockw, 0, s, g, G, in, out, MT);
      else if (which == 15)
        hipLaunchKernelGGL((minor_src_kernel<2, 3>), gridp, block, 0, s, g, G, in, out, MT);
      else
        hipLaunchKernelGGL((minor_src_kernel<1, 0>), gridp, block, 0, s, g, G, in, out, MT);
      break;
    }
    case 3:
    case 4:
    case 9:
    case 17:
    case 21:
    case 29:
    case 31:
    case 32:
    case 33: {  // O+: bufs = {dj bdotu bvel tpop diffj djbz opbm2 bdopj w2, op, optm1, opout,
                //           optm1out, bd_diff bd_diffj bd_bvel vdotn bd_djbz opts2, oploss, opprod}
      OpWork W{D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), which >= 31 ? D(21) : nullptr};
      const double *op = D(9), *optm1 = D(10);
      double *opout = D(11), *optm1out = D(12);
      OpTerms OT{D(13), D(14), D(15), D(16), D(17), D(18)};
      if (which == 17) {
        hipLaunchKernelGGL(oplus_chem_kernel, gridp, block, 0, s, g, G, in, D(19), D(20), out.xiop2p,
                           out.xiop2d);
      } else if (which == 3 || which == 31) {  // 31: the O+-dependent values only
        const int nrun = oplus_runs(ncol, G.nk, 1);
        dim3 grid1((unsigned)((ncol * nrun + 255) / 256));
        if (which == 31) {
          if (nrun > 1)
            hipLaunchKernelGGL((oplus_stage1<true, false>), grid1, block, 0, s, g, G, in, op, W, out.Fe, out.Fn, nrun);
          else
            hipLaunchKernelGGL((oplus_stage1<false, false>), grid1, block, 0, s, g, G, in, op, W, out.Fe, out.Fn, 1);
        } else if (nrun > 1)
          hipLaunchKernelGGL((oplus_stage1<true, true>), grid1, block, 0, s, g, G, in, op, W, out.Fe, out.Fn, nrun);
        else
          hipLaunchKernelGGL((oplus_stage1<false, true>), grid1, block, 0, s, g, G, in, op, W, out.Fe, out.Fn, 1);
      } else if (which == 4) {
        const int nrun = oplus_runs(ncol, G.nk, 2);
        dim3 grid2((unsigned)((ncol * nrun + 255) / 256));
        if (nrun > 1)
          hipLaunchKernelGGL(oplus_stage2<true>, grid2, block, 0, s, g, G, in, W, nrun);
        else
          hipLaunchKernelGGL(oplus_stage2<false>, grid2, block, 0, s, g, G, in, W, 1);
      } else if (which == 9) {
        StencilJobs J{};
        J.n = 6;
        J.job[0] = job(ST_BDOTDH, W.bdopj, D(13), 0.);
        J.job[1] = job(ST_BDOTDH, W.diffj, D(14), 0.);
        J.job[2] = job(ST_BDOTDH, W.bvel, D(15), 0.);
        J.job[3] = job(ST_VDOTN, W.opbm2, D(16), 0.);
        J.job[4] = job(ST_BDOTDH, W.djbz, D(17), 0.);
        J.job[5] = job(ST_SMOOTH, optm1, D(18), g.shapiro / g.nstep_sub);
        launch_stencil_run<KindsOplus>(g, G, in, J, s);
      } else {
        const double *lonb = in.lonb_op + (size_t)sub * 2 * g.nk * (size_t)G.nlatd;
        const double *latb = in.latb_op + (size_t)sub * 2 * g.nk * (size_t)G.ld;
        dim3 gridt((unsigned)(((G.nlond + TILE_X - 1) / TILE_X) * ((G.nlatd + TILE_Y - 1) / TILE_Y)));
        if (which == 32)  // first of several sub-steps: keeps the factorisation
          hipLaunchKernelGGL((oplus_tile_kernel<false, 1>), gridt, block, 0, s, g, G, in, op, optm1, lonb, latb, W,
                             (const double *)D(19), (const double *)D(20), opout, optm1out);
        else if (which == 33)  // a later sub-step: right-hand side and sweeps
          hipLaunchKernelGGL((oplus_tile_kernel<false, 2>), gridt, block, 0, s, g, G, in, op, optm1, lonb, latb, W,
                             (const double *)D(19), (const double *)D(20), opout, optm1out);
        else if (which == 29)  // stencils fused: one block per column tile
          if (g.capture)
            hipLaunchKernelGGL((oplus_tile_kernel<true, 0>), gridt, block, 0, s, g, G, in, op, optm1, lonb, latb, W,
                               (const double *)D(19), (const double *)D(20), opout, optm1out);
          else
            hipLaunchKernelGGL((oplus_tile_kernel<false, 0>), gridt, block, 0, s, g, G, in, op, optm1, lonb, latb, W,
                               (const double *)D(19), (const double *)D(20), opout, optm1out);
        else if (G.nk <= 64)
          hipLaunchKernelGGL(oplus_wave_stage3<1>, gridw, blockw, 0, s, g, G, in, op, optm1, lonb, latb, W,
                             OT, (const double *)D(19), (const double *)D(20), opout, optm1out);
        else
          hipLaunchKernelGGL(oplus_wave_stage3<2>, gridw, blockw, 0, s, g, G, in, op, optm1, lonb, latb, W,
                             OT, (const double *)D(19), (const double *)D(20), opout, optm1out);
      }
      break;
    }
    case 6:  // duv stage 1: bufs = {ztmp}
      if (G.nk <= 64 && small_grid(ncol))
        hipLaunchKernelGGL(duv_stage1_wave, gridw, blockw, 0, s, g, G, in, (const double *)out.tn, D(0));
      else
        hipLaunchKernelGGL(duv_stage1, grid, block, 0, s, g, G, in, (const double *)out.tn, D(0));
      break;
    case 22: {  // duv stage 2 (PCR): bufs = {ztmp, gamma, divx_ztmp, divy_z, unms, vnms, advu, advv}
      UvTerms UT{D(2), D(3), D(4), D(5), D(6), D(7)};
      if (G.nk <= 64)
        hipLaunchKernelGGL(duv_wave_stage2<1>, gridw, blockw, 0, s, g, G, in, out, UT);
      else
        hipLaunchKernelGGL(duv_wave_stage2<2>, gridw, blockw, 0, s, g, G, in, out, UT);
      break;
    }
    case 8:
    case 14:
    case 23: {  // comp: bufs = {elimination tiles (12 values), comp input tiles (7), fs(:,1,0), fs(:,2,0)}
      CompTerms CT{D(1), {D(2), D(3)}};
      if (which == 8) {
        if (g.capture)
          hipLaunchKernelGGL(comp_kernel<true>, grid, block, 0, s, g, G, in, out, D(0), CT);
        else
          hipLaunchKernelGGL(comp_kernel<false>, grid, block, 0, s, g, G, in, out, D(0), CT);
      } else if (which == 23) {
        if (G.nk <= 64)
          hipLaunchKernelGGL(comp_wave_kernel<1>, gridw, blockw, 0, s, g, G, in, out, CT);
        else
          hipLaunchKernelGGL(comp_wave_kernel<2>, gridw, blockw, 0, s, g, G, in, out, CT);
      } else {
        dim3 grido((unsigned)((ncol * (G.nk - 1) + 255) / 256));
        hipLaunchKernelGGL(o2o_kernel, grido, block, 0, s, g, G, in, CT);
      }
      break;
    }
    case 34:
    case 35:
    case 36:
    case 37: {  // comp precision form: bufs = {elimination tiles (12 floats), comp input tiles (7),
                // fs(:,1,0), fs(:,2,0), u tiles (CM_NV)}
      CompTerms CT{D(1), {D(2), D(3)}};
      float *gs = (float *)D(0);
      if (which == 34)
        hipLaunchKernelGGL(comp_mixed_solve, grid, block, 0, s, g, G, in, gs, D(4), CT);
      else if (which == 35)
        hipLaunchKernelGGL((comp_mixed_refine<true, false>), grid, block, 0, s, g, G, in, out, gs, D(4), CT);
      else if (which == 36)
        hipLaunchKernelGGL((comp_mixed_refine<true, true>), grid, block, 0, s, g, G, in, out, gs, D(4), CT);
      else
        hipLaunchKernelGGL((comp_mixed_refine<false, true>), grid, block, 0, s, g, G, in, out, gs, D(4), CT);
      break;
    }
    case 10: {  // dt prelude: bufs = {w2, hadv_tn, smooth_tn_nm}
      StencilJobs J{};
      J.n = 2;
      J.job[0] = job(ST_ADVEC, in.tn, D(1), 0.);
      J.job[1] = job(ST_SMOOTH, in.tn_nm, D(2), g.shapiro);
      launch_stencil_run<KindsDt>(g, G, in, J, s);
      break;
    }
    case 11: {  // minors prelude (species 0..3 = NO, N4S, N2D, AR)
      StencilJobs J{};
      J.n = 8;
      const double *fc[4] = {in.no, in.n4s, in.n2d, in.ar};
      const double *fm[4] = {in.no_nm, in.n4s_nm, in.n2d_nm, in.ar_nm};
      for (int q = 0; q < 4; ++q) {  // combined into the right-hand-side part of each species
        J.job[q] = job(ST_SMOOTH, fm[q], nullptr, g.shapiro);
        J.job[4 + q] = job(ST_ADVEC_R, fc[q], nullptr, 0.);
      }
      J.ci = D(1);
      launch_stencil_run<KindsMinor, STC_MINOR>(g, G, in, J, s);
      break;
    }
    case 12: {  // comp prelude
      StencilJobs J{};
      J.n = 6;
      J.job[0] = job(ST_ADVEC, in.o2, nullptr, 0.);  // combined into J.ci
      J.job[1] = job(ST_ADVEC, in.o1, nullptr, 0.);
      J.job[2] = job(ST_ADVEC, in.he, nullptr, 0.);
      J.job[3] = job(ST_SMOOTH, in.o2_nm, nullptr, g.shapiro);
      J.job[4] = job(ST_SMOOTH, in.o1_nm, nullptr, g.shapiro);
      J.job[5] = job(ST_SMOOTH, in.he_nm, nullptr, g.shapiro);
      J.aux[0] = D(2);  // fs(:,1,0), fs(:,2,0) of comp_o2o_ng
      J.aux[1] = D(3);
      J.aux[2] = nullptr;
      J.ci = D(1);
      launch_stencil_run<KindsComp, STC_COMP>(g, G, in, J, s);
      break;
    }
    case 13: {  // duv prelude (after stage 1 wrote ztmp)
      StencilJobs J{};
      J.n = 6;
      J.job[0] = job(ST_DIVX, D(0), D(2), 0.);
      J.job[1] = job(ST_DIVY, in.z, D(3), 0.);
      J.job[2] = job(ST_SMOOTH, in.un_nm, D(4), g.shapiro);
      J.job[3] = job(ST_SMOOTH, in.vn_nm, D(5), g.shapiro);
      J.job[4] = job(ST_ADVEC, in.un, D(6), 0.);
      J.job[5] = job(ST_ADVEC, in.vn, D(7), 0.);
      launch_stencil_run<KindsDuv>(g, G, in, J, s);
      break;
    }
    default:
      return tgx_set_error(1, "tgx_ng_launch: unknown kernel %d", which);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tgx_set_error(2, "launch %d: %s", which, hipGetErrorString(e));
  return 0;
}
