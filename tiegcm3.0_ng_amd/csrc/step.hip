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
  size_t P;
  __device__ __forceinline__ size_t a3(int k, int i, int j) const {
    return (size_t)k * P + (size_t)j * ld + i;
  }
  __device__ __forceinline__ size_t a2(int i, int j) const { return (size_t)j * ld + i; }
};

static Geo make_geo(const tgx_ng_grid &g) {
  Geo G;
  G.nk = g.nk; G.nlon = g.nlon; G.nlat = g.nlat;
  G.nlond = g.nlon + 4; G.nlatd = g.nlat + 4; G.ld = g.ld;
  G.P = (size_t)g.ld * (size_t)(g.nlat + 4);
  return G;
}

// ---- stencils on one level plane (dffm_ng.F90, single task) ----------------
// Branch-free: every case of the reference's edge rules is a pair of loads at
// clamped indices and a scale, so a wavefront issues all its loads at once.
// Scaling by 0.5/0.25 is exactly the reference's division by 2/4.
#define PV(f, i, j) (f)[(size_t)(j) * G.ld + (i)]
// df_3d order 1 in longitude: edges one-sided, interior centred / 2
__device__ __forceinline__ double dflon1(const double *f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  int ip = i < L1 ? i + 1 : L1, im = i > 0 ? i - 1 : 0;
  double d = PV(f, ip, j) - PV(f, im, j);
  return (i == 0 || i == L1) ? d : d * 0.5;
}
// df_3d order 2: edges one-sided, next-to-edge centred over 2 cells / 2,
// interior (f(i+2)-f(i-2)) / 4
__device__ __forceinline__ double dflon2(const double *f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
  int ip = edge ? (i == 0 ? 1 : L1) : (near ? i + 1 : i + 2);
  int im = edge ? (i == 0 ? 0 : L1 - 1) : (near ? i - 1 : i - 2);
  double d = PV(f, ip, j) - PV(f, im, j);
  return edge ? d : (near ? d * 0.5 : d * 0.25);
}
// fm_3d order 1/2: edges return f itself ((f+f)*0.5 == f exactly)
__device__ __forceinline__ double fmlon1(const double *f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1);
  int ip = edge ? i : i + 1, im = edge ? i : i - 1;
  return (PV(f, ip, j) + PV(f, im, j)) * 0.5;
}
__device__ __forceinline__ double fmlon2(const double *f, const Geo &G, int i, int j) {
  int L1 = G.nlond - 1;
  bool edge = (i == 0 || i == L1), near = (i == 1 || i == L1 - 1);
  int ip = edge ? i : (near ? i + 1 : i + 2), im = edge ? i : (near ? i - 1 : i - 2);
  return (PV(f, ip, j) + PV(f, im, j)) * 0.5;
}
__device__ __forceinline__ double dflat1(const double *f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  int jp = j < J1 ? j + 1 : J1, jm = j > 0 ? j - 1 : 0;
  double d = PV(f, i, jp) - PV(f, i, jm);
  return (j == 0 || j == J1) ? d : d * 0.5;
}
__device__ __forceinline__ double dflat2(const double *f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1), near = (j == 1 || j == J1 - 1);
  int jp = edge ? (j == 0 ? 1 : J1) : (near ? j + 1 : j + 2);
  int jm = edge ? (j == 0 ? 0 : J1 - 1) : (near ? j - 1 : j - 2);
  double d = PV(f, i, jp) - PV(f, i, jm);
  return edge ? d : (near ? d * 0.5 : d * 0.25);
}
__device__ __forceinline__ double fmlat1(const double *f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1);
  int jp = edge ? j : j + 1, jm = edge ? j : j - 1;
  return (PV(f, i, jm) + PV(f, i, jp)) * 0.5;
}
__device__ __forceinline__ double fmlat2(const double *f, const Geo &G, int i, int j) {
  int J1 = G.nlatd - 1;
  bool edge = (j == 0 || j == J1), near = (j == 1 || j == J1 - 1);
  int jp = edge ? j : (near ? j + 1 : j + 2), jm = edge ? j : (near ? j - 1 : j - 2);
  return (PV(f, i, jm) + PV(f, i, jp)) * 0.5;
}

// advec_ng.F90:244-265 at (i, j) of the plane
__device__ __forceinline__ double advec(const double *f, const double *u, const double *v,
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

// smooth_ng.F90:283-294: latitude pass at (i, j); rows 0/J1 copy, rows 1/J1-1
// use the 3-point form
__device__ __forceinline__ double smooth_lat(const double *f, const Geo &G, int i, int j, double s) {
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
__device__ __forceinline__ double smooth(const double *f, const Geo &G, int i, int j, double s) {
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
__device__ __forceinline__ double divx(const double *f, const Geo &G, int i, int j) {
  double d1 = dflon1(f, G, i, j);
  double d2 = dflon2(f, G, i, j);
  double gen = d1 * 2. * 2. / 3. - d2 * 4. / 12.;
  return (i <= 1 || i >= G.nlon + 2) ? d1 : gen;
}
__device__ __forceinline__ double divy(const double *f, const Geo &G, int i, int j) {
  double d1 = dflat1(f, G, i, j);
  double d2 = dflat2(f, G, i, j);
  double gen = d1 * 2. * 2. / 3. - d2 * 4. / 12.;
  return (j <= 1 || j >= G.nlat + 2) ? d1 : gen;
}

// oplus_ng.F90:344-365 (bdotdh)
__device__ __forceinline__ double bdotdh(const double *f, const Geo &G, int i, int j, double bx,
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
  // first, then lat ghosts (corners take the lat data).
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
enum { ST_ADVEC = 0, ST_SMOOTH, ST_DIVX, ST_DIVY, ST_BDOTDH, ST_VDOTN };
struct StencilJob {
  int kind;
  const double *f;
  double *out;
  double s;   // Shapiro coefficient for ST_SMOOTH
};
struct StencilJobs {
  int n;
  StencilJob job[8];
};

__global__ void __launch_bounds__(256) stencil_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                      StencilJobs J) {
  int nb = gridDim.x;
  long t = (long)tgx_xcd_block(nb) * blockDim.x + threadIdx.x;
  long npt = (long)G.nlond * G.nlatd * G.nk;
  if (t >= npt) return;
  int i = (int)(t % G.nlond);
  long r = t / G.nlond;
  int j = (int)(r % G.nlatd);
  int k = (int)(r / G.nlatd);
  const size_t off = (size_t)k * G.P, p = off + (size_t)j * G.ld + i, c2 = (size_t)j * G.ld + i;
  for (int n = 0; n < J.n; ++n) {
    const StencilJob &jb = J.job[n];
    const double *f = jb.f + off;
    double v;
    switch (jb.kind) {
      case ST_ADVEC:
        v = advec(f, in.un + off, in.vn + off, G, i, j, g.racs[j], g.dlamda, g.dphi);
        break;
      case ST_SMOOTH:
        v = smooth(f, G, i, j, jb.s);
        break;
      case ST_DIVX:
        v = divx(f, G, i, j);
        break;
      case ST_DIVY:
        v = divy(f, G, i, j);
        break;
      case ST_BDOTDH:
        v = bdotdh(f, G, i, j, in.bx[c2], in.by[c2], g.cs[j], g.dlamda, g.dphi);
        break;
      default: {  // ST_VDOTN: oplus_ng.F90:196-209, f = op/bmod2**2
        double uii = k < G.nk - 1 ? .5 * (in.ui[p] + in.ui[p + G.P]) : 1.5 * in.ui[p] - .5 * in.ui[p - G.P];
        double vii = k < G.nk - 1 ? .5 * (in.vi[p] + in.vi[p + G.P]) : 1.5 * in.vi[p] - .5 * in.vi[p - G.P];
        double vx = uii * dflon1(f, G, i, j), vy = vii * dflat1(f, G, i, j);
        vx = vx / g.cs[j];
        double vh = 1. / RE * (vx / g.dlamda + vy / g.dphi);
        double bm = in.bmod2[c2];
        v = vh * (bm * bm);
      }
    }
    jb.out[p] = v;
  }
}

// =============================================================================
// dt_ng.F90:54-213
__global__ void __launch_bounds__(256) dt_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                 tgx_ng_out out, double *__restrict__ w2s,
                                                 const double *__restrict__ hadv_tn,
                                                 const double *__restrict__ smooth_tn_nm) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
  const double t_lbc = in.t_lbc[c2], tl = in.tlbc[c2];
  out.tlbc_nm[c2] = tl;
  out.tlbc[c2] = t_lbc;
  double v;
  if (lateral_value(G, in.lonb_tn, in.latb_tn, 0, i, j, v)) {
    for (int k = 0; k < nk; ++k) {
      lateral_value(G, in.lonb_tn, in.latb_tn, k, i, j, v);
      size_t p = G.a3(k, i, j);
      out.tn_nm[p] = DTSMOOTH_DIV2 * (in.tn_nm[p] + v) + DTSMOOTH * in.tn[p];
      out.tn[p] = v > 100. ? v : 100.;
    }
    return;
  }
  const double racs = g.racs[j];
#define F3(f, k) (f)[G.a3((k), i, j)]
  // g(L), f(L) of dt_ng.F90:121-139
  auto gf = [&](int L, double tni, double &gg, double &ff) {
    double h = F3(in.schti, L);
    double rho = P0 * g.expzmid_inv / (h * GRAV);
    rho = rho * g.expz[L];
    double gpart = h * h * rho * F3(in.cp, L);
    double gpart_a = gpart * g.dift[L];
    gg = GRAV * (F3(in.kt, L) + gpart_a) / (P0 * h * (dz * dz));
    ff = GRAV * GRAV * (h * h) * rho / (tni * P0 * 2. * dz);
    ff = ff * g.dift[L];
  };
  double tn_m = 0., tn_c = F3(in.tn, 0), tn_p = F3(in.tn, 1);
  double gk, fk;
  gf(0, tl, gk, fk);
  double w1 = 0., w3 = 0., rprev = 0., dudz_prev = 0., dvdz_prev = 0.;
  for (int k = 0; k <= nk - 2; ++k) {
    const double w_k = F3(in.w_upd, k), w_k1 = F3(in.w_upd, k + 1);
    // veradv (:57-65)
    double raw_k = k == 0 ? (tn_c - t_lbc) * w_k * 2. : (tn_c - tn_m) * w_k;
    double raw_k1 = (k + 1 <= nk - 2) ? (tn_p - tn_c) * w_k1 : 0.;
    double veradv = (raw_k + raw_k1) * .5;
    veradv = veradv / dz;
    // interface means (:71-78)
    double cpi = .5 * (F3(in.cp, k) + F3(in.cp, k + 1));
    double kmi = .5 * (F3(in.km, k) + F3(in.km, k + 1));
    double wi = .5 * (w_k + w_k1);
    // cptn (:67, :80-83)
    double horadv = F3(hadv_tn, k);
    double tnsmooth = F3(smooth_tn_nm, k);
    double advec_tn = horadv + veradv;
    double cptn = cpi * (advec_tn - dtx2inv * tnsmooth);
    cptn = cptn * g.expz[k];
    // dudz (:85-94)
    double dudz, dvdz;
    if (k == 0) {
      dudz = (F3(in.un, 0) + 1. / 3. * F3(in.un, 1) - 4. / 3. * in.ulbc[c2]) / dz;
      dvdz = (F3(in.vn, 0) + 1. / 3. * F3(in.vn, 1) - 4. / 3. * in.vlbc[c2]) / dz;
    } else if (k <= nk - 3) {
      dudz = (F3(in.un, k + 1) - F3(in.un, k - 1)) / (2. * dz);
      dvdz = (F3(in.vn, k + 1) - F3(in.vn, k - 1)) / (2. * dz);
    } else {
      dudz = dudz_prev / 3.;
      dvdz = dvdz_prev / 3.;
    }
    dudz_prev = dudz;
    dvdz_prev = dvdz;
    // heating (:96-113)
    double qsolar = .5 * (F3(in.qtotal, k) + F3(in.qtotal, k + 1));
    double t = F3(in.xnmbar, k) * F3(in.o1, k) * RIO1;
    double qrecomb = TSURPLUS * F3(in.rkm12, k) * (t * t) * AVO / F3(in.mbar, k);
    double qm = GRAV * kmi / (P0 * F3(in.scht, k)) * (dudz * dudz + dvdz * dvdz);
    qm = qm / g.expz[k];
    double total_heat = qsolar + F3(in.hdt, k) + qrecomb + F3(in.qji_tn, k) + qm;
    double total_heat_a = total_heat * g.expz[k];
    cptn = cptn - total_heat_a;
    // coefficients (:141-163)
    double p = gk - fk, q, r;
    double g1 = 0., f1 = 0.;
    if (k <= nk - 3) {
      double tni1 = .5 * (tn_c + tn_p);
      gf(k + 1, tni1, g1, f1);
      q = -(g1 - f1 + gk + fk);
      r = g1 + f1;
    } else {
      q = -gk - fk;
      r = 0.;
    }
    double rhs = cptn;
    double qpart = cpi * (dtx2inv + F3(in.cool_imp, k)) + wi * GASK / F3(in.mbar, k);
    rhs = rhs + F3(in.cool_exp, k);
    double qpart_a = qpart * g.expz[k];
    q = q - qpart_a;
    if (k == 0) {
      q = q - p;
      rhs = rhs - 2. * p * t_lbc;
      p = 0.;
    }
    // Thomas forward (trsolv_ng.F90:14-22)
    size_t pk = G.a3(k, i, j);
    if (k == 0) {
      w1 = q;
      w3 = rhs / w1;
    } else {
      double w2 = rprev / w1;
      w2s[G.a3(k - 1, i, j)] = w2;
      w1 = q - p * w2;
      w3 = (rhs - p * w3) / w1;
    }
    out.tn[pk] = w3;
    rprev = r;
    gk = g1; fk = f1;
    tn_m = tn_c; tn_c = tn_p;
    if (k + 2 <= nk - 1) tn_p = F3(in.tn, k + 2);
  }
  // back substitution (:165-167) fused with the output stage (:208-213)
  double xn = w3, xnn = 0.;
  auto emit = [&](int k, double x) {
    size_t p = G.a3(k, i, j);
    out.tn_nm[p] = DTSMOOTH_DIV2 * (in.tn_nm[p] + x) + DTSMOOTH * in.tn[p];
    out.tn[p] = x > 100. ? x : 100.;
  };
  for (int k = nk - 2; k >= 0; --k) {
    double x = k == nk - 2 ? xn : out.tn[G.a3(k, i, j)] - w2s[G.a3(k, i, j)] * xn;
    if (k == nk - 3) emit(nk - 1, xn * xn / x);
    if (k < nk - 2) xnn = xn;
    (void)xnn;
    emit(k, x);
    xn = x;
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
  double pso1, pso2, pshe, dmdz, do1dz, do2dz, dhedz, tni, barm, inv[9];
};
struct MinorSpecies {  // per species and level
  double ax, thdiff, sloss, s0prod, ftm1s, hadvec;
};

struct MinorTerms {  // prelude planes: smoothed fcomp_tm1 and advection of fcomp, by species id
  const double *ftm1s[4], *hadvec[4];
};

template <int SA, int SB>
__global__ void __launch_bounds__(256) minor_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                    tgx_ng_out out, double *__restrict__ w2a,
                                                    double *__restrict__ w2b, MinorTerms MT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dzp = g.dz, dtx2inv = g.dtx2inv;
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

  auto tni_at = [&](int L) -> double {  // :89-93
    if (L == 0) return in.tlbc[c2];
    if (L == nk - 1) return F3(in.tn, nk - 2);
    return .5 * (F3(in.tn, L) + F3(in.tn, L - 1));
  };
  auto shared_at = [&](int L, MinorShared &S) {
    if (L == 0) {
      S.dmdz = (F3(in.mbar, 0) - xmbari) / dzp;
      S.pso1 = .5 * (bo1 + o1_0);
      S.pso2 = .5 * (bo2 + o2_0);
      S.pshe = .5 * (bhe + he_0);
      S.do1dz = (o1_0 - bo1) / dzp;
      S.do2dz = (o2_0 - bo2) / dzp;
      S.dhedz = (he_0 - bhe) / dzp;
    } else {
      S.dmdz = (F3(in.mbar, L) - F3(in.mbar, L - 1)) / dzp;
      S.pso1 = .5 * (F3(in.o1, L) + F3(in.o1, L - 1));
      S.pso2 = .5 * (F3(in.o2, L) + F3(in.o2, L - 1));
      S.pshe = .5 * (F3(in.he, L) + F3(in.he, L - 1));
      S.do1dz = (F3(in.o1, L) - F3(in.o1, L - 1)) / dzp;
      S.do2dz = (F3(in.o2, L) - F3(in.o2, L - 1)) / dzp;
      S.dhedz = (F3(in.he, L) - F3(in.he, L - 1)) / dzp;
    }
    S.tni = tni_at(L);
    S.barm = F3(in.barm, L);
    double df[3], al[9];
    for (int n = 0; n < 3; ++n) df[n] = pow(S.tni / T00, 1.75 - ssn(n));
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
    matinv3_ng(al, S.inv);
  };
  // sources of species s at level L (comp_*_ng.F90)
  auto sources = [&](int s, int L, double &sloss, double &sprod) {
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
  };
  auto species_at = [&](int q, int L, const MinorShared &S, double dtnidz, MinorSpecies &T) {
    const SpeciesConst &c = sc[q];
    double sai[3];
    for (int n = 0; n < 3; ++n) {
      double s = 0.;
      for (int m = 0; m < 3; ++m) s = s + c.salfax[m] * S.inv[m + 3 * n];
      sai[n] = s;
    }
    double ex = 1. - (c.rmx + S.dmdz) / S.barm +
                sai[0] * (S.do2dz - (1. - (RMO2 + S.dmdz) / S.barm) * S.pso2) +
                sai[1] * (S.do1dz - (1. - (RMO1 + S.dmdz) / S.barm) * S.pso1) +
                sai[2] * (S.dhedz - (1. - (RMHE + S.dmdz) / S.barm) * S.pshe);
    T.ax = -(S.barm / (TAU * RMN2) * pow(T00 / S.tni, 0.25) /
             (c.phix[2] + c.salfax[0] * S.pso2 + c.salfax[1] * S.pso1 + c.salfax[2] * S.pshe));
    T.thdiff = ex - c.alfax * dtnidz / (dzp * S.tni);
    double sprod;
    sources(SP[q], L, T.sloss, sprod);
    T.s0prod = sprod * c.rmx / F3(in.xnmbar, L);
    T.ftm1s = F3(MT.ftm1s[SP[q]], L);
    T.hadvec = F3(MT.hadvec[SP[q]], L);
  };
  auto wi_at = [&](int L) -> double {
    if (L < nk - 1) return .5 * (F3(in.w_upd, L) + F3(in.w_upd, L + 1));
    return 1.5 * F3(in.w_upd, nk - 1) - .5 * F3(in.w_upd, nk - 2);
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

  // sliding window: levels k (cur) and k+1 (nxt)
  MinorShared Sc, Sn;
  MinorSpecies Tc[2], Tn[2];
  double tni_m = 0.;
  shared_at(0, Sc);
  double tni_1 = tni_at(1);
#pragma unroll
  for (int q = 0; q < 2; ++q) species_at(q, 0, Sc, tni_1 - Sc.tni, Tc[q]);
  double w1[2] = {0., 0.}, w3[2] = {0., 0.}, rprev[2] = {0., 0.};
  double *w2s[2] = {w2a, w2b};
  double wic = wi_at(0);
  for (int k = 0; k < nk; ++k) {
    double wi_n = 0.;
    if (k + 1 <= nk - 1) {
      shared_at(k + 1, Sn);
      double tni_p = (k + 2 <= nk - 1) ? tni_at(k + 2) : 0.;
      double dtnidz = (k + 1 == nk - 1) ? Sn.tni - Sc.tni : (tni_p - Sc.tni) / 2.;
    #pragma unroll
  for (int q = 0; q < 2; ++q) species_at(q, k + 1, Sn, dtnidz, Tn[q]);
      wi_n = wi_at(k + 1);
    }
    (void)tni_m;
    const double dmdzc = Sc.dmdz / Sc.barm;
    const double dmdzn = k + 1 <= nk - 1 ? Sn.dmdz / Sn.barm : 0.;
  #pragma unroll
  for (int q = 0; q < 2; ++q) {
      double pc, qc, rc, fr;
      if (k <= nk - 2) {
        double molp = Tc[q].ax * (1. / dzp + .5 * Tc[q].thdiff);
        double molr = Tn[q].ax * (1. / dzp - .5 * Tn[q].thdiff);
        double molq = Tc[q].ax * (1. / dzp - .5 * Tc[q].thdiff) + Tn[q].ax * (1. / dzp + .5 * Tn[q].thdiff);
        double eddyp = g.expzmid_inv * g.difk[k] * (1. / dzp - .5 * dmdzc);
        double eddyr = g.expzmid * g.difk[k + 1] * (1. / dzp + .5 * dmdzn);
        double eddyq = g.expzmid_inv * g.difk[k] * (1. / dzp + .5 * dmdzc) +
                       g.expzmid * g.difk[k + 1] * (1. / dzp - .5 * dmdzn);
        pc = molp / dzp - g.expz[k] * (eddyp * dfac + .5 * wic) / dzp;
        rc = molr / dzp - g.expz[k] * (eddyr * dfac - .5 * wic) / dzp;
        qc = -molq / dzp + g.expz[k] * (eddyq * dfac / dzp + dtx2inv - Tc[q].sloss);
        fr = Tc[q].ftm1s * dtx2inv - Tc[q].hadvec + Tc[q].s0prod;
        fr = fr * g.expz[k];
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
      // Thomas forward, rows 1..nk (minor_ng.F90:200)
      if (k == 0) {
        w1[q] = qc;
        w3[q] = fr / w1[q];
      } else {
        double w2 = rprev[q] / w1[q];
        w2s[q][G.a3(k - 1, i, j)] = w2;
        w1[q] = qc - pc * w2;
        w3[q] = (fr - pc * w3[q]) / w1[q];
      }
      outs[q][G.a3(k, i, j)] = w3[q];
      rprev[q] = rc;
    }
    Sc = Sn;
    Tc[0] = Tn[0];
    Tc[1] = Tn[1];
    wic = wi_n;
  }
  // back substitution and output stage (:234-237)
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    double xn = w3[q];
    for (int k = nk - 1; k >= 0; --k) {
      size_t p = G.a3(k, i, j);
      double x = k == nk - 1 ? xn : outs[q][p] - w2s[q][p] * xn;
      xn = x;
      double nm = DTSMOOTH * fc[q][p] + DTSMOOTH_DIV2 * (fctm1[q][p] + x);
      outs[q][p] = x > small ? x : small;
      outs_nm[q][p] = nm > small ? nm : small;
    }
  }
#undef F3
}


// =============================================================================
// comp_o2o_ng.F90:44-81 (fused) + comp_ng.F90:71-457
struct CompTerms {  // prelude planes: advection of O2, O, He and smoothed *_nm
  const double *o2adv, *o1adv, *headv, *o2s, *o1s, *hes;
};

__global__ void __launch_bounds__(256) comp_kernel(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                   tgx_ng_out out, double *__restrict__ gs,
                                                   CompTerms CT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv, thdiffalpha = -0.38, small = 1.e-9;
  const size_t c2 = G.a2(i, j);
  const double fhe = g.calc_helium == 0 ? 0. : in.flx_he[c2];
  out.flx_he[c2] = fhe;
#define F3(f, k) (f)[G.a3((k), i, j)]
  double v;
  const bool bnd = lateral_value(G, in.lonb_o2, in.latb_o2, 0, i, j, v);
  if (bnd) {
    for (int k = 0; k < nk; ++k) {
      size_t p = G.a3(k, i, j);
      lateral_value(G, in.lonb_o2, in.latb_o2, k, i, j, v); out.o2[p] = v;
      lateral_value(G, in.lonb_o1, in.latb_o1, k, i, j, v); out.o1[p] = v;
      lateral_value(G, in.lonb_he, in.latb_he, k, i, j, v); out.he[p] = v;
    }
  } else {
    const double dfac = dfactor_at(g, j), racs = g.racs[j];
    const double tl = in.t_lbc[c2];  // flds%tlbc after dt_ng of this step
    const double rmass[3] = {RMO2, RMO1, RMHE};
    const double o2_0 = F3(in.o2, 0), o1_0 = F3(in.o1, 0), he_0 = F3(in.he, 0);
    double ps0[3];
    for (int l = 0; l < 3; ++l)
      ps0[l] = g.b[3 * l + 0] * o2_0 + g.b[3 * l + 1] * o1_0 + g.b[3 * l + 2] * he_0 + g.fb[l];
    const double embar0 =
        1. / (ps0[0] * RIO2 + ps0[1] * RIO1 + ps0[2] * RIHE + (1. - ps0[0] - ps0[1] - ps0[2]) * RIN2);
    struct Lev { double o2i, o1i, hei, embari, tni, dembardz, dtndz, df[3]; };
    auto lev = [&](int L, Lev &Q) {  // comp_ng.F90:80-97, :124-126
      if (L == 0) {
        Q.o2i = .5 * (ps0[0] + o2_0);
        Q.o1i = .5 * (ps0[1] + o1_0);
        Q.hei = .5 * (ps0[2] + he_0);
        Q.embari = .5 * (embar0 + F3(in.mbar, 0));
        Q.tni = tl;
        Q.dembardz = F3(in.mbar, 0) - embar0;
        Q.dtndz = (F3(in.tn, 0) - tl) * 2;
      } else {
        Q.o2i = .5 * (F3(in.o2, L - 1) + F3(in.o2, L));
        Q.o1i = .5 * (F3(in.o1, L - 1) + F3(in.o1, L));
        Q.hei = .5 * (F3(in.he, L - 1) + F3(in.he, L));
        Q.embari = .5 * (F3(in.mbar, L - 1) + F3(in.mbar, L));
        Q.tni = .5 * (F3(in.tn, L - 1) + F3(in.tn, L));
        Q.dembardz = F3(in.mbar, L) - F3(in.mbar, L - 1);
        Q.dtndz = F3(in.tn, L) - F3(in.tn, L - 1);
      }
      Q.dembardz = Q.dembardz / dz;
      Q.dtndz = Q.dtndz / dz;
      for (int n = 0; n < 3; ++n) Q.df[n] = pow(Q.tni / T00, 1.75 - ssn(n));
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
    auto wi_at = [&](int L) -> double {
      if (L < nk - 1) return .5 * (F3(in.w, L) + F3(in.w, L + 1));
      return 1.5 * F3(in.w, nk - 1) - .5 * F3(in.w, nk - 2);
    };
    // fs(L, isp, m) of comp_o2o_ng.F90:44-81 (nonzero planes only)
    auto fs_at = [&](int L, double fs[2][3]) {
      size_t p = G.a3(L, i, j);
      double rji, qo2pi, qopi, beta8i, nei;
      if (L < nk - 1) {
        rji = 0.5 * (in.rj[p] + in.rj[p + G.P]);
        qo2pi = 0.5 * (in.qo2p[p] + in.qo2p[p + G.P]);
        qopi = 0.5 * (in.qop[p] + in.qop[p + G.P]);
        beta8i = 0.5 * (in.beta8[p] + in.beta8[p + G.P]);
        nei = sqrt(in.ne[p] * in.ne[p + G.P]);
      } else {
        rji = 1.5 * in.rj[p] - .5 * in.rj[p - G.P];
        qo2pi = 1.5 * in.qo2p[p] - .5 * in.qo2p[p - G.P];
        qopi = 1.5 * in.qop[p] - .5 * in.qop[p - G.P];
        beta8i = 1.5 * in.beta8[p] - .5 * in.beta8[p - G.P];
        nei = sqrt(in.ne[p] * in.ne[p] * in.ne[p] / in.ne[p - G.P]);
      }
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
      fs[0][1] = -lo21;
      fs[0][2] = xn * po21 * in.o1[p] * (RIO1 * RIO1) * RMO2;
      fs[1][1] = pox2 * RMO1 * RIO2;
      fs[1][2] = -lox2 - lox1 * in.o1[p] * RIO1 * xn;
      fs[0][0] = (po23 - lo22) * RMO2 / xn;
      fs[1][0] = (pox1 - lox3) * RMO1 / xn;
    };

    Lev Qc, Qn;
    lev(0, Qc);
    double epm[3], epp[3], akm[9], akp[9], alpha[9];  // levels k (km) and k+1 (kp)
    double wks3, wks4 = Qc.dembardz / (Qc.embari * 2.);
    for (int l = 0; l < 3; ++l) epp[l] = 1. - 1. / Qc.embari * (rmass[l] + Qc.dembardz);
    epp[2] = epp[2] - thdiffalpha * Qc.dtndz / Qc.tni;
    alpha_of(Qc, alpha);
    matinv3_ng(alpha, akp);
    double wks1 = Qc.embari * RIN2 * pow(T00 / tl, 0.25) / TAU;
    for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
    double gama[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, zz[3] = {0, 0, 0};
    double alpha22 = 0, alpha33 = 0, alpha23 = 0, alpha32 = 0, flx00 = 0;
    for (int k = 0; k < nk - 1; ++k) {
      for (int n = 0; n < 9; ++n) akm[n] = akp[n];
      for (int n = 0; n < 3; ++n) epm[n] = epp[n];
      lev(k + 1, Qn);
      for (int l = 0; l < 3; ++l) epp[l] = 1. - 1. / Qn.embari * (rmass[l] + Qn.dembardz);
      if (k != nk - 2) epp[2] = epp[2] - thdiffalpha * Qn.dtndz / Qn.tni;
      alpha_of(Qn, alpha);
      matinv3_ng(alpha, akp);
      wks1 = Qn.embari * RIN2 * pow(T00 / Qn.tni, 0.25) / TAU;
      wks3 = wks4;
      wks4 = Qn.dembardz / (Qn.embari * 2.);
      for (int n = 0; n < 9; ++n) akp[n] = akp[n] * wks1;
      double molp[9], molr[9], molq[9], pk[9], rk[9], qk[9], fk[3], fs[2][3];
      for (int isp = 0; isp < 3; ++isp)
        for (int m = 0; m < 3; ++m) {
          int e = isp + 3 * m;
          molp[e] = akm[e] * (1. / dz + epm[m] / 2.);
          molr[e] = akp[e] * (1. / dz - epp[m] / 2.);
          molq[e] = akm[e] * (1. / dz - epm[m] / 2.) + akp[e] * (1. / dz + epp[m] / 2.);
        }
      double eddyp = g.expzmid_inv * g.difk[k] * (1. / dz - wks3);
      double eddyr = g.expzmid * g.difk[k + 1] * (1. / dz + wks4);
      double eddyq = g.expzmid_inv * g.difk[k] * (1. / dz + wks3) + g.expzmid * g.difk[k + 1] * (1. / dz - wks4);
      fs_at(k, fs);
      const double wik = wi_at(k);
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) {
          int e = isp + 3 * m;
          double delta = isp == m ? 1. : 0.;
          double fsv = (isp < 2 && m < 2) ? fs[isp][m + 1] : 0.;
          pk[e] = molp[e] / dz - g.expz[k] * delta / dz * (eddyp * dfac + .5 * wik);
          rk[e] = molr[e] / dz - g.expz[k] * delta / dz * (eddyr * dfac - .5 * wik);
          qk[e] = -molq[e] / dz + g.expz[k] * delta * (eddyq * dfac / dz + dtx2inv) - g.expz[k] * fsv;
        }
      double o2adv = F3(CT.o2adv, k), o1adv = F3(CT.o1adv, k), headv = F3(CT.headv, k);
      double o2s = F3(CT.o2s, k), o1s = F3(CT.o1s, k), hes = F3(CT.hes, k);
      fk[0] = g.expz[k] * (o2s * dtx2inv - (o2adv - fs[0][0]) + F3(in.hdo2, k));
      fk[1] = g.expz[k] * (o1s * dtx2inv - (o1adv - fs[1][0]) + F3(in.hdo1, k));
      fk[2] = g.expz[k] * (hes * dtx2inv - (headv - 0.) + F3(in.hdhe, k));
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
        flx00 = Qc.embari * RIN2 * P0 * pow(T00 / F3(in.tn, k), 0.25) / (TAU * GRAV);
        for (int isp = 0; isp < 3; ++isp)
          fk[isp] = fk[isp] - rk[isp + 3 * 1] * (alpha23 - alpha22) * fhe / (flx00 * (1. / dz - 0.5 * epp[1])) -
                    rk[isp + 3 * 2] * (alpha33 - alpha32) * fhe / (flx00 * (1. / dz - 0.5 * epp[2]));
        for (int n = 0; n < 9; ++n) rk[n] = 0.;
      }
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp) qk[isp + 3 * m] = qk[isp + 3 * m] - pk[isp + 3 * n] * gama[n + 3 * m];
      double wkm1[9], wkv1[3], ng[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, nz[3] = {0, 0, 0};
      matinv3_ng(qk, wkm1);
      for (int isp = 0; isp < 3; ++isp) wkv1[isp] = fk[isp];
      for (int m = 0; m < 3; ++m) {
        for (int isp = 0; isp < 3; ++isp) wkv1[isp] = wkv1[isp] - pk[isp + 3 * m] * zz[m];
        for (int n = 0; n < 3; ++n)
          for (int isp = 0; isp < 3; ++isp) ng[isp + 3 * m] = ng[isp + 3 * m] + wkm1[isp + 3 * n] * rk[n + 3 * m];
      }
      for (int m = 0; m < 3; ++m)
        for (int isp = 0; isp < 3; ++isp) nz[isp] = nz[isp] + wkm1[isp + 3 * m] * wkv1[m];
      for (int n = 0; n < 9; ++n) gama[n] = ng[n];
      for (int n = 0; n < 9; ++n) gs[(size_t)n * ((size_t)nk * G.P) + G.a3(k, i, j)] = ng[n];
      for (int isp = 0; isp < 3; ++isp) zz[isp] = nz[isp];
      out.o2[G.a3(k, i, j)] = nz[0];
      out.o1[G.a3(k, i, j)] = nz[1];
      out.he[G.a3(k, i, j)] = nz[2];
      Qc = Qn;
    }
    // :317-332 back substitution and the closed-form top level
    double u1[3] = {0., 0., 0.};
    for (int k = nk - 2; k >= 0; --k) {
      size_t p = G.a3(k, i, j);
      double z0 = out.o2[p], z1 = out.o1[p], z2 = out.he[p], u[3];
      double zr[3] = {z0, z1, z2};
      for (int isp = 0; isp < 3; ++isp) {
        double uu = zr[isp];
        for (int m = 0; m < 3; ++m) uu = uu - gs[(size_t)(isp + 3 * m) * ((size_t)nk * G.P) + p] * u1[m];
        u[isp] = uu;
      }
      out.o2[p] = u[0]; out.o1[p] = u[1]; out.he[p] = u[2];
      for (int isp = 0; isp < 3; ++isp) u1[isp] = u[isp];
    }
    double top[3];
    for (int isp = 0; isp < 3; ++isp)
      top[isp] = (1. + .5 * epp[isp] * dz) / (1. - .5 * epp[isp] * dz) * (isp == 0 ? F3(out.o2, nk - 2) : isp == 1 ? F3(out.o1, nk - 2) : F3(out.he, nk - 2));
    top[1] = top[1] + (alpha23 - alpha22) * fhe / (flx00 * (1. / dz - 0.5 * epp[1]));
    top[2] = top[2] + (alpha33 - alpha32) * fhe / (flx00 * (1. / dz - 0.5 * epp[2]));
    size_t pt = G.a3(nk - 1, i, j);
    out.o2[pt] = top[0]; out.o1[pt] = top[1]; out.he[pt] = top[2];
    if (g.calc_helium == 0)
      for (int k = 0; k < nk; ++k) out.he[G.a3(k, i, j)] = g.pshelb;
  }
  // :415-457 time smoothing, N2 above the O-dominated top, clamps, normalisation
  bool found = false, extrap = false;
  double n2top = 0., dn2dz = 0., zmtop = 0., n2prev = 0.;
  for (int k = 0; k < nk; ++k) {
    size_t p = G.a3(k, i, j);
    double o2u = out.o2[p], o1u = out.o1[p], heu = out.he[p];
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
          n2top = log(n2 * in.xnmbar[p]);
          double n2top1 = log(n2prev * in.xnmbar[p - G.P]);
          dn2dz = (n2top - n2top1) / dz;
        }
      }
    } else if (extrap) {
      n2 = exp(n2top + dn2dz * (g.zpmid[k] - zmtop)) / in.xnmbar[p];
    }
    n2prev = n2;
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
#undef F3
}


// =============================================================================
// oplus_ng.F90, sub-step stage 1 (every column, ghosts included): the
// pointwise and vertical quantities whose horizontal derivatives are needed
// (:106-170).  Writes dj, bdotu, bvel, tp_op, diffj, djbz, op/bmod2**2 to the
// workspace and Fe, Fn to the outputs.
struct OpWork {
  double *dj, *bdotu, *bvel, *tpop, *diffj, *djbz, *opbm2, *bdopj, *w2;
};
struct OpTerms {  // prelude planes of one sub-step
  const double *bd_diff, *bd_diffj, *bd_bvel, *vdotn, *bd_djbz, *opts2;
};

__global__ void __launch_bounds__(256) oplus_stage1(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                    const double *__restrict__ op, OpWork W,
                                                    double *__restrict__ Fe, double *__restrict__ Fn) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dz = g.dz;
  const size_t c2 = G.a2(i, j);
  const double bx = in.bx[c2], by = in.by[c2], bz = in.bz[c2], bm = in.bmod2[c2];
  const double sncsdip = sin(in.dipmag[c2]) * cos(in.dipmag[c2]);
  const double sndec = in.sndec[c2], csdec = in.csdec[c2];
#define F3(f, k) (f)[G.a3((k), i, j)]
  auto tp_at = [&](int L) { return (F3(in.te, L) + F3(in.ti, L)) * F3(op, L); };
  double tpm = 0., tpc = tp_at(0), tpn = tp_at(1);
  for (int k = 0; k < nk; ++k) {
    size_t p = G.a3(k, i, j);
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
    double hj = in.scht[p];
    double bdotu = bx * in.un[p] + by * in.vn[p] + hj * bz * wni;
    double dtp;
    if (k == 0) dtp = tpn - tpc;
    else if (k <= nk - 3) dtp = (tpn - tpm) / 2.;
    else dtp = tpc - tpm;
    double diffj = 1. / (hj * dz) * dtp + RMOP * GRAV / GASK * op[p];
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

// stage 2 (every column): bdotdh_opj = bdotdh(tp_op) * dj (:175-176)
__global__ void __launch_bounds__(256) oplus_stage2(tgx_ng_grid g, Geo G, tgx_ng_fields in, OpWork W) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const size_t c2 = G.a2(i, j);
  const double bx = in.bx[c2], by = in.by[c2], cs = g.cs[j];
  for (int k = 0; k < G.nk; ++k) {
    size_t off = (size_t)k * G.P, p = G.a3(k, i, j);
    W.bdopj[p] = bdotdh(W.tpop + off, G, i, j, bx, by, cs, g.dlamda, g.dphi) * W.dj[p];
  }
}

// stage 3: assembly, solve and outputs of one sub-step (:172-340)
__global__ void __launch_bounds__(256) oplus_stage3(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                    const double *__restrict__ op,
                                                    const double *__restrict__ optm1,
                                                    const double *__restrict__ lonb,
                                                    const double *__restrict__ latb, OpWork W,
                                                    OpTerms OT,
                                                    double *__restrict__ opout,
                                                    double *__restrict__ optm1out,
                                                    double *__restrict__ xiop2p,
                                                    double *__restrict__ xiop2d) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk, nsub = g.nstep_sub;
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
#define F3(f, k) (f)[G.a3((k), i, j)]
  // xiop2p, xiop2d and the O+ loss (:288-293), every column and level
  auto chem = [&](int L, double &x2p, double &x2d, double &oploss) {
    size_t p = G.a3(L, i, j);
    double xn = in.xnmbar[p], ne = in.ne[p];
    double qop2pi = L < nk - 1 ? .5 * (in.qop2p[p] + in.qop2p[p + G.P]) : 1.5 * in.qop2p[p] - .5 * in.qop2p[p - G.P];
    double qop2di = L < nk - 1 ? .5 * (in.qop2d[p] + in.qop2d[p + G.P]) : 1.5 * in.qop2d[p] - .5 * in.qop2d[p - G.P];
    x2p = qop2pi / (xn * ((RK16 + RK17) * in.n2[p] * RIN2 + RK18 * in.o1[p] * RIO1) +
                    (in.rk19[p] + in.rk20[p]) * ne + RK21 + RK22);
    x2d = (qop2di + (in.rk20[p] * ne + RK22) * x2p) /
          (xn * (RK23 * in.n2[p] * RIN2 + RK24 * in.o1[p] * RIO1 + RK26 * in.o2[p] * RIO2) + in.rk25[p] * ne + RK27);
    oploss = xn * (in.rk1[p] * in.o2[p] * RIO2 + in.rk2[p] * in.n2[p] * RIN2 + RK10 * in.n2d[p] * RIN2D);
  };
  double v;
  if (lateral_value(G, lonb, latb, 0, i, j, v)) {
    for (int k = 0; k < nk; ++k) {
      size_t p = G.a3(k, i, j);
      double x2p, x2d, ol;
      chem(k, x2p, x2d, ol);
      xiop2p[p] = x2p;
      xiop2d[p] = x2d;
      lateral_value(G, lonb, latb, k, i, j, v);
      double nm = DTSMOOTH * op[p] + DTSMOOTH_DIV2 * (optm1[p] + v);
      opout[p] = v > 0. ? v : 0.;
      optm1out[p] = nm > 0. ? nm : 0.;
    }
    return;
  }
  const double pi = 4.0 * atan(1.0), rtd = 180. / pi;
  const double bx = in.bx[c2], by = in.by[c2], bz = in.bz[c2], bm = in.bmod2[c2], cs = g.cs[j];
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
  auto hdz_at = [&](int L) { return 1. / (F3(in.scht, L) * dz); };
  auto tp1_at = [&](int L) { return 0.5 * (F3(in.ti, L) + F3(in.te, L)); };
  auto hdzi_at = [&](int L) {
    if (L == 0) return 1.5 * hdz_at(0) - 0.5 * hdz_at(1);
    if (L == nk - 1) return 1.5 * hdz_at(nk - 2) - 0.5 * hdz_at(nk - 3);
    return 0.5 * (hdz_at(L - 1) + hdz_at(L));
  };
  auto tphdz1_at = [&](int L) {
    if (L == nk - 1) return 2. * (2. * tp1_at(nk - 2) - tp1_at(nk - 3)) * hdzi_at(nk - 1) + gmr;
    return 2. * tp1_at(L) * hdzi_at(L) + gmr;
  };
  auto tphdz0_at = [&](int L) {
    if (L == 0) return 2. * (2. * tp1_at(0) - tp1_at(1)) * hdzi_at(0) - gmr;
    return 2. * tp1_at(L - 1) * hdzi_at(L) - gmr;
  };
  auto djint_at = [&](int L) {
    if (L == 0) return 1.5 * F3(W.dj, 0) - 0.5 * F3(W.dj, 1);
    if (L == nk - 1) return 1.5 * F3(W.dj, nk - 2) - 0.5 * F3(W.dj, nk - 3);
    return 0.5 * (F3(W.dj, L - 1) + F3(W.dj, L));
  };
  auto hdzmp_at = [&](int L, double &hm, double &hp) {  // :239-248
    double bd_djbz = F3(OT.bd_djbz, L);
    double divbz = dvbz + bd_djbz / (F3(W.dj, L) * (bz * bz));
    double h = hdz_at(L);
    hm = h - 0.5 * divbz;
    hp = h + 0.5 * divbz;
    hm = hm * (bz * bz);
    hp = hp * (bz * bz);
  };
  const double ss2 = g.shapiro / nsub;
  double w1 = 0., w3 = 0., rprev = 0.;
  double ol_next_cache = 0.;
  for (int k = 0; k <= nk - 2; ++k) {
    size_t p = G.a3(k, i, j), off = (size_t)k * G.P;
    // explicit terms (:181-211, :250)
    double dbd;
    if (k == 0) dbd = F3(W.bdopj, 1) - F3(W.bdopj, 0);
    else if (k <= nk - 3) dbd = (F3(W.bdopj, k + 1) - F3(W.bdopj, k - 1)) / 2.;
    else dbd = F3(W.bdopj, nk - 2) - F3(W.bdopj, nk - 3);
    double bdzdvb = bz / (in.scht[p] * dz) * dbd + dvb * W.bdopj[p];
    double bd_diff = OT.bd_diff[p];
    double bd_op = W.djbz[p] * OT.bd_diffj[p];
    double diffexp = bdzdvb + bd_diff + bd_op;
    double expl = -1. * diffexp;
    double vh = OT.vdotn[p];
    double bd_bvel = OT.bd_bvel[p];
    expl = expl + vh + bd_bvel;
    double opts2 = OT.opts2[p];
    expl = expl - opts2 * dtx2inv * nsub;
    // diffusion, drift and wind coefficients (:252-278)
    double hm, hp;
    hdzmp_at(k, hm, hp);
    double dji = djint_at(k), dji1 = djint_at(k + 1);
    double t0 = tphdz0_at(k), t01 = tphdz0_at(k + 1), t1 = tphdz1_at(k), t11 = tphdz1_at(k + 1);
    double diffp = hm * dji * t0;
    double diffq = -(hp * dji1 * t01 + hm * dji * t1);
    double diffr = hp * dji1 * t11;
    double hdzk = hdz_at(k);
    double wii = .5 * (in.wi[p] + in.wi[p + G.P]);
    double driftp = wii * 0.5 * hdzk;
    double driftq = -wii * 6. / RE;
    double driftr = -wii * 0.5 * hdzk;
    double bu = W.bdotu[p];
    double windp = k == 0 ? bz * (2. * F3(W.bdotu, 0) - F3(W.bdotu, 1)) * 0.5 * hdzk
                          : bz * F3(W.bdotu, k - 1) * 0.5 * hdzk;
    double windq = -bu * dvbz * bz;
    double windr = k < nk - 2 ? -bz * F3(W.bdotu, k + 1) * 0.5 * hdzk
                              : bz * (2. * F3(W.bdotu, nk - 2) - F3(W.bdotu, nk - 3)) * 0.5 * hdzk;
    double pcf = diffp + driftp + windp;
    double qcf = diffq + driftq + windq - dtx2inv * nsub;
    double rcf = diffr + driftr + windr;
    if (k == nk - 2) {  // upper boundary (:280-286)
      double djn = djint_at(nk - 1);
      double ubcb = -(bz * bz) * djn * tphdz0_at(nk - 1);
      double ubca = -(bz * bz) * djn * tphdz1_at(nk - 1);
      qcf = qcf + ubcb / ubca * rcf;
      expl = expl - opflux * rcf / ubca;
      rcf = 0.;
    }
    double x2p, x2d, oploss;
    chem(k, x2p, x2d, oploss);
    xiop2p[p] = x2p;
    xiop2d[p] = x2d;
    qcf = qcf - oploss;
    double qopi = .5 * (in.qop[p] + in.qop[p + G.P]);
    double ne = in.ne[p], xn = in.xnmbar[p];
    double op_prod = qopi + (in.rk19[p] * ne + RK21) * x2p + (in.rk25[p] * ne + RK27) * x2d +
                     (RK18 * x2p + RK24 * x2d) * in.o1[p] * RIO1 * xn;
    expl = expl - op_prod;
    if (k == 0) {  // lower boundary (:301-303)
      double a1, a2, ol1;
      chem(1, a1, a2, ol1);
      qcf = qcf - pcf;
      expl = expl - 2. * pcf * in.qop[p] / (1.5 * oploss - 0.5 * ol1);
      pcf = 0.;
    }
    (void)ol_next_cache;
    // Thomas forward (trsolv_ng, rows 1..nk-1)
    if (k == 0) {
      w1 = qcf;
      w3 = expl / w1;
    } else {
      double w2 = rprev / w1;
      W.w2[G.a3(k - 1, i, j)] = w2;
      w1 = qcf - pcf * w2;
      w3 = (expl - pcf * w3) / w1;
    }
    opout[p] = w3;
    rprev = rcf;
  }
  {  // xiop2p/xiop2d at the top level
    double x2p, x2d, ol;
    chem(nk - 1, x2p, x2d, ol);
    xiop2p[G.a3(nk - 1, i, j)] = x2p;
    xiop2d[G.a3(nk - 1, i, j)] = x2d;
  }
  // back substitution, top extrapolation (:307) and outputs (:337-340)
  auto emit = [&](int k, double x) {
    size_t p = G.a3(k, i, j);
    double nm = DTSMOOTH * op[p] + DTSMOOTH_DIV2 * (optm1[p] + x);
    opout[p] = x > 0. ? x : 0.;
    optm1out[p] = nm > 0. ? nm : 0.;
  };
  double xn = w3;
  for (int k = nk - 2; k >= 0; --k) {
    double x = k == nk - 2 ? xn : opout[G.a3(k, i, j)] - W.w2[G.a3(k, i, j)] * xn;
    if (k == nk - 3) emit(nk - 1, xn * xn / x);
    emit(k, x);
    xn = x;
  }
#undef F3
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

// stage 2: assembly of the 2x2 (UN,VN) system, block Thomas, outputs (:83-321)
struct UvTerms {  // prelude planes
  const double *divx_ztmp, *divy_z, *unms, *vnms, *advu, *advv;
};

__global__ void __launch_bounds__(256) duv_stage2(tgx_ng_grid g, Geo G, tgx_ng_fields in,
                                                  tgx_ng_out out, double *__restrict__ gs,
                                                  UvTerms UT) {
  int i, j;
  if (!column_of(G, i, j)) return;
  const int nk = G.nk;
  const double dz = g.dz, dtx2inv = g.dtx2inv;
  const size_t c2 = G.a2(i, j);
  out.ulbc_nm[c2] = in.ulbc[c2];
  out.vlbc_nm[c2] = in.vlbc[c2];
  out.ulbc[c2] = in.u_lbc[c2];
  out.vlbc[c2] = in.v_lbc[c2];
  double v;
  if (lateral_value(G, in.lonb_un, in.latb_un, 0, i, j, v)) {
    for (int k = 0; k < nk; ++k) {
      size_t p = G.a3(k, i, j);
      double vu, vv;
      lateral_value(G, in.lonb_un, in.latb_un, k, i, j, vu);
      lateral_value(G, in.lonb_vn, in.latb_vn, k, i, j, vv);
      out.un[p] = vu;
      out.vn[p] = vv;
      out.un_nm[p] = DTSMOOTH_DIV2 * (in.un_nm[p] + vu) + DTSMOOTH * in.un[p];
      out.vn_nm[p] = DTSMOOTH_DIV2 * (in.vn_nm[p] + vv) + DTSMOOTH * in.vn[p];
    }
    return;
  }
#define F3(f, k) (f)[G.a3((k), i, j)]
  const double racs = g.racs[j], cor = g.cor[j], tanphi = g.tanphi[j];
  auto zl_at = [&](int L) {
    double zl = F3(UT.divx_ztmp, L);
    if (L < nk - 1) zl = (zl + F3(UT.divx_ztmp, L + 1)) * .5;
    zl = zl * racs;
    return zl * GRAV / g.dlamda;
  };
  auto zp_at = [&](int L) {
    double zp = F3(UT.divy_z, L);
    if (L < nk - 1) zp = (zp + F3(UT.divy_z, L + 1)) * .5;
    return zp * GRAV * RE_INV / g.dphi;
  };
  auto iavg = [&](const double *f, int L) {
    size_t p = G.a3(L, i, j);
    return L < nk - 1 ? .5 * (f[p] + f[p + G.P]) : 1.5 * f[p] - .5 * f[p - G.P];
  };
  // eddy viscosity (:143-152)
  auto e1b = [&](int L) -> double {
    double e = P0 * F3(in.scht, L) / GRAV;
    return e * g.expz[L] * g.xmue[L];
  };
  auto e1 = [&](int L) -> double {
    if (L <= nk - 2) return e1b(L);
    double a = e1b(nk - 2), b = e1b(nk - 3);
    return sqrt(a * a * a / b);
  };
  auto evf = [&](int L) -> double {
    if (L >= 1) return sqrt(e1(L - 1) * e1(L));
    double e0 = e1(0);
    return e0 * e0 / sqrt(e1(0) * e1(1));
  };
  auto gg_at = [&](int L) { return GRAV * (F3(in.km, L) + evf(L)) / (P0 * F3(in.schti, L) * (dz * dz)); };
  double ulbc = in.u_lbc[c2], vlbc = in.v_lbc[c2];
  // block Thomas state (duv_ng.F90:218-243): beta(k-1), yy(k-1), rr(k-1)
  double be11 = 0, be12 = 0, be21 = 0, be22 = 0, y1 = 0, y2 = 0;
  double r11p = 0, r12p = 0, r21p = 0, r22p = 0;
  double gk = gg_at(0);
  for (int k = 0; k <= nk - 2; ++k) {
    size_t p = G.a3(k, i, j), off = (size_t)k * G.P;
    double gk1 = gg_at(k + 1);
    double g_1 = gk1;  // k <= nk-2
    double wi = iavg(in.w_upd, k);
    double dwdz = wi / (2. * dz);
    dwdz = dwdz * g.expz[k];
    double lxxi = iavg(in.lxx, k), lyyi = iavg(in.lyy, k), lxyi = iavg(in.lxy, k), lyxi = iavg(in.lyx, k);
    double uii = iavg(in.ui, k), vii = iavg(in.vi, k);
    double udc = out.Fe[p] + lxxi * uii + lxyi * vii;
    double vdc = out.Fn[p] + lyyi * vii - lyxi * uii;
    double unms = UT.unms[p], vnms = UT.vnms[p], advu = UT.advu[p], advv = UT.advv[p];
    double s1 = dtx2inv * unms + in.hdu[p] + udc - advu - zl_at(k);
    double s2 = dtx2inv * vnms + in.hdv[p] + vdc - advv - zp_at(k);
    s1 = s1 * g.expz[k];
    s2 = s2 * g.expz[k];
    double p11 = -(gk + dwdz), p12 = 0., p21 = 0., p22 = p11;
    double r11 = -(g_1 - dwdz), r12 = 0., r21 = 0., r22 = r11;
    double q11 = gk + g_1;
    double qa = dtx2inv + lxxi;
    qa = qa * g.expz[k];
    double qq11 = q11 + qa;
    qa = dtx2inv + lyyi;
    qa = qa * g.expz[k];
    double qq22 = q11 + qa;
    double q12 = cor + F3(in.un, k) / RE * tanphi;
    double q21 = q12;
    qa = q12 - lxyi;
    qa = qa * g.expz[k];
    double qq12 = -qa;
    qa = q21 - lyxi;
    qa = qa * g.expz[k];
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
      size_t pm = G.a3(k - 1, i, j), NP = (size_t)nk * G.P;
      gs[pm] = g11; gs[pm + NP] = g21; gs[pm + 2 * NP] = g12; gs[pm + 3 * NP] = g22;
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
    out.un[p] = y1;
    out.vn[p] = y2;
    r11p = r11; r12p = r12; r21p = r21; r22p = r22;
    gk = gk1;
  }
  // back substitution (:244-250) and outputs (:320-321)
  auto emit = [&](int k, double u, double vv) {
    size_t p = G.a3(k, i, j);
    out.un[p] = u;
    out.vn[p] = vv;
    out.un_nm[p] = DTSMOOTH_DIV2 * (in.un_nm[p] + u) + DTSMOOTH * in.un[p];
    out.vn_nm[p] = DTSMOOTH_DIV2 * (in.vn_nm[p] + vv) + DTSMOOTH * in.vn[p];
  };
  double xn1 = y1, xn2 = y2, xp1 = 0., xp2 = 0.;
  const size_t NP = (size_t)nk * G.P;
  for (int k = nk - 2; k >= 0; --k) {
    size_t p = G.a3(k, i, j);
    double x1, x2;
    if (k == nk - 2) { x1 = xn1; x2 = xn2; }
    else {
      x1 = out.un[p] - gs[p] * xn1 - gs[p + 2 * NP] * xn2;
      x2 = out.vn[p] - gs[p + NP] * xn1 - gs[p + 3 * NP] * xn2;
    }
    if (k == nk - 3) emit(nk - 1, 2 * xn1 - x1, 2 * xn2 - x2);
    emit(k, x1, x2);
    xp1 = xn1; xp2 = xn2;
    xn1 = x1; xn2 = x2;
  }
  (void)xp1; (void)xp2;
#undef F3
}

}  // namespace ng
}  // namespace tgx

// ---------------------------------------------------------------------------
// Launcher used by step_api.cpp.  bufs holds the workspace planes (and for
// O+ the input/output state of the sub-step) in the order documented there.
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
  dim3 grid(nb), gridp(nbp), block(256);
  auto D = [&](int n) { return (double *)bufs[n]; };
  auto job = [](int kind, const double *f, double *o, double sh) {
    StencilJob jb;
    jb.kind = kind; jb.f = f; jb.out = o; jb.s = sh;
    return jb;
  };
  switch (which) {
    case 0:  // dt: bufs = {w2, hadv_tn, smooth_tn_nm}
      hipLaunchKernelGGL(dt_kernel, grid, block, 0, s, g, G, in, out, D(0), D(1), D(2));
      break;
    case 1:
    case 2: {  // minors: bufs = {w2a, w2b, ftm1s[4], hadvec[4]} (species 0..3)
      MinorTerms MT;
      for (int q = 0; q < 4; ++q) { MT.ftm1s[q] = D(2 + q); MT.hadvec[q] = D(6 + q); }
      if (which == 1)
        hipLaunchKernelGGL((minor_kernel<2, 3>), grid, block, 0, s, g, G, in, out, D(0), D(1), MT);
      else
        hipLaunchKernelGGL((minor_kernel<1, 0>), grid, block, 0, s, g, G, in, out, D(0), D(1), MT);
      break;
    }
    case 3:
    case 4:
    case 5:
    case 9: {  // O+: bufs = {dj bdotu bvel tpop diffj djbz opbm2 bdopj w2, op, optm1, opout,
               //           optm1out, bd_diff bd_diffj bd_bvel vdotn bd_djbz opts2}
      OpWork W{D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8)};
      const double *op = D(9), *optm1 = D(10);
      double *opout = D(11), *optm1out = D(12);
      OpTerms OT{D(13), D(14), D(15), D(16), D(17), D(18)};
      if (which == 3) {
        hipLaunchKernelGGL(oplus_stage1, grid, block, 0, s, g, G, in, op, W, out.Fe, out.Fn);
      } else if (which == 4) {
        hipLaunchKernelGGL(oplus_stage2, grid, block, 0, s, g, G, in, W);
      } else if (which == 9) {
        StencilJobs J;
        J.n = 6;
        J.job[0] = job(ST_BDOTDH, W.bdopj, D(13), 0.);
        J.job[1] = job(ST_BDOTDH, W.diffj, D(14), 0.);
        J.job[2] = job(ST_BDOTDH, W.bvel, D(15), 0.);
        J.job[3] = job(ST_VDOTN, W.opbm2, D(16), 0.);
        J.job[4] = job(ST_BDOTDH, W.djbz, D(17), 0.);
        J.job[5] = job(ST_SMOOTH, optm1, D(18), g.shapiro / g.nstep_sub);
        hipLaunchKernelGGL(stencil_kernel, gridp, block, 0, s, g, G, in, J);
      } else {
        const double *lonb = in.lonb_op + (size_t)sub * 2 * g.nk * (size_t)G.nlatd;
        const double *latb = in.latb_op + (size_t)sub * 2 * g.nk * (size_t)G.ld;
        hipLaunchKernelGGL(oplus_stage3, grid, block, 0, s, g, G, in, op, optm1, lonb, latb, W, OT,
                           opout, optm1out, out.xiop2p, out.xiop2d);
      }
      break;
    }
    case 6:  // duv stage 1: bufs = {ztmp}
      hipLaunchKernelGGL(duv_stage1, grid, block, 0, s, g, G, in, (const double *)out.tn, D(0));
      break;
    case 7: {  // duv stage 2: bufs = {ztmp, gamma, divx_ztmp, divy_z, unms, vnms, advu, advv}
      UvTerms UT{D(2), D(3), D(4), D(5), D(6), D(7)};
      hipLaunchKernelGGL(duv_stage2, grid, block, 0, s, g, G, in, out, D(1), UT);
      break;
    }
    case 8: {  // comp: bufs = {gamma, o2adv, o1adv, headv, o2s, o1s, hes}
      CompTerms CT{D(1), D(2), D(3), D(4), D(5), D(6)};
      hipLaunchKernelGGL(comp_kernel, grid, block, 0, s, g, G, in, out, D(0), CT);
      break;
    }
    case 10: {  // dt prelude: bufs = {w2, hadv_tn, smooth_tn_nm}
      StencilJobs J;
      J.n = 2;
      J.job[0] = job(ST_ADVEC, in.tn, D(1), 0.);
      J.job[1] = job(ST_SMOOTH, in.tn_nm, D(2), g.shapiro);
      hipLaunchKernelGGL(stencil_kernel, gridp, block, 0, s, g, G, in, J);
      break;
    }
    case 11: {  // minors prelude (species 0..3 = NO, N4S, N2D, AR)
      StencilJobs J;
      J.n = 8;
      const double *fc[4] = {in.no, in.n4s, in.n2d, in.ar};
      const double *fm[4] = {in.no_nm, in.n4s_nm, in.n2d_nm, in.ar_nm};
      for (int q = 0; q < 4; ++q) {
        J.job[q] = job(ST_SMOOTH, fm[q], D(2 + q), g.shapiro);
        J.job[4 + q] = job(ST_ADVEC, fc[q], D(6 + q), 0.);
      }
      hipLaunchKernelGGL(stencil_kernel, gridp, block, 0, s, g, G, in, J);
      break;
    }
    case 12: {  // comp prelude
      StencilJobs J;
      J.n = 6;
      J.job[0] = job(ST_ADVEC, in.o2, D(1), 0.);
      J.job[1] = job(ST_ADVEC, in.o1, D(2), 0.);
      J.job[2] = job(ST_ADVEC, in.he, D(3), 0.);
      J.job[3] = job(ST_SMOOTH, in.o2_nm, D(4), g.shapiro);
      J.job[4] = job(ST_SMOOTH, in.o1_nm, D(5), g.shapiro);
      J.job[5] = job(ST_SMOOTH, in.he_nm, D(6), g.shapiro);
      hipLaunchKernelGGL(stencil_kernel, gridp, block, 0, s, g, G, in, J);
      break;
    }
    case 13: {  // duv prelude (after stage 1 wrote ztmp)
      StencilJobs J;
      J.n = 6;
      J.job[0] = job(ST_DIVX, D(0), D(2), 0.);
      J.job[1] = job(ST_DIVY, in.z, D(3), 0.);
      J.job[2] = job(ST_SMOOTH, in.un_nm, D(4), g.shapiro);
      J.job[3] = job(ST_SMOOTH, in.vn_nm, D(5), g.shapiro);
      J.job[4] = job(ST_ADVEC, in.un, D(6), 0.);
      J.job[5] = job(ST_ADVEC, in.vn, D(7), 0.);
      hipLaunchKernelGGL(stencil_kernel, gridp, block, 0, s, g, G, in, J);
      break;
    }
    default:
      return tgx_set_error(1, "tgx_ng_launch: unknown kernel %d", which);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tgx_set_error(2, "launch %d: %s", which, hipGetErrorString(e));
  return 0;
}
