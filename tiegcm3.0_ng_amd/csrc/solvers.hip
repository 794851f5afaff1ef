// Plain column solvers on the reference's level-fastest layout:
//   trsolv / trsolv_ng  (scalar tridiagonal, trsolv.F:31-101, trsolv_ng.F90:14-26)
//   blktri              (2x2 block tridiagonal, duv.F:801-908)
//   comp 3x3 elimination (comp_ng.F90:281-326 on pre-assembled blocks)
//
// Two schemes for the scalar system:
//   * Thomas, one thread per column, in the reference's operation order
//     (bit-compatible with the Fortran when FP contraction is off);
//   * parallel cyclic reduction (PCR), one wavefront (or half) per column,
//     coefficients held in registers and exchanged with cross-lane shuffles.
// and a mixed-precision PCR: FP32 reduction, FP64 residual r = f - A x,
// iterative refinement sweeps that reuse the FP32 reduction multipliers.
#include "tgx_common.h"

namespace tgx {

// ---------------------------------------------------------------------------
// Thomas, thread per column.  Column c of the batch starts at base + c*ldk;
// rows k1..k2 (0-based, inclusive) are solved.  w2 (the c/w1 ratios) go to a
// scratch array of the same shape as x; w3 is parked in x and overwritten by
// the back substitution.
__global__ void __launch_bounds__(256) thomas_kfast_f64(
    const double *__restrict__ a, const double *__restrict__ b,
    const double *__restrict__ c, const double *__restrict__ f, double *__restrict__ x,
    double *__restrict__ w2s, int ldk, int k1, int k2, int col0, int ncol) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol) return;
  size_t o = (size_t)(col0 + t) * (size_t)ldk;
  // the ratios of batch column t: w2s holds ncol * ldk doubles
  double *__restrict__ w2c = w2s + (size_t)t * (size_t)ldk;
  double w1 = b[o + k1];
  double w3 = f[o + k1] / w1;
  x[o + k1] = w3;
  for (int k = k1 + 1; k <= k2; ++k) {
    double w2 = c[o + k - 1] / w1;
    w2c[k - 1] = w2;
    double ak = a[o + k];
    w1 = b[o + k] - ak * w2;
    w3 = (f[o + k] - ak * w3) / w1;
    x[o + k] = w3;
  }
  double xn = w3;
  for (int k = k2 - 1; k >= k1; --k) {
    double xk = x[o + k] - w2c[k] * xn;
    x[o + k] = xk;
    xn = xk;
  }
}

// ---------------------------------------------------------------------------
// PCR helpers.  A segment of SEG lanes owns one column; lane l holds rows
// r = l*NPL + e (e < NPL).  Rows outside [0, n) are identity rows.
template <int SEG>
__device__ __forceinline__ double shfl_from(double v, int src_in_seg, int lane_in_seg) {
  (void)lane_in_seg;
  return __shfl(v, src_in_seg, SEG);
}

template <typename T, int SEG, int NPL>
struct PcrRows {
  T a[NPL], b[NPL], c[NPL], d[NPL];
};

// One PCR step of stride s on blocked rows; also records the multipliers
// (k1, k2) so the same reduction can be replayed on a new right-hand side.
template <typename T, int SEG, int NPL>
__device__ __forceinline__ void pcr_step(PcrRows<T, SEG, NPL> &R, int s, int lane, T *m1, T *m2) {
  T na[NPL], nb[NPL], nc[NPL], nd[NPL];
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    // neighbour r - s
    int dn = e - s;  // element index in the source lane space
    int lo_lane = lane + (dn >= 0 ? 0 : -((-dn + NPL - 1) / NPL));
    int lo_e = dn - (lo_lane - lane) * NPL;
    int up = e + s;
    int hi_lane = lane + up / NPL;
    int hi_e = up % NPL;
    T am = 0, bm = 1, cm = 0, dm = 0, ap = 0, bp = 1, cp = 0, dp = 0;
    // Every lane must take part in the shuffles; gather all NPL candidates.
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      T ta = __shfl(R.a[q], lo_lane, SEG), tb = __shfl(R.b[q], lo_lane, SEG);
      T tc = __shfl(R.c[q], lo_lane, SEG), td = __shfl(R.d[q], lo_lane, SEG);
      if (q == lo_e) { am = ta; bm = tb; cm = tc; dm = td; }
      T ua = __shfl(R.a[q], hi_lane, SEG), ub = __shfl(R.b[q], hi_lane, SEG);
      T uc = __shfl(R.c[q], hi_lane, SEG), ud = __shfl(R.d[q], hi_lane, SEG);
      if (q == hi_e) { ap = ua; bp = ub; cp = uc; dp = ud; }
    }
    if (lo_lane < 0) { am = 0; bm = 1; cm = 0; dm = 0; }
    if (hi_lane >= SEG) { ap = 0; bp = 1; cp = 0; dp = 0; }
    T k1 = R.a[e] / bm;
    T k2 = R.c[e] / bp;
    if (m1) { m1[e] = k1; m2[e] = k2; }
    na[e] = -am * k1;
    nc[e] = -cp * k2;
    nb[e] = R.b[e] - cm * k1 - ap * k2;
    nd[e] = R.d[e] - dm * k1 - dp * k2;
  }
#pragma unroll
  for (int e = 0; e < NPL; ++e) { R.a[e] = na[e]; R.b[e] = nb[e]; R.c[e] = nc[e]; R.d[e] = nd[e]; }
}

// Replay of a recorded reduction on a new right-hand side d.
template <typename T, int SEG, int NPL>
__device__ __forceinline__ void pcr_rhs_step(T *d, int s, int lane, const T *m1, const T *m2) {
  T nd[NPL];
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    int dn = e - s;
    int lo_lane = lane + (dn >= 0 ? 0 : -((-dn + NPL - 1) / NPL));
    int lo_e = dn - (lo_lane - lane) * NPL;
    int up = e + s;
    int hi_lane = lane + up / NPL;
    int hi_e = up % NPL;
    T dm = 0, dp = 0;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      T td = __shfl(d[q], lo_lane, SEG);
      if (q == lo_e) dm = td;
      T ud = __shfl(d[q], hi_lane, SEG);
      if (q == hi_e) dp = ud;
    }
    if (lo_lane < 0) dm = 0;
    if (hi_lane >= SEG) dp = 0;
    nd[e] = d[e] - dm * m1[e] - dp * m2[e];
  }
#pragma unroll
  for (int e = 0; e < NPL; ++e) d[e] = nd[e];
}

template <int SEG, int NPL>
constexpr int pcr_nsteps() {
  int n = SEG * NPL, s = 0;
  while ((1 << s) < n) ++s;
  return s;
}

// FP64 PCR: the whole column in registers, coalesced loads (a column is
// contiguous in the level-fastest layout).
template <int SEG, int NPL>
__global__ void __launch_bounds__(256) pcr_kfast_f64(
    const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    const double *__restrict__ f, double *__restrict__ x, int ldk, int k1, int k2, int col0,
    int ncol) {
  const int lane = threadIdx.x % SEG;
  const int col = (blockIdx.x * blockDim.x + threadIdx.x) / SEG;
  const bool live = col < ncol;
  const int n = k2 - k1 + 1;
  size_t o = (size_t)(col0 + (live ? col : 0)) * (size_t)ldk + k1;
  PcrRows<double, SEG, NPL> R;
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    int r = lane * NPL + e;
    bool in = live && r < n;
    R.a[e] = (in && r > 0) ? a[o + r] : 0.;
    R.b[e] = in ? b[o + r] : 1.;
    R.c[e] = (in && r < n - 1) ? c[o + r] : 0.;
    R.d[e] = in ? f[o + r] : 0.;
  }
  constexpr int NS = pcr_nsteps<SEG, NPL>();
#pragma unroll
  for (int st = 0; st < NS; ++st) pcr_step<double, SEG, NPL>(R, 1 << st, lane, nullptr, nullptr);
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    int r = lane * NPL + e;
    if (live && r < n) x[o + r] = R.d[e] / R.b[e];
  }
}

// Mixed precision: FP32 PCR reduction of the FP64 system, FP64 residual with
// the FP64 coefficients kept in registers, nsweep refinement sweeps replaying
// the FP32 reduction on the residual.
template <int SEG, int NPL>
__global__ void __launch_bounds__(256) pcr_kfast_mixed(
    const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    const double *__restrict__ f, double *__restrict__ x, int ldk, int k1, int k2, int col0,
    int ncol, int nsweep) {
  const int lane = threadIdx.x % SEG;
  const int col = (blockIdx.x * blockDim.x + threadIdx.x) / SEG;
  const bool live = col < ncol;
  const int n = k2 - k1 + 1;
  size_t o = (size_t)(col0 + (live ? col : 0)) * (size_t)ldk + k1;
  double A[NPL], B[NPL], Cc[NPL], Fv[NPL];
  PcrRows<float, SEG, NPL> R;
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    int r = lane * NPL + e;
    bool in = live && r < n;
    A[e] = (in && r > 0) ? a[o + r] : 0.;
    B[e] = in ? b[o + r] : 1.;
    Cc[e] = (in && r < n - 1) ? c[o + r] : 0.;
    Fv[e] = in ? f[o + r] : 0.;
    R.a[e] = (float)A[e]; R.b[e] = (float)B[e]; R.c[e] = (float)Cc[e]; R.d[e] = (float)Fv[e];
  }
  constexpr int NS = pcr_nsteps<SEG, NPL>();
  float m1[NS][NPL], m2[NS][NPL];
#pragma unroll
  for (int st = 0; st < NS; ++st) pcr_step<float, SEG, NPL>(R, 1 << st, lane, m1[st], m2[st]);
  float binv[NPL];
  double X[NPL];
#pragma unroll
  for (int e = 0; e < NPL; ++e) { binv[e] = 1.f / R.b[e]; X[e] = (double)(R.d[e] * binv[e]); }
  for (int sw = 0; sw < nsweep; ++sw) {
    // FP64 residual: r = f - (a x[r-1] + b x[r] + c x[r+1])
    double xl[NPL], xr[NPL];
#pragma unroll
    for (int e = 0; e < NPL; ++e) {
      int dn = e - 1, lo_lane = lane + (dn >= 0 ? 0 : -1), lo_e = dn >= 0 ? dn : NPL - 1;
      int up = e + 1, hi_lane = lane + up / NPL, hi_e = up % NPL;
      double vl = 0., vr = 0.;
#pragma unroll
      for (int q = 0; q < NPL; ++q) {
        double tl = __shfl(X[q], lo_lane, SEG);
        if (q == lo_e) vl = tl;
        double tr = __shfl(X[q], hi_lane, SEG);
        if (q == hi_e) vr = tr;
      }
      xl[e] = lo_lane < 0 ? 0. : vl;
      xr[e] = hi_lane >= SEG ? 0. : vr;
    }
    float d[NPL];
#pragma unroll
    for (int e = 0; e < NPL; ++e) {
      double res = Fv[e] - (A[e] * xl[e] + B[e] * X[e] + Cc[e] * xr[e]);
      d[e] = (float)res;
    }
#pragma unroll
    for (int st = 0; st < NS; ++st) pcr_rhs_step<float, SEG, NPL>(d, 1 << st, lane, m1[st], m2[st]);
#pragma unroll
    for (int e = 0; e < NPL; ++e) X[e] += (double)(d[e] * binv[e]);
  }
#pragma unroll
  for (int e = 0; e < NPL; ++e) {
    int r = lane * NPL + e;
    if (live && r < n) x[o + r] = X[e];
  }
}

// ---------------------------------------------------------------------------
// 2x2 block Thomas (duv.F:830-906), thread per column.  a,b,c: (2,2,ldk,col),
// f,x: (2,ldk,col).  gamma goes to scratch (4 per level); y is parked in x.
__global__ void __launch_bounds__(256) blktri2_kfast_f64(
    const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    const double *__restrict__ f, double *__restrict__ x, double *__restrict__ gs, int ldk,
    int nrow, int col0, int ncol) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol) return;
  size_t o = (size_t)(col0 + t) * (size_t)ldk;
  const double *A = a + 4 * o, *B = b + 4 * o, *C = c + 4 * o, *F = f + 2 * o;
  double *X = x + 2 * o, *G = gs + 4 * o;
  double det = B[0] * B[3] - B[2] * B[1];
  double be11 = B[3] / det, be12 = -B[2] / det, be21 = -B[1] / det, be22 = B[0] / det;
  double y1 = be11 * F[0] + be12 * F[1];
  double y2 = be21 * F[0] + be22 * F[1];
  X[0] = y1; X[1] = y2;
  for (int k = 1; k < nrow; ++k) {
    const double *Cm = C + 4 * (k - 1), *Ak = A + 4 * k, *Bk = B + 4 * k, *Fk = F + 2 * k;
    double g11 = be11 * Cm[0] + be12 * Cm[1];
    double g12 = be11 * Cm[2] + be12 * Cm[3];
    double g21 = be21 * Cm[0] + be22 * Cm[1];
    double g22 = be21 * Cm[2] + be22 * Cm[3];
    double *Gm = G + 4 * (k - 1);
    Gm[0] = g11; Gm[1] = g21; Gm[2] = g12; Gm[3] = g22;
    double h11 = Bk[0] - Ak[0] * g11 - Ak[2] * g21;
    double h12 = Bk[2] - Ak[0] * g12 - Ak[2] * g22;
    double h21 = Bk[1] - Ak[1] * g11 - Ak[3] * g21;
    double h22 = Bk[3] - Ak[1] * g12 - Ak[3] * g22;
    double dk = h11 * h22 - h12 * h21;
    be11 = h22 / dk; be12 = -h12 / dk; be21 = -h21 / dk; be22 = h11 / dk;
    double x1 = Fk[0] - Ak[0] * y1 - Ak[2] * y2;
    double x2 = Fk[1] - Ak[1] * y1 - Ak[3] * y2;
    y1 = be11 * x1 + be12 * x2;
    y2 = be21 * x1 + be22 * x2;
    X[2 * k] = y1; X[2 * k + 1] = y2;
  }
  double xn1 = y1, xn2 = y2;
  for (int k = nrow - 2; k >= 0; --k) {
    const double *Gk = G + 4 * k;
    double x1 = X[2 * k] - Gk[0] * xn1 - Gk[2] * xn2;
    double x2 = X[2 * k + 1] - Gk[1] * xn1 - Gk[3] * xn2;
    X[2 * k] = x1; X[2 * k + 1] = x2;
    xn1 = x1; xn2 = x2;
  }
}

// 3x3 adjugate / determinant inverse (matutil.F90:57-73, :110-118), column-major.
__device__ __forceinline__ void matinv3(const double *A, double *Bm) {
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

// comp's 3x3 block elimination (comp_ng.F90:281-326), thread per column.
// p,q,r: (3,3,ldk,col), f,x: (3,ldk,col); gama to scratch (9 per level),
// zz parked in x.
__global__ void __launch_bounds__(256) blktri3_kfast_f64(
    const double *__restrict__ p, const double *__restrict__ q, const double *__restrict__ r,
    const double *__restrict__ f, double *__restrict__ x, double *__restrict__ gs, int ldk,
    int nrow, int col0, int ncol) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol) return;
  size_t o = (size_t)(col0 + t) * (size_t)ldk;
  const double *P = p + 9 * o, *Q = q + 9 * o, *Rr = r + 9 * o, *F = f + 3 * o;
  double *X = x + 3 * o, *G = gs + 9 * o;
  double gam[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, zz[3] = {0, 0, 0};
  for (int k = 0; k < nrow; ++k) {
    double pk[9], qk[9], rk[9], w[9], v[3];
#pragma unroll
    for (int n = 0; n < 9; ++n) { pk[n] = P[9 * k + n]; qk[n] = Q[9 * k + n]; rk[n] = Rr[9 * k + n]; }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int i = 0; i < 3; ++i) qk[i + 3 * m] = qk[i + 3 * m] - pk[i + 3 * n] * gam[n + 3 * m];
    matinv3(qk, w);
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = F[3 * k + i];
    double ng[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
#pragma unroll
      for (int i = 0; i < 3; ++i) v[i] = v[i] - pk[i + 3 * m] * zz[m];
#pragma unroll
      for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int i = 0; i < 3; ++i) ng[i + 3 * m] = ng[i + 3 * m] + w[i + 3 * n] * rk[n + 3 * m];
    }
    double nz[3] = {0, 0, 0};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int i = 0; i < 3; ++i) nz[i] = nz[i] + w[i + 3 * m] * v[m];
#pragma unroll
    for (int n = 0; n < 9; ++n) { gam[n] = ng[n]; G[9 * k + n] = ng[n]; }
#pragma unroll
    for (int i = 0; i < 3; ++i) { zz[i] = nz[i]; X[3 * k + i] = nz[i]; }
  }
  double un[3] = {0, 0, 0};
  for (int k = nrow - 1; k >= 0; --k) {
    double u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u[i] = X[3 * k + i];
      if (k < nrow - 1)
#pragma unroll
        for (int m = 0; m < 3; ++m) u[i] = u[i] - G[9 * k + i + 3 * m] * un[m];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) { X[3 * k + i] = u[i]; un[i] = u[i]; }
  }
}

}  // namespace tgx

// ---------------------------------------------------------------------------
// Launchers (called from api.cpp)
namespace tgx {
// ---------------------------------------------------------------------------
// Mixed-precision block Thomas for N x N block rows (N = 1, 2, 3):
// A x(k-1) + B x(k) + C x(k+1) = F.  The elimination runs in FP32 (the
// coefficients rounded to float, Schur complements and multipliers in float);
// each block pivot is inverted in FP64 and applied in FP64; the residual
// r = F - A x is formed in FP64 from the FP64 coefficients and nsweep
// refinement sweeps solve for the correction with the stored FP32
// factorisation.  Layout: element e (row-major n*N+m) of A for (col, k) at
// A[e*pstride + col*nk + k]; rows 0..nrow-1; columns with valid[col] == 0
// are skipped.  Work per column and row: N*N floats (multipliers), N*N
// doubles (pivot inverses), 2N doubles.
template <int N>
__device__ __forceinline__ void inv_nd(const double *M, double *R) {
  if (N == 1) {
    R[0] = 1. / M[0];
  } else if (N == 2) {
    double d = M[0] * M[3] - M[1] * M[2];
    R[0] = M[3] / d; R[1] = -M[1] / d; R[2] = -M[2] / d; R[3] = M[0] / d;
  } else {
    double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
    double r = 1. / (M[0] * c00 + M[1] * c01 + M[2] * c02);
    R[0] = c00 * r; R[3] = c01 * r; R[6] = c02 * r;
    R[1] = (M[2] * M[7] - M[1] * M[8]) * r; R[4] = (M[0] * M[8] - M[2] * M[6]) * r;
    R[7] = (M[1] * M[6] - M[0] * M[7]) * r;
    R[2] = (M[1] * M[5] - M[2] * M[4]) * r; R[5] = (M[2] * M[3] - M[0] * M[5]) * r;
    R[8] = (M[0] * M[4] - M[1] * M[3]) * r;
  }
}

template <int N>
__global__ void __launch_bounds__(256) blk_mixed_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                        const double *__restrict__ C, const double *__restrict__ F,
                                                        double *__restrict__ X, long pstride, int nk, int nrow,
                                                        const unsigned char *__restrict__ valid, long ncol,
                                                        int nsweep, float *__restrict__ gam,
                                                        double *__restrict__ ginv, double *__restrict__ yv) {
  long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol || !valid[col]) return;
  constexpr int NN = N * N;
  const size_t base = (size_t)col * nk;
  auto el = [&](const double *M, int e, int k) { return M[(size_t)e * pstride + base + k]; };
  float *Gm = gam + (size_t)col * nrow * NN;      // multipliers gamma_k = Gamma_k^-1 C_k (float)
  double *Gi = ginv + (size_t)col * nrow * NN;    // pivot inverses Gamma_k^-1 (double)
  double *Y = yv + (size_t)col * nrow * 2 * N;    // y and the right-hand side of a sweep
  // factorisation: Gamma_k = B_k - A_k gamma_{k-1} in FP32
  float gp[NN];
  for (int e = 0; e < NN; ++e) gp[e] = 0.f;
  for (int k = 0; k < nrow; ++k) {
    float Af[NN], Gf[NN];
    for (int e = 0; e < NN; ++e) { Af[e] = (float)el(A, e, k); Gf[e] = (float)el(B, e, k); }
    for (int n = 0; n < N; ++n)
      for (int m = 0; m < N; ++m) {
        float t = Gf[n * N + m];
        for (int l = 0; l < N; ++l) t -= Af[n * N + l] * gp[l * N + m];
        Gf[n * N + m] = t;
      }
    double Gd[NN], Id[NN];
    for (int e = 0; e < NN; ++e) Gd[e] = (double)Gf[e];
    inv_nd<N>(Gd, Id);  // the block pivot inverted in FP64
    for (int e = 0; e < NN; ++e) Gi[(size_t)k * NN + e] = Id[e];
    float Cf[NN];
    for (int e = 0; e < NN; ++e) Cf[e] = (float)el(C, e, k);
    for (int n = 0; n < N; ++n)
      for (int m = 0; m < N; ++m) {
        double t = 0.;
        for (int l = 0; l < N; ++l) t += Id[n * N + l] * (double)Cf[l * N + m];
        gp[n * N + m] = (float)t;
      }
    for (int e = 0; e < NN; ++e) Gm[(size_t)k * NN + e] = gp[e];
  }
  // solve with the factorisation: rhs in Y[.. + N + n], result into X (add != 0: X += solution)
  auto solve = [&](bool add) {
    float yp[N];
    for (int n = 0; n < N; ++n) yp[n] = 0.f;
    for (int k = 0; k < nrow; ++k) {
      float v[N];
      for (int n = 0; n < N; ++n) {
        float t = (float)Y[((size_t)k * 2 + 1) * N + n];
        for (int l = 0; l < N; ++l) t -= (float)el(A, n * N + l, k) * yp[l];
        v[n] = t;
      }
      for (int n = 0; n < N; ++n) {
        double t = 0.;
        for (int l = 0; l < N; ++l) t += Gi[(size_t)k * NN + n * N + l] * (double)v[l];
        yp[n] = (float)t;
        Y[(size_t)k * 2 * N + n] = t;
      }
    }
    float xn[N];
    for (int n = 0; n < N; ++n) xn[n] = 0.f;
    for (int k = nrow - 1; k >= 0; --k) {
      float xv[N];
      for (int n = 0; n < N; ++n) {
        float t = (float)Y[(size_t)k * 2 * N + n];
        for (int l = 0; l < N; ++l) t -= Gm[(size_t)k * NN + n * N + l] * xn[l];
        xv[n] = t;
      }
      for (int n = 0; n < N; ++n) {
        double* xo = X + (size_t)n * pstride + base + k;
        *xo = add ? *xo + (double)xv[n] : (double)xv[n];
        xn[n] = xv[n];
      }
    }
  };
  for (int k = 0; k < nrow; ++k)
    for (int n = 0; n < N; ++n) Y[((size_t)k * 2 + 1) * N + n] = el(F, n, k);
  solve(false);
  for (int sw = 0; sw < nsweep; ++sw) {
    // FP64 residual r = F - (A x(k-1) + B x(k) + C x(k+1))
    for (int k = 0; k < nrow; ++k)
      for (int n = 0; n < N; ++n) {
        double r = el(F, n, k);
        for (int l = 0; l < N; ++l) {
          if (k > 0) r -= el(A, n * N + l, k) * X[(size_t)l * pstride + base + k - 1];
          r -= el(B, n * N + l, k) * X[(size_t)l * pstride + base + k];
          if (k + 1 < nrow) r -= el(C, n * N + l, k) * X[(size_t)l * pstride + base + k + 1];
        }
        Y[((size_t)k * 2 + 1) * N + n] = r;
      }
    solve(true);
  }
}
}  // namespace tgx

extern "C" int tgx_launch_thomas_f64(const double *a, const double *b, const double *c,
                                     const double *f, double *x, double *w2s, int ldk, int k1,
                                     int k2, int col0, int ncol, hipStream_t s) {
  if (ncol <= 0 || k2 < k1) return 0;
  int nb = (ncol + 255) / 256;
  hipLaunchKernelGGL(tgx::thomas_kfast_f64, dim3(nb), dim3(256), 0, s, a, b, c, f, x, w2s, ldk,
                     k1, k2, col0, ncol);
  TGX_HIP(hipGetLastError());
  return 0;
}

template <int SEG, int NPL>
static int launch_pcr(const double *a, const double *b, const double *c, const double *f,
                      double *x, int ldk, int k1, int k2, int col0, int ncol, int nsweep,
                      bool mixed, hipStream_t s) {
  int cols_per_block = 256 / SEG;
  int nb = (ncol + cols_per_block - 1) / cols_per_block;
  if (mixed)
    hipLaunchKernelGGL((tgx::pcr_kfast_mixed<SEG, NPL>), dim3(nb), dim3(256), 0, s, a, b, c, f,
                       x, ldk, k1, k2, col0, ncol, nsweep);
  else
    hipLaunchKernelGGL((tgx::pcr_kfast_f64<SEG, NPL>), dim3(nb), dim3(256), 0, s, a, b, c, f, x,
                       ldk, k1, k2, col0, ncol);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_launch_pcr(const double *a, const double *b, const double *c,
                              const double *f, double *x, int ldk, int k1, int k2, int col0,
                              int ncol, int nsweep, int mixed, hipStream_t s) {
  if (ncol <= 0 || k2 < k1) return 0;
  int n = k2 - k1 + 1;
  if (n <= 32) return launch_pcr<32, 1>(a, b, c, f, x, ldk, k1, k2, col0, ncol, nsweep, mixed, s);
  if (n <= 64) return launch_pcr<64, 1>(a, b, c, f, x, ldk, k1, k2, col0, ncol, nsweep, mixed, s);
  if (n <= 128) return launch_pcr<64, 2>(a, b, c, f, x, ldk, k1, k2, col0, ncol, nsweep, mixed, s);
  if (n <= 256) return launch_pcr<64, 4>(a, b, c, f, x, ldk, k1, k2, col0, ncol, nsweep, mixed, s);
  return tgx_set_error(1, "PCR supports at most 256 rows per column (got %d)", n);
}

extern "C" int tgx_launch_blktri2_f64(const double *a, const double *b, const double *c,
                                      const double *f, double *x, double *gs, int ldk, int nrow,
                                      int col0, int ncol, hipStream_t s) {
  if (ncol <= 0 || nrow <= 0) return 0;
  hipLaunchKernelGGL(tgx::blktri2_kfast_f64, dim3((ncol + 255) / 256), dim3(256), 0, s, a, b, c,
                     f, x, gs, ldk, nrow, col0, ncol);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_launch_blktri3_f64(const double *p, const double *q, const double *r,
                                      const double *f, double *x, double *gs, int ldk, int nrow,
                                      int col0, int ncol, hipStream_t s) {
  if (ncol <= 0 || nrow <= 0) return 0;
  hipLaunchKernelGGL(tgx::blktri3_kfast_f64, dim3((ncol + 255) / 256), dim3(256), 0, s, p, q, r,
                     f, x, gs, ldk, nrow, col0, ncol);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" int tgx_blk_mixed_dev(int nblk, const double *A, const double *B, const double *C,
                                 const double *F, double *X, long pstride, int nk, int nrow,
                                 const unsigned char *valid, long ncol, int nsweep, void *work,
                                 void *stream) {
  if (nblk < 1 || nblk > 3) return tgx_set_error(1, "tgx_blk_mixed_dev: block size %d not in 1..3", nblk);
  if (nrow < 1 || nrow > nk || ncol < 0 || nsweep < 0 || nsweep > 8 || !work || !valid)
    return tgx_set_error(1, "tgx_blk_mixed_dev: bad arguments");
  if (ncol == 0) return 0;
  const size_t nn = (size_t)nblk * nblk, per = (size_t)ncol * nrow;
  float *gam = (float *)work;
  double *ginv = (double *)((char *)work + ((per * nn * sizeof(float) + 255) / 256) * 256);
  double *yv = ginv + per * nn;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((unsigned)((ncol + 255) / 256)), block(256);
  if (nblk == 1)
    hipLaunchKernelGGL(tgx::blk_mixed_kernel<1>, grid, block, 0, s, A, B, C, F, X, pstride, nk, nrow, valid,
                       ncol, nsweep, gam, ginv, yv);
  else if (nblk == 2)
    hipLaunchKernelGGL(tgx::blk_mixed_kernel<2>, grid, block, 0, s, A, B, C, F, X, pstride, nk, nrow, valid,
                       ncol, nsweep, gam, ginv, yv);
  else
    hipLaunchKernelGGL(tgx::blk_mixed_kernel<3>, grid, block, 0, s, A, B, C, F, X, pstride, nk, nrow, valid,
                       ncol, nsweep, gam, ginv, yv);
  TGX_HIP(hipGetLastError());
  return 0;
}

extern "C" unsigned long long tgx_blk_mixed_work_bytes(int nblk, int nrow, long ncol) {
  const size_t nn = (size_t)nblk * nblk, per = (size_t)ncol * nrow;
  return ((per * nn * sizeof(float) + 255) / 256) * 256 + per * nn * sizeof(double) +
         per * 2 * nblk * sizeof(double);
}
