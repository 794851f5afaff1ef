// The global grid's passes around the column kernels: comp's post pass
// after its ring filter (the time smoothing, N2 and the normalisation) on the
// latitudes the filter reaches (the column kernels apply the post passes
// themselves on the others, and the ring filter launch applies those of dt,
// duv, the minor species and O+), the periodic points of the outputs
// (mp_periodic_f3d) and the halos of an input field (mk_polelat.F,
// mp_periodic_f4d).  Level planes as in tgx_step.h; the elementwise passes
// take one thread per (point, level), the longitudes fastest.
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "../../include/tgx_global.h"
#include "tgx_common.h"

namespace tgx {
namespace glb {

constexpr double DTSMOOTH = 0.95, DTSMOOTH_DIV2 = 0.5 * (1. - DTSMOOTH);

struct Dims {
  int nk, nlon, nlat, ld;
  int nfilt;     // latitudes the ring filter reaches next to each pole
  size_t plane;  // ld * (nlat + 4)
};

// Row j of the r-th filtered latitude (r < 2 nfilt): the nfilt next to the
// south pole, then the nfilt next to the north pole.
__device__ __forceinline__ int filtered_row(const Dims &D, int r) {
  return r < D.nfilt ? 2 + r : D.nlat + 2 - 2 * D.nfilt + r;
}

// comp.F:936-1022 on one column per thread: the time smoothing of the
// filtered O2, O, He, N2 above its lowest level under 0.01 extrapolated in
// log(N2 n) (:975-990), the floors and the normalisation.
struct CompIO {
  double *o2, *o1, *he, *o2nm, *o1nm, *henm;
  const double *o2i, *o1i, *hei, *o2nmi, *o1nmi, *henmi, *xnmbar;
  const double *zpmid;
  double dz;
};

__global__ void __launch_bounds__(256) comp_post_kernel(Dims D, CompIO C) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)D.nlon * 2 * D.nfilt) return;
  const int row = (int)(t / D.nlon);
  const int j = filtered_row(D, row), i = 2 + (int)(t - (long)row * D.nlon);
  const size_t c = (size_t)j * D.ld + i;
  const int nk = D.nk;
  const double small = 1.e-9;
  // the lowest level with N2 < 0.01 and the level below it
  int ktop = nk;
  for (int k = 0; k < nk; ++k) {
    const size_t p = (size_t)k * D.plane + c;
    const double n2 = 1. - C.o2[p] - C.o1[p] - C.he[p];
    if (n2 < 0.01) {
      ktop = k;
      break;
    }
  }
  const bool extrap = ktop + 1 < nk && ktop >= 1;
  double n2top = 0., dn2dz = 0., zmtop = 0.;
  if (extrap) {
    const size_t pt = (size_t)ktop * D.plane + c, pb = pt - D.plane;
    zmtop = C.zpmid[ktop];
    n2top = log((1. - C.o2[pt] - C.o1[pt] - C.he[pt]) * C.xnmbar[pt]);
    const double n2top1 = log((1. - C.o2[pb] - C.o1[pb] - C.he[pb]) * C.xnmbar[pb]);
    dn2dz = (n2top - n2top1) / C.dz;
  }
  for (int k = 0; k < nk; ++k) {
    const size_t p = (size_t)k * D.plane + c;
    const double o2 = C.o2[p], o1 = C.o1[p], he = C.he[p];
    const double o2nm = DTSMOOTH_DIV2 * (C.o2nmi[p] + o2) + DTSMOOTH * C.o2i[p];
    const double o1nm = DTSMOOTH_DIV2 * (C.o1nmi[p] + o1) + DTSMOOTH * C.o1i[p];
    const double henm = DTSMOOTH_DIV2 * (C.henmi[p] + he) + DTSMOOTH * C.hei[p];
    double n2 = 1. - o2 - o1 - he;
    if (extrap && k > ktop) n2 = exp(n2top + dn2dz * (C.zpmid[k] - zmtop)) / C.xnmbar[p];
    const double a = o2 > small ? o2 : small, b = o1 > small ? o1 : small;
    const double cc = he > small ? he : small, d = n2 > small ? n2 : small;
    const double norm = a + b + cc + d;
    C.o2[p] = a / norm;
    C.o1[p] = b / norm;
    C.he[p] = cc / norm;
    const double n2nm = 1. - o2nm - o1nm - henm;
    const double an = o2nm > small ? o2nm : small, bn = o1nm > small ? o1nm : small;
    const double cn = henm > small ? henm : small, dn = n2nm > small ? n2nm : small;
    const double normn = an + bn + cn + dn;
    C.o2nm[p] = an / normn;
    C.o1nm[p] = bn / normn;
    C.henm[p] = cn / normn;
  }
}

// mp_periodic_f3d on the interior latitudes of up to eight fields: thread
// per (periodic point, interior latitude, level, field).
struct Periodic {
  double *f[8];
  int n, nk;
};

__global__ void __launch_bounds__(256) periodic_kernel(Dims D, Periodic P) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = 4L * D.nlat * P.nk;
  if (t >= per * P.n) return;
  const int q = (int)(t / per);
  long r = t - (long)q * per;
  const int k = (int)(r / (4L * D.nlat));
  r -= (long)k * 4L * D.nlat;
  const int j = 2 + (int)(r / 4), c = (int)(r & 3);
  const int dst = c == 0 ? 0 : c == 1 ? 1 : c == 2 ? D.nlon + 2 : D.nlon + 3;
  const int src = c == 0 ? D.nlon : c == 1 ? D.nlon + 1 : c == 2 ? 2 : 3;
  double *row = P.f[q] + (size_t)k * D.plane + (size_t)j * D.ld;
  row[dst] = row[src];
}

// The rows across the poles of `dst` from those of `src` (O+ between
// sub-cycles): thread per (column, ghost row, level).
__global__ void __launch_bounds__(256) ghost_rows_kernel(Dims D, const double *__restrict__ src,
                                                         double *__restrict__ dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nlond = D.nlon + 4;
  if (t >= 4L * nlond * D.nk) return;
  const int k = (int)(t / (4L * nlond));
  const long r = t - (long)k * 4L * nlond;
  const int g = (int)(r / nlond), i = (int)(r - (long)g * nlond);
  const int j = g < 2 ? g : D.nlat + g;  // 0, 1, nlat+2, nlat+3
  const size_t p = (size_t)k * D.plane + (size_t)j * D.ld + i;
  dst[p] = src[p];
}

// tgx_glb_halos: periodic points of the interior rows and the rows across
// the poles, thread per (column, row to fill, level); the pole rows read the
// interior columns of interior rows only, so the two kinds of writes are
// independent.
__global__ void __launch_bounds__(256) halos_kernel(Dims D, double *__restrict__ f, double sign) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nlond = D.nlon + 4;
  const long nper = 4L * D.nlat, npole = 4L * nlond;
  if (t >= (nper + npole) * D.nk) return;
  const int k = (int)(t / (nper + npole));
  long r = t - (long)k * (nper + npole);
  double *pl = f + (size_t)k * D.plane;
  if (r < nper) {
    const int j = 2 + (int)(r / 4), c = (int)(r & 3);
    const int dst = c == 0 ? 0 : c == 1 ? 1 : c == 2 ? D.nlon + 2 : D.nlon + 3;
    const int src = c == 0 ? D.nlon : c == 1 ? D.nlon + 1 : c == 2 ? 2 : 3;
    pl[(size_t)j * D.ld + dst] = pl[(size_t)j * D.ld + src];
    return;
  }
  r -= nper;
  const int g = (int)(r / nlond), i = (int)(r - (long)g * nlond);
  // latitude 0 <- 1, -1 <- 2, nlat+1 <- nlat, nlat+2 <- nlat-1 (rows j = lat + 1)
  const int jnew = g == 0 ? 1 : g == 1 ? 0 : g == 2 ? D.nlat + 2 : D.nlat + 3;
  const int jold = g == 0 ? 2 : g == 1 ? 3 : g == 2 ? D.nlat + 1 : D.nlat;
  const int isrc = ((i - 2 + D.nlon / 2) % D.nlon + D.nlon) % D.nlon + 2;
  pl[(size_t)jnew * D.ld + i] = sign * pl[(size_t)jold * D.ld + isrc];
}

inline unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

Dims dims_of(const tgx_ng_grid *g, int nfilt) {
  Dims D;
  D.nfilt = nfilt;
  D.nk = g->nk;
  D.nlon = g->nlon;
  D.nlat = g->nlat;
  D.ld = g->ld;
  D.plane = (size_t)g->ld * (size_t)(g->nlat + 4);
  return D;
}

int periodic(const Dims &D, double *const *f, int n, int nk, hipStream_t s) {
  Periodic P = {};
  for (int q = 0; q < n; ++q) P.f[q] = f[q];
  P.n = n;
  P.nk = nk;
  hipLaunchKernelGGL(periodic_kernel, dim3(nblocks(4L * D.nlat * nk * n)), dim3(256), 0, s, D, P);
  TGX_HIP(hipGetLastError());
  return 0;
}

}  // namespace glb
}  // namespace tgx

using namespace tgx::glb;

// What remains of a global routine after its ring filter launch
// (step_api.cpp), which has applied the post pass of dt, duv, the minor
// species and O+ on the filtered latitudes: which 0 dt, 1 duv, 2 minor
// species N2D and Ar, 3 minor species NO and N4S, 4 O+ sub-step (op_in,
// nm_in: the sub-step's inputs, op_o, nm_o: its outputs, whose rows across
// the poles take those of in->op / in->op_nm), 5 comp (its post pass on the
// nfilt filtered latitudes next to each pole, which needs whole columns).
// Then the periodic points of the outputs on every interior latitude.
extern "C" int tgx_glb_post(int which, int nfilt, const tgx_ng_grid *g, const tgx_ng_fields *in,
                            const tgx_ng_out *out, const double *op_in, const double *nm_in, double *op_o,
                            double *nm_o, hipStream_t s) {
  if (which < 0 || which > 5) return tgx_set_error(TGX_ERR_ARG, "tgx_glb_post: unknown pass %d", which);
  const Dims D = dims_of(g, nfilt);
  if (which == 5 && nfilt > 0) {
    CompIO C;
    C.o2 = out->o2; C.o1 = out->o1; C.he = out->he;
    C.o2nm = out->o2_nm; C.o1nm = out->o1_nm; C.henm = out->he_nm;
    C.o2i = in->o2; C.o1i = in->o1; C.hei = in->he;
    C.o2nmi = in->o2_nm; C.o1nmi = in->o1_nm; C.henmi = in->he_nm;
    C.xnmbar = in->xnmbar;
    C.zpmid = g->zpmid;
    C.dz = g->dz;
    hipLaunchKernelGGL(comp_post_kernel, dim3(nblocks((long)D.nlon * 2 * nfilt)), dim3(256), 0, s, D, C);
  }
  TGX_HIP(hipGetLastError());
  // mp_periodic_f3d of every output the pass wrote
  double *per[8];
  int n = 0;
  switch (which) {
    case 0: per[n++] = out->tn; per[n++] = out->tn_nm; break;
    case 1: per[n++] = out->un; per[n++] = out->vn; per[n++] = out->un_nm; per[n++] = out->vn_nm; break;
    case 2: per[n++] = out->n2d; per[n++] = out->n2d_nm; per[n++] = out->ar; per[n++] = out->ar_nm; break;
    case 3: per[n++] = out->no; per[n++] = out->no_nm; per[n++] = out->n4s; per[n++] = out->n4s_nm; break;
    case 4: per[n++] = op_o; per[n++] = nm_o; break;
    default:
      per[n++] = out->o2; per[n++] = out->o1; per[n++] = out->he;
      per[n++] = out->o2_nm; per[n++] = out->o1_nm; per[n++] = out->he_nm;
  }
  if (int rc = periodic(D, per, n, D.nk, s)) return rc;
  if (which == 1) {  // the lower-boundary winds (duv.F:473-500)
    double *per2[4] = {out->ulbc, out->vlbc, out->ulbc_nm, out->vlbc_nm};
    if (int rc = periodic(D, per2, 4, 1, s)) return rc;
  }
  if (which == 4) {
    hipLaunchKernelGGL(ghost_rows_kernel, dim3(nblocks(4L * (D.nlon + 4) * D.nk)), dim3(256), 0, s, D, in->op, op_o);
    hipLaunchKernelGGL(ghost_rows_kernel, dim3(nblocks(4L * (D.nlon + 4) * D.nk)), dim3(256), 0, s, D, in->op_nm,
                       nm_o);
    TGX_HIP(hipGetLastError());
  }
  return 0;
}

extern "C" int tgx_glb_halos(double *f, int nk, int nlon, int nlat, int ld, double sign, void *stream) {
  if (!f || nk < 1 || nlon < 4 || (nlon & 1) || nlat < 4 || ld < nlon + 4)
    return tgx_set_error(TGX_ERR_ARG, "tgx_glb_halos: bad arguments (nk %d nlon %d (even) nlat %d ld %d)", nk, nlon,
                         nlat, ld);
  Dims D;
  D.nfilt = 0;
  D.nk = nk;
  D.nlon = nlon;
  D.nlat = nlat;
  D.ld = ld;
  D.plane = (size_t)ld * (size_t)(nlat + 4);
  const long n = (4L * nlat + 4L * (nlon + 4)) * nk;
  hipLaunchKernelGGL(halos_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, D, f, sign);
  TGX_HIP(hipGetLastError());
  return 0;
}
