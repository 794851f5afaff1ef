// Host side of the fused column step: workspace layout, launch order and the
// O+ sub-cycle ping-pong (dynamics_ng.F90:99-121).
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "../../include/tgx_step.h"
#include "tgx_common.h"

namespace tgx {
namespace ng {
struct Geo {
  int nk, nlon, nlat, nlond, nlatd, ld;
  size_t P;
};
struct OpWork {
  double *dj, *bdotu, *bvel, *tpop, *diffj, *djbz, *opbm2, *bdopj, *w2;
};
}  // namespace ng
}  // namespace tgx

// launchers defined in step.hip
extern "C" int tgx_ng_launch(int which, const tgx_ng_grid *g, const tgx_ng_fields *in,
                             const tgx_ng_out *out, void *const *bufs, int sub, hipStream_t s);

namespace {
// Workspace planes, each nk * ld * (nlat+4) doubles.
enum {
  W_DT = 0, W_DT_T = 1,                  // dt ratios; 2 prelude planes
  W_MIN_A = 3, W_MIN_B = 4, W_MIN_T = 5, // minor ratios (two species per pass); 8 prelude planes
  W_OP = 13,                             // O+: 9 intermediates, 2 ping-pong, 6 prelude planes
  W_OP_X = W_OP + 9, W_OP_XNM = W_OP + 10, W_OP_T = W_OP + 11,
  W_UV = W_OP_T + 6,                     // wind: ztmp, 4 gamma, 6 prelude planes
  W_CO = W_UV + 11,                      // comp: 9 gamma, 6 prelude planes
  W_NPLANES = W_CO + 15
};

size_t plane_elems(const tgx_ng_grid *g) { return (size_t)g->nk * (size_t)g->ld * (size_t)(g->nlat + 4); }

int check(const tgx_ng_grid *g, const void *work) {
  if (!g) return tgx_set_error(1, "tgx_ng: null grid");
  if (g->nk < 4 || g->nlon < 2 || g->nlat < 2 || g->ld < g->nlon + 4)
    return tgx_set_error(1, "tgx_ng: bad grid nk=%d nlon=%d nlat=%d ld=%d", g->nk, g->nlon, g->nlat, g->ld);
  if (g->nstep_sub < 1) return tgx_set_error(1, "tgx_ng: nstep_sub must be >= 1");
  if ((size_t)g->ld * (size_t)(g->nlat + 4) * (size_t)g->nk > ((size_t)1 << 40))
    return tgx_set_error(1, "tgx_ng: grid too large");
  if (!work) return tgx_set_error(1, "tgx_ng: null workspace");
  return 0;
}

double *plane(void *work, const tgx_ng_grid *g, int idx) {
  return (double *)work + (size_t)idx * plane_elems(g);
}
}  // namespace

extern "C" unsigned long long tgx_ng_workspace_bytes(const tgx_ng_grid *g) {
  if (!g) return 0;
  return (unsigned long long)W_NPLANES * plane_elems(g) * sizeof(double);
}

enum { K_DT = 0, K_MIN_A, K_MIN_B, K_OP1, K_OP2, K_OP3, K_UV1, K_UV2, K_COMP, K_OP_PRE,
       K_DT_PRE, K_MIN_PRE, K_CO_PRE, K_UV_PRE };

extern "C" int tgx_ng_dt(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                         void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  void *b[3] = {plane(work, g, W_DT), plane(work, g, W_DT_T), plane(work, g, W_DT_T + 1)};
  if (int rc = tgx_ng_launch(K_DT_PRE, g, in, out, b, 0, s)) return rc;
  return tgx_ng_launch(K_DT, g, in, out, b, 0, s);
}

extern "C" int tgx_ng_minors(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                             void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  void *b[10];
  b[0] = plane(work, g, W_MIN_A);
  b[1] = plane(work, g, W_MIN_B);
  for (int q = 0; q < 8; ++q) b[2 + q] = plane(work, g, W_MIN_T + q);
  if (int rc = tgx_ng_launch(K_MIN_PRE, g, in, out, b, 0, s)) return rc;
  // N2D and AR first; NO and N4S use this step's N2D (dynamics_ng.F90:147-165)
  if (int rc = tgx_ng_launch(K_MIN_A, g, in, out, b, 0, s)) return rc;
  return tgx_ng_launch(K_MIN_B, g, in, out, b, 0, s);
}

extern "C" int tgx_ng_oplus(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                            void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int nsub = g->nstep_sub;
  double *tmp_op = plane(work, g, W_OP_X), *tmp_nm = plane(work, g, W_OP_XNM);
  const double *op_in = in->op, *nm_in = in->op_nm;
  for (int sub = 0; sub < nsub; ++sub) {
    // the last sub-step writes the outputs; earlier ones alternate so that
    // a sub-step never reads the buffer it writes
    bool to_out = ((nsub - 1 - sub) % 2) == 0;
    double *op_o = to_out ? out->op : tmp_op, *nm_o = to_out ? out->op_nm : tmp_nm;
    void *b[19];
    for (int q = 0; q < 9; ++q) b[q] = plane(work, g, W_OP + q);
    b[9] = (void *)op_in;
    b[10] = (void *)nm_in;
    b[11] = op_o;
    b[12] = nm_o;
    for (int q = 0; q < 6; ++q) b[13 + q] = plane(work, g, W_OP_T + q);
    if (int rc = tgx_ng_launch(K_OP1, g, in, out, b, sub, s)) return rc;
    if (int rc = tgx_ng_launch(K_OP2, g, in, out, b, sub, s)) return rc;
    if (int rc = tgx_ng_launch(K_OP_PRE, g, in, out, b, sub, s)) return rc;
    if (int rc = tgx_ng_launch(K_OP3, g, in, out, b, sub, s)) return rc;
    op_in = op_o;
    nm_in = nm_o;
  }
  return 0;
}

extern "C" int tgx_ng_duv(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                          void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  void *b[8];
  b[0] = plane(work, g, W_UV);
  b[1] = plane(work, g, W_UV + 1);
  for (int q = 0; q < 6; ++q) b[2 + q] = plane(work, g, W_UV + 5 + q);
  if (int rc = tgx_ng_launch(K_UV1, g, in, out, b, 0, s)) return rc;
  if (int rc = tgx_ng_launch(K_UV_PRE, g, in, out, b, 0, s)) return rc;
  return tgx_ng_launch(K_UV2, g, in, out, b, 0, s);
}

extern "C" int tgx_ng_comp(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                           void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  void *b[7];
  b[0] = plane(work, g, W_CO);
  for (int q = 0; q < 6; ++q) b[1 + q] = plane(work, g, W_CO + 9 + q);
  if (int rc = tgx_ng_launch(K_CO_PRE, g, in, out, b, 0, s)) return rc;
  return tgx_ng_launch(K_COMP, g, in, out, b, 0, s);
}

extern "C" int tgx_ng_step(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                           void *work, void *stream) {
  // dynamics_ng.F90: oplus (:99-121), minors (:147-165), dt (:203), duv (:210),
  // comp_o2o + comp (:221-233)
  if (int rc = tgx_ng_oplus(g, in, out, work, stream)) return rc;
  if (int rc = tgx_ng_minors(g, in, out, work, stream)) return rc;
  if (int rc = tgx_ng_dt(g, in, out, work, stream)) return rc;
  if (int rc = tgx_ng_duv(g, in, out, work, stream)) return rc;
  return tgx_ng_comp(g, in, out, work, stream);
}
