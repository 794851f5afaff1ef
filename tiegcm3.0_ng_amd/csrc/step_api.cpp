// Host side of the fused column step: workspace layout, launch order and the
// O+ sub-cycle ping-pong (dynamics_ng.F90:99-121).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/tgx.h"
#include "../../include/tgx_global.h"
#include "../../include/tgx_step.h"
#include "tgx_common.h"

// launchers defined in step.hip and global.hip
extern "C" int tgx_ng_launch(int which, const tgx_ng_grid *g, const tgx_ng_fields *in,
                             const tgx_ng_out *out, void *const *bufs, int sub, hipStream_t s);
extern "C" void tgx_ng_set_global_mode(int on, int nfilt);
extern "C" void tgx_ng_set_wave_columns(int cols);
// filter.hip
extern "C" int tgx_ring_table_dev(double *tab, int nlon, int tord, double dlamda, hipStream_t s);
extern "C" int tgx_ringfilter_post_dev(int nf, double *const *f, double *const *xnm, const double *const *fin,
                                       const double *const *fnmi, int kind, int nlev, int nk, int nlon, int nlat,
                                       int ld, int order, int oplus, double dlamda, int nlat_filter,
                                       const int *chunk, const double *gtab, int tord, hipStream_t s);
extern "C" int tgx_glb_post(int which, int nfilt, const tgx_ng_grid *g, const tgx_ng_fields *in,
                            const tgx_ng_out *out, const double *op_in, const double *nm_in, double *op_o,
                            double *nm_o, hipStream_t s);

namespace {
// The ring filters' cos/sin table of the global routines: harmonics 1..3
// (the largest order they filter with), at the end of the workspace.
constexpr int kGlbTableOrder = 3;
// The global grid of the tgx_glb_* call running on this host thread, or null
// for a nest domain.  The routines below then run the Thomas scheme with the
// global grid's column kernels and, after each column kernel, the ring filter
// and the global post pass of that routine.
thread_local const tgx_glb_grid *t_glb = nullptr;
// Its ring filters' cos/sin table in the workspace (filled by glb_call).
thread_local const double *t_glb_tab = nullptr;

// After a global routine's column kernel, its ring filters and post pass
// (which: as tgx_glb_post).  dt.F:389-424 (filter_tn below the top level),
// duv.F:490-514 (filter_uv of UN and VN below the top level),
// minor.F:546-597 (filter_minor, all levels; a pair of species per pass),
// oplus.F:896-941 (filter_op below the top level, one sub-step), comp.F:925-1022
// (filter_o2o of O2, O, He, all levels, order 1).  All the routine's fields
// are filtered in one launch, with the post pass applied to each filtered
// value (comp's post pass, which needs the whole column, stays its own
// kernel); tgx_glb_post then adds the periodic points.
int glb_finish(int which, const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
               const double *op_in, const double *nm_in, double *op_o, double *nm_o, hipStream_t s) {
  enum { NONE = 0, DT, DUV, MINOR, OPLUS };  // the post kinds of tgx_ringfilter_post_dev
  double *f[4] = {}, *xnm[4] = {};
  const double *fi[4] = {}, *fm[4] = {};
  int nf = 0, kind = NONE, nlev = g->nk - 1, order = 3, oplus = 0;
  auto add = [&](double *x, double *xn, const double *a, const double *b) {
    f[nf] = x;
    xnm[nf] = xn;
    fi[nf] = a;
    fm[nf] = b;
    ++nf;
  };
  switch (which) {
    case 0:
      kind = DT;
      add(out->tn, out->tn_nm, in->tn, in->tn_nm);
      break;
    case 1:
      kind = DUV;
      add(out->un, out->un_nm, in->un, in->un_nm);
      add(out->vn, out->vn_nm, in->vn, in->vn_nm);
      break;
    case 2:
      kind = MINOR;
      nlev = g->nk;
      add(out->n2d, out->n2d_nm, in->n2d, in->n2d_nm);
      add(out->ar, out->ar_nm, in->ar, in->ar_nm);
      break;
    case 3:
      kind = MINOR;
      nlev = g->nk;
      add(out->no, out->no_nm, in->no, in->no_nm);
      add(out->n4s, out->n4s_nm, in->n4s, in->n4s_nm);
      break;
    case 4:
      kind = OPLUS;
      oplus = 1;
      add(op_o, nm_o, op_in, nm_in);
      break;
    case 5:
      nlev = g->nk;
      order = 1;
      add(out->o2, nullptr, nullptr, nullptr);
      add(out->o1, nullptr, nullptr, nullptr);
      add(out->he, nullptr, nullptr, nullptr);
      break;
    default:
      return tgx_set_error(TGX_ERR_ARG, "glb_finish: unknown routine %d", which);
  }
  const int nfilt = t_glb->nlat_filter;
  if (int rc = tgx_ringfilter_post_dev(nf, f, xnm, fi, fm, kind, nlev, g->nk, g->nlon, g->nlat, g->ld, order, oplus,
                                       g->dlamda, nfilt, t_glb->chunk, t_glb_tab, kGlbTableOrder, s))
    return rc;
  return tgx_glb_post(which, nfilt, g, in, out, op_in, nm_in, op_o, nm_o, s);
}
}  // namespace

namespace {
// Workspace planes, each nk * ld * (nlat+4) doubles.
enum {
  W_DT_T = 0,                            // dt: 2 prelude planes (PCR)
  W_OP = 2,                              // O+: 8 intermediates, 2 ping-pong, 6 prelude planes
  W_OP_X = W_OP + 8, W_OP_XNM = W_OP + 9, W_OP_T = W_OP + 10,
  W_OP_C = W_OP_T + 6,                   // O+ loss and production (once per step)
  W_UV = W_OP_C + 2,                     // wind: ztmp, 6 prelude planes
  W_CO = W_UV + 7,                       // comp: fs(:,1,0), fs(:,2,0)
  W_NPLANES = W_CO + 2
};
// Tiled regions after the planes, in units of tile_elems() doubles per value
// (the kernels' Geo::tile and Geo::wtile give the element order).
enum {
  T_CI = 0,        // comp inputs: 3 right-hand-side parts, 4 fs values
  T_CE = T_CI + 7, // comp elimination: 9 gamma, 3 forward solution
  T_MR = T_CE + 12, // minor species: right-hand-side parts of NO, N4S, N2D, AR
  T_MW = T_MR + 4,  // minor species: Thomas ratios and forward solutions of a pass's two species
  T_UW = T_MW + 4,  // winds: 2x2 block gamma and forward solution
  T_DW = T_UW + 6,  // TN: Thomas ratio and forward solution
  T_OW = T_DW + 2,  // O+: Thomas ratio and forward solution
  T_MS = T_OW + 2,  // minor species: loss and scaled production of NO or N2D, and of N4S
  T_ML = T_MS + 4,  // minor species: NO / N4S source coefficients linear in N2D
  T_OF = T_ML + 6,  // O+: factorisation kept across the sub-cycles of a step
  T_CU = T_OF + 3,  // comp precision form: u (O2, O, He) between refinement sweeps
  T_N = T_CU + 3
};

size_t plane_elems(const tgx_ng_grid *g) { return (size_t)g->nk * (size_t)g->ld * (size_t)(g->nlat + 4); }

int check(const tgx_ng_grid *g, const void *work) {
  if (!g) return tgx_set_error(1, "tgx_ng: null grid");
  if (g->nk < 4 || g->nlon < 2 || g->nlat < 2 || g->ld < g->nlon + 4)
    return tgx_set_error(1, "tgx_ng: bad grid nk=%d nlon=%d nlat=%d ld=%d", g->nk, g->nlon, g->nlat, g->ld);
  if (g->nstep_sub < 1) return tgx_set_error(1, "tgx_ng: nstep_sub must be >= 1");
  if (g->scheme < TGX_SCHEME_AUTO || g->scheme > TGX_SCHEME_PCR)
    return tgx_set_error(1, "tgx_ng: unknown scheme %d", g->scheme);
  if (g->scheme == TGX_SCHEME_PCR && g->nk > 128)
    return tgx_set_error(1, "tgx_ng: the PCR scheme holds at most 128 levels (nk=%d)", g->nk);
  if ((size_t)g->ld * (size_t)(g->nlat + 4) * (size_t)g->nk > ((size_t)1 << 40))
    return tgx_set_error(1, "tgx_ng: grid too large");
  if (!work) return tgx_set_error(1, "tgx_ng: null workspace");
  return 0;
}

double *plane(void *work, const tgx_ng_grid *g, int idx) {
  return (double *)work + (size_t)idx * plane_elems(g);
}

// One value of a tiled region: nk levels of either 64-column groups (column
// kernels) or the wavefronts of a tile-kernel launch, whichever is larger.
size_t tile_elems(const tgx_ng_grid *g) {
  const size_t nlond = (size_t)(g->nlon + 4), nlatd = (size_t)(g->nlat + 4);
  const size_t ncol = (nlond * nlatd + 63) / 64 * 64;
  const size_t nwt = 256 * ((nlond + TGX_TILE_X - 1) / TGX_TILE_X) * ((nlatd + TGX_TILE_Y - 1) / TGX_TILE_Y);
  return (size_t)g->nk * (ncol > nwt ? ncol : nwt);
}

double *tiles(void *work, const tgx_ng_grid *g, int idx) {
  return (double *)work + (size_t)W_NPLANES * plane_elems(g) + (size_t)idx * tile_elems(g);
}
}  // namespace

extern "C" unsigned long long tgx_ng_workspace_bytes(const tgx_ng_grid *g) {
  if (!g) return 0;
  return ((unsigned long long)W_NPLANES * plane_elems(g) + (unsigned long long)T_N * tile_elems(g)) *
         sizeof(double);
}

// Optional per-routine timing: hipEvent pairs recorded on the stream each
// routine runs on, accumulated until tgx_ng_profile_end.  Off by default.
namespace {
enum { R_OPLUS = 0, R_MINORS, R_DT, R_DUV, R_COMP, R_N, P_N = 2 * R_N };  // routines, then their column kernels
struct Prof {
  bool on = false;
  int cap = 0;
  int used[P_N] = {};
  std::vector<hipEvent_t> ev;  // [P_N][cap][2]
};
std::mutex g_prof_mu;
Prof g_prof;

struct ProfScope {
  hipEvent_t e1 = nullptr;
  hipStream_t s;
  ProfScope(int r, hipStream_t st) : s(st) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof.on || g_prof.used[r] >= g_prof.cap) return;
    size_t at = ((size_t)r * g_prof.cap + g_prof.used[r]++) * 2;
    if (hipEventRecord(g_prof.ev[at], s) == hipSuccess) e1 = g_prof.ev[at + 1];
  }
  ~ProfScope() {
    if (e1) (void)hipEventRecord(e1, s);
  }
};
}  // namespace

extern "C" int tgx_ng_profile_begin(int max_calls) {
  if (max_calls < 1 || max_calls > 100000) return tgx_set_error(1, "tgx_ng_profile_begin: bad max_calls %d", max_calls);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  size_t need = (size_t)P_N * max_calls * 2;
  while (g_prof.ev.size() < need) {
    hipEvent_t e;
    TGX_HIP(hipEventCreate(&e));
    g_prof.ev.push_back(e);
  }
  g_prof.cap = max_calls;
  for (int r = 0; r < P_N; ++r) g_prof.used[r] = 0;
  g_prof.on = true;
  return 0;
}

extern "C" int tgx_ng_profile_end(double *ms_total, int *calls) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof.on) return tgx_set_error(1, "tgx_ng_profile_end: profiling not started");
  g_prof.on = false;
  for (int r = 0; r < P_N; ++r) {
    double tot = 0.;
    for (int c = 0; c < g_prof.used[r]; ++c) {
      size_t at = ((size_t)r * g_prof.cap + c) * 2;
      TGX_HIP(hipEventSynchronize(g_prof.ev[at + 1]));
      float ms = 0.f;
      TGX_HIP(hipEventElapsedTime(&ms, g_prof.ev[at], g_prof.ev[at + 1]));
      tot += ms;
    }
    if (ms_total) ms_total[r] = tot;
    if (calls) calls[r] = g_prof.used[r];
  }
  return 0;
}

// Kernel ids of tgx_ng_launch (0, 5 and 7 are retired: the Thomas dt, O+ and
// duv column kernels are the tile kernels K_DT_TILE, K_OP_TILE, K_UV_TILE; 26
// too: the NO / N4S column kernels form their sources, K_MIN_B_LIN and
// K_MIN_WB_LIN).
enum { K_DT = 0, K_MIN_A, K_MIN_B, K_OP1, K_OP2, K_OP3, K_UV1, K_UV2, K_COMP, K_OP_PRE,
       K_DT_PRE, K_MIN_PRE, K_CO_PRE, K_UV_PRE, K_O2O, K_MIN_SRC_A, K_MIN_SRC_B, K_OP_CHEM,
       K_MIN_WA, K_MIN_WB, K_DT_W, K_OP3_W, K_UV2_W, K_COMP_W, K_DT_TILE, K_CHEM, K_MIN_LIN,
       K_MIN_B_LIN, K_UV_TILE, K_OP_TILE, K_MIN_WB_LIN, K_OP1_LITE, K_OP_TILE_KEEP, K_OP_TILE_REUSE,
       K_COMP_M0, K_COMP_M_REFINE, K_COMP_M_LAST, K_COMP_M_POST };

namespace {
// Solve scheme of each routine.  The wavefront-per-column cyclic reduction
// holds at most 128 levels.  TGX_SCHEME_AUTO uses the choice timed by
// tgx_ng_autotune for this device and shape (PCR until then).
typedef std::tuple<int, int, int, int, int, int, int> TuneKey;  // dev nk nlon nlat ld nsub routine
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
std::map<TuneKey, int> g_tune_cols;  // columns per block of the wavefront-per-column kernels
thread_local int t_forced = -1;       // scheme forced while the autotuner times a routine
thread_local int t_forced_cols = 0;   // and its columns per block (PCR)

TuneKey tune_key(const tgx_ng_grid *g, int routine) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return TuneKey(dev, g->nk, g->nlon, g->nlat, g->ld, g->nstep_sub, routine);
}
int tuned_choice(const tgx_ng_grid *g, int routine) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(g, routine));
  return it == g_tune.end() ? TGX_SCHEME_AUTO : it->second;
}
// Whether the routine runs the wavefront-per-column scheme; if so, its
// columns per block (the autotuner's choice, else 4) are set for the launches
// that follow on this host thread.
bool use_pcr(const tgx_ng_grid *g, int routine) {
  bool pcr;
  if (g->nk > 128 || t_glb) pcr = false;
  else if (t_forced >= 0) pcr = t_forced == TGX_SCHEME_PCR;
  else if (g->scheme != TGX_SCHEME_AUTO) pcr = g->scheme == TGX_SCHEME_PCR;
  else pcr = tuned_choice(g, routine) != TGX_SCHEME_THOMAS;
  if (pcr) {
    int cols = 4;
    if (t_forced_cols > 0) {
      cols = t_forced_cols;
    } else {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cols.find(tune_key(g, routine));
      if (it != g_tune_cols.end()) cols = it->second;
    }
    tgx_ng_set_wave_columns(cols);
  }
  return pcr;
}
}  // namespace

extern "C" int tgx_ng_dt(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                         void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof(R_DT, s);
  void *b[3] = {tiles(work, g, T_DW), plane(work, g, W_DT_T), plane(work, g, W_DT_T + 1)};
  if (!use_pcr(g, R_DT)) {  // Thomas: the stencils are fused into the column kernel
    {
      ProfScope kern(R_N + R_DT, s);
      if (int rc = tgx_ng_launch(K_DT_TILE, g, in, out, b, 0, s)) return rc;
    }
    if (!t_glb) return 0;
    return glb_finish(0, g, in, out, nullptr, nullptr, nullptr, nullptr, s);
  }
  if (int rc = tgx_ng_launch(K_DT_PRE, g, in, out, b, 0, s)) return rc;
  ProfScope kern(R_N + R_DT, s);
  return tgx_ng_launch(K_DT_W, g, in, out, b, 0, s);
}

namespace {
// Buffers of chem_kernel (the step-wide chemistry pass).
void chem_bufs(void *work, const tgx_ng_grid *g, void **b) {
  b[0] = tiles(work, g, T_CI);                                       // comp fs(:,1:2,1:2)
  b[1] = plane(work, g, W_CO + 0);                                   // comp fs(:,1,0)
  b[2] = plane(work, g, W_CO + 1);                                   // comp fs(:,2,0)
  b[3] = b[4] = b[5] = nullptr;
  b[6] = tiles(work, g, T_MS);  // N2D loss, s0prod (then NO's), N4S loss, s0prod
  b[7] = b[8] = nullptr;
  b[9] = tiles(work, g, T_ML);  // NO / N4S source coefficients
  for (int q = 10; q < 15; ++q) b[q] = nullptr;
  b[15] = plane(work, g, W_OP_C);
  b[16] = plane(work, g, W_OP_C + 1);
  b[17] = b[18] = b[19] = nullptr;
}
int run_chem(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
             hipStream_t s) {
  void *b[20];
  chem_bufs(work, g, b);
  return tgx_ng_launch(K_CHEM, g, in, out, b, 0, s);
}
int run_minors(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
               void *stream, bool chem_done, hipEvent_t chem_ready = nullptr);
int run_oplus(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
              void *stream, bool chem_done, int nrun, hipEvent_t fe_ready = nullptr);
int run_comp(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
             void *stream, bool chem_done);
}  // namespace

extern "C" int tgx_ng_minors(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                             void *work, void *stream) {
  return run_minors(g, in, out, work, stream, false);
}
extern "C" int tgx_ng_oplus(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                            void *work, void *stream) {
  return run_oplus(g, in, out, work, stream, false, g ? g->nstep_sub : 0);
}
// nrun (1..nstep_sub) O+ sub-steps of a step whose matrix carries the full
// nstep_sub time factor, with the lateral data blocks 0..nrun-1.  The host
// context runs the reference's one-sub-step oplus_ng calls with nrun = 1.
int tgx_ng_oplus_substeps(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                          void *stream, int nrun) {
  if (g && (nrun < 1 || nrun > g->nstep_sub))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_oplus_substeps: nrun %d outside 1..nstep_sub (%d)", nrun, g->nstep_sub);
  return run_oplus(g, in, out, work, stream, false, nrun);
}
extern "C" int tgx_ng_comp(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                           void *work, void *stream) {
  return run_comp(g, in, out, work, stream, false);
}

namespace {
// chem_ready, if given, is waited for after the stencil prelude (which reads
// no output of the chemistry pass) and before the column kernels.
int run_minors(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
               void *stream, bool chem_done, hipEvent_t chem_ready) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof(R_MINORS, s);
  void *b[14];
  b[0] = tiles(work, g, T_MW);
  b[1] = tiles(work, g, T_MR);
  for (int q = 2; q < 10; ++q) b[q] = nullptr;
  b[10] = tiles(work, g, T_MS);
  for (int q = 11; q < 14; ++q) b[q] = nullptr;
  if (int rc = tgx_ng_launch(K_MIN_PRE, g, in, out, b, 0, s)) return rc;
  if (chem_ready) TGX_HIP(hipStreamWaitEvent(s, chem_ready, 0));
  // N2D and AR first; NO and N4S use this step's N2D (dynamics_ng.F90:147-165)
  const bool pcr = use_pcr(g, R_MINORS);
  if (!chem_done)
    if (int rc = tgx_ng_launch(K_MIN_SRC_A, g, in, out, b, 0, s)) return rc;
  {
    ProfScope kern(R_N + R_MINORS, s);  // the N2D + Ar column kernel
    if (int rc = tgx_ng_launch(pcr ? K_MIN_WA : K_MIN_A, g, in, out, b, 0, s)) return rc;
  }
  // minor.F:546-597 for N2D and Ar before NO and N4S read this step's N2D
  // (dynamics.F:485-600): filter_minor, then the post pass
  if (t_glb)
    if (int rc = glb_finish(2, g, in, out, nullptr, nullptr, nullptr, nullptr, s)) return rc;
  if (chem_done) {  // NO and N4S sources formed in the column kernel from the
                    // step's chemistry pass and this N2D
    void *b2[20];
    for (int q = 0; q < 14; ++q) b2[q] = b[q];
    b2[14] = tiles(work, g, T_ML);
    for (int q = 15; q < 20; ++q) b2[q] = nullptr;
    if (int rc = tgx_ng_launch(pcr ? K_MIN_WB_LIN : K_MIN_B_LIN, g, in, out, b2, 0, s)) return rc;
  } else {
    if (int rc = tgx_ng_launch(K_MIN_SRC_B, g, in, out, b, 0, s)) return rc;
    if (int rc = tgx_ng_launch(pcr ? K_MIN_WB : K_MIN_B, g, in, out, b, 0, s)) return rc;
  }
  if (!t_glb) return 0;
  return glb_finish(3, g, in, out, nullptr, nullptr, nullptr, nullptr, s);
}

// fe_ready, if given, is recorded on the stream once the last sub-step's
// stage 1 has written Fe and Fn (all that duv reads of this routine).
int run_oplus(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
              void *stream, bool chem_done, int nrun, hipEvent_t fe_ready) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof(R_OPLUS, s);
  const int nsub = nrun;  // sub-steps run here (the time factor is g->nstep_sub)
  double *tmp_op = plane(work, g, W_OP_X), *tmp_nm = plane(work, g, W_OP_XNM);
  const double *op_in = in->op, *nm_in = in->op_nm;
  double *oploss = plane(work, g, W_OP_C), *opprod = plane(work, g, W_OP_C + 1);
  // The O+ matrix does not depend on O+ (oplus_ng.F90:252-303, inputs fixed
  // at itp, dynamics_ng.F90:103-119): with the Thomas scheme and several
  // sub-cycles, the first sub-step keeps its factorisation and the later
  // ones rebuild only the right-hand side (bitwise the same values).  The
  // middle sub-steps also skip the O+-independent part of stage 1 and the
  // Fe/Fn outputs, which the last sub-step writes.  Capture mode (analysis)
  // assembles every sub-step in full.
  const bool pcr = use_pcr(g, R_OPLUS);
  const bool reuse = !pcr && nsub > 1 && !g->capture;
  for (int sub = 0; sub < nsub; ++sub) {
    // the last sub-step writes the outputs; earlier ones alternate so that
    // a sub-step never reads the buffer it writes
    bool to_out = ((nsub - 1 - sub) % 2) == 0;
    double *op_o = to_out ? out->op : tmp_op, *nm_o = to_out ? out->op_nm : tmp_nm;
    void *b[22];
    for (int q = 0; q < 8; ++q) b[q] = plane(work, g, W_OP + q);
    b[8] = tiles(work, g, T_OW);  // Thomas ratio and forward solution (K_OP_TILE)
    b[9] = (void *)op_in;
    b[10] = (void *)nm_in;
    b[11] = op_o;
    b[12] = nm_o;
    for (int q = 0; q < 6; ++q) b[13 + q] = plane(work, g, W_OP_T + q);
    b[19] = oploss;
    b[20] = opprod;
    b[21] = tiles(work, g, T_OF);
    // the O+ chemistry does not depend on O+: once for all sub-cycles
    if (sub == 0 && !chem_done)
      if (int rc = tgx_ng_launch(K_OP_CHEM, g, in, out, b, sub, s)) return rc;
    const bool lite = reuse && sub > 0 && sub < nsub - 1;
    if (int rc = tgx_ng_launch(lite ? K_OP1_LITE : K_OP1, g, in, out, b, sub, s)) return rc;
    if (fe_ready && sub == nsub - 1) TGX_HIP(hipEventRecord(fe_ready, s));
    if (int rc = tgx_ng_launch(K_OP2, g, in, out, b, sub, s)) return rc;
    if (pcr)  // Thomas: the stencils are fused into the column kernel
      if (int rc = tgx_ng_launch(K_OP_PRE, g, in, out, b, sub, s)) return rc;
    {
      ProfScope kern(R_N + R_OPLUS, s);
      const int col = pcr ? K_OP3_W : !reuse ? K_OP_TILE : sub == 0 ? K_OP_TILE_KEEP : K_OP_TILE_REUSE;
      if (int rc = tgx_ng_launch(col, g, in, out, b, sub, s)) return rc;
    }
    if (t_glb)  // oplus.F:896-941: filter_op below the top level, then the post pass
      if (int rc = glb_finish(4, g, in, out, op_in, nm_in, op_o, nm_o, s)) return rc;
    op_in = op_o;
    nm_in = nm_o;
  }
  return 0;
}

// The precision form of comp: FP32 assembly and elimination with FP64 block
// pivots, then nsweep (0..2) refinement sweeps with FP64 residuals; the last
// kernel applies the post pass.  Same inputs, outputs and workspace as
// tgx_ng_comp.
namespace {
// Refinement sweeps of comp's precision form inside tgx_ng_step_mixed on
// this host thread (-1: the FP64 routine).
thread_local int t_comp_nsweep = -1;

int check_mixed(const tgx_ng_grid *g, int nsweep, const char *fn) {
  if (nsweep < 0 || nsweep > 2) return tgx_set_error(TGX_ERR_ARG, "%s: nsweep %d outside 0..2", fn, nsweep);
  if (g->capture) return tgx_set_error(TGX_ERR_ARG, "%s: no capture mode", fn);
  if (t_glb) return tgx_set_error(TGX_ERR_ARG, "%s: nest domains only", fn);
  return 0;
}

int run_comp_mixed(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                   hipStream_t s, bool chem_done, int nsweep) {
  ProfScope prof(R_COMP, s);
  void *b[5];
  b[0] = tiles(work, g, T_CE);  // FP32 gamma and forward solution (12 floats per level)
  b[1] = tiles(work, g, T_CI);
  b[2] = plane(work, g, W_CO + 0);
  b[3] = plane(work, g, W_CO + 1);
  b[4] = tiles(work, g, T_CU);
  if (!chem_done)
    if (int rc = tgx_ng_launch(K_O2O, g, in, out, b, 0, s)) return rc;
  if (int rc = tgx_ng_launch(K_CO_PRE, g, in, out, b, 0, s)) return rc;
  ProfScope kern(R_N + R_COMP, s);
  if (int rc = tgx_ng_launch(K_COMP_M0, g, in, out, b, 0, s)) return rc;
  for (int sw = 0; sw + 1 < nsweep; ++sw)
    if (int rc = tgx_ng_launch(K_COMP_M_REFINE, g, in, out, b, 0, s)) return rc;
  return tgx_ng_launch(nsweep > 0 ? K_COMP_M_LAST : K_COMP_M_POST, g, in, out, b, 0, s);
}
}  // namespace

extern "C" int tgx_ng_comp_mixed(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                                 void *work, void *stream, int nsweep) {
  if (int rc = check(g, work)) return rc;
  if (int rc = check_mixed(g, nsweep, "tgx_ng_comp_mixed")) return rc;
  return run_comp_mixed(g, in, out, work, (hipStream_t)stream, false, nsweep);
}

extern "C" int tgx_ng_duv(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                          void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof(R_DUV, s);
  void *b[8];
  b[0] = plane(work, g, W_UV);
  b[1] = tiles(work, g, T_UW);  // Thomas gamma and forward solution (K_UV_TILE)
  for (int q = 0; q < 6; ++q) b[2 + q] = plane(work, g, W_UV + 1 + q);
  if (!use_pcr(g, R_DUV)) {  // Thomas: stage 1 and the stencils are fused into the column kernel
    {
      ProfScope kern(R_N + R_DUV, s);
      if (int rc = tgx_ng_launch(K_UV_TILE, g, in, out, b, 0, s)) return rc;
    }
    if (!t_glb) return 0;
    return glb_finish(1, g, in, out, nullptr, nullptr, nullptr, nullptr, s);
  }
  if (int rc = tgx_ng_launch(K_UV1, g, in, out, b, 0, s)) return rc;
  if (int rc = tgx_ng_launch(K_UV_PRE, g, in, out, b, 0, s)) return rc;
  ProfScope kern(R_N + R_DUV, s);
  return tgx_ng_launch(K_UV2_W, g, in, out, b, 0, s);
}

int run_comp(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
             void *stream, bool chem_done) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (t_comp_nsweep >= 0) return run_comp_mixed(g, in, out, work, s, chem_done, t_comp_nsweep);
  ProfScope prof(R_COMP, s);
  void *b[4];
  b[0] = tiles(work, g, T_CE);  // Thomas gamma and forward solution
  b[1] = tiles(work, g, T_CI);  // right-hand-side parts and fs(:,1:2,1:2)
  b[2] = plane(work, g, W_CO + 0);
  b[3] = plane(work, g, W_CO + 1);
  // comp_o2o_ng's fs planes first: the prelude folds fs(:,:,0) into the
  // right-hand sides
  if (!chem_done)
    if (int rc = tgx_ng_launch(K_O2O, g, in, out, b, 0, s)) return rc;
  if (int rc = tgx_ng_launch(K_CO_PRE, g, in, out, b, 0, s)) return rc;
  {
    ProfScope kern(R_N + R_COMP, s);
    if (int rc = tgx_ng_launch(use_pcr(g, R_COMP) ? K_COMP_W : K_COMP, g, in, out, b, 0, s)) return rc;
  }
  if (!t_glb) return 0;
  return glb_finish(5, g, in, out, nullptr, nullptr, nullptr, nullptr, s);
}
}  // namespace

// Concurrency of the step: oplus, the minor species, dt and comp read only
// the step's inputs and write disjoint outputs and workspace planes, so they
// run on four streams forked from the caller's stream; duv needs dt's TN and
// oplus' Fe/Fn and joins both.  All streams have the default priority.
namespace {
// The fork streams and events belong to the calling host thread (one set per
// thread and device): calls from several threads never share an event, and
// calls from one thread are ordered through its streams.  A worker thread's
// set is released when the thread exits, once its streams have drained; the
// main thread's lives until the process ends.
struct Fork {
  hipStream_t s[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev[8] = {};
};
struct ForkSet {
  std::map<int, Fork> m;
  ~ForkSet() {
    if (tgx_on_main_thread()) return;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    for (auto &kv : m) {
      if (hipSetDevice(kv.first) != hipSuccess) continue;
      for (hipStream_t st : kv.second.s)
        if (st) (void)hipStreamSynchronize(st);
      for (hipEvent_t e : kv.second.ev)
        if (e) (void)hipEventDestroy(e);
      for (hipStream_t st : kv.second.s)
        if (st) (void)hipStreamDestroy(st);
    }
    (void)hipSetDevice(cur);
  }
};
thread_local ForkSet t_fork;

int get_fork(Fork **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tgx_set_error(3, "no HIP device: %s", hipGetErrorString(e));
  Fork &f = t_fork.m[dev];
  if (!f.s[0]) {
    for (int i = 0; i < 3; ++i) TGX_HIP(hipStreamCreateWithFlags(&f.s[i], hipStreamNonBlocking));
    for (int i = 0; i < 8; ++i) TGX_HIP(hipEventCreateWithFlags(&f.ev[i], hipEventDisableTiming));
  }
  *out = &f;
  return 0;
}
}  // namespace

typedef int (*RoutineFn)(const tgx_ng_grid *, const tgx_ng_fields *, const tgx_ng_out *, void *, void *);

// Times every routine under both schemes on `stream` (reference order, so
// duv sees this step's dt and oplus outputs) and keeps the faster scheme per
// routine for this device and shape.  The outputs are left as a full step
// computed with the last scheme timed; the caller runs its step afterwards.
extern "C" int tgx_ng_autotune(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                               void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const RoutineFn fns[R_N] = {tgx_ng_oplus, tgx_ng_minors, tgx_ng_dt, tgx_ng_duv, tgx_ng_comp};
  const int user_cols = t_forced_cols;  // tgx_ng_force_wave_columns, restored below
  hipEvent_t e0, e1;
  TGX_HIP(hipEventCreate(&e0));
  TGX_HIP(hipEventCreate(&e1));
  int rc = 0;
  for (int r = 0; r < R_N && rc == 0; ++r) {
    // the candidates: Thomas, and the wavefront-per-column scheme with 1, 2
    // or 4 columns (wavefronts) per block
    const int schemes[4] = {TGX_SCHEME_THOMAS, TGX_SCHEME_PCR, TGX_SCHEME_PCR, TGX_SCHEME_PCR};
    const int cols[4] = {0, 1, 2, 4};
    float best[4] = {0.f, 0.f, 0.f, 0.f};
    const int nsch = g->nk > 128 ? 1 : 4;
    for (int q = 0; q < nsch && rc == 0; ++q) {
      t_forced = schemes[q];
      t_forced_cols = cols[q];
      float tmin = 1e30f;
      for (int rep = 0; rep < 3 && rc == 0; ++rep) {  // the first run is a warm-up
        if (hipEventRecord(e0, s) != hipSuccess) { rc = tgx_set_error(2, "tgx_ng_autotune: event record"); break; }
        rc = fns[r](g, in, out, work, stream);
        if (rc) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = tgx_set_error(2, "tgx_ng_autotune: event wait");
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < tmin) tmin = ms;
      }
      best[q] = tmin;
    }
    t_forced = -1;
    t_forced_cols = user_cols;
    if (rc == 0) {
      int q = 0;
      for (int c = 1; c < nsch; ++c)
        if (best[c] < best[q]) q = c;
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[tune_key(g, r)] = schemes[q];
      if (q > 0) g_tune_cols[tune_key(g, r)] = cols[q];
    }
  }
  t_forced = -1;
  t_forced_cols = user_cols;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" void tgx_ng_force_wave_columns(int cols) { t_forced_cols = cols == 1 || cols == 2 || cols == 4 ? cols : 0; }

extern "C" int tgx_ng_wave_columns_choice(const tgx_ng_grid *g, int routine) {
  if (!g || routine < 0 || routine >= R_N) return tgx_set_error(1, "tgx_ng_wave_columns_choice: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_cols.find(tune_key(g, routine));
  return it == g_tune_cols.end() ? 4 : it->second;
}

extern "C" int tgx_ng_scheme_choice(const tgx_ng_grid *g, int routine) {
  if (!g || routine < 0 || routine >= R_N) return tgx_set_error(1, "tgx_ng_scheme_choice: bad arguments");
  if (g->scheme != TGX_SCHEME_AUTO) return g->scheme;
  if (g->nk > 128) return TGX_SCHEME_THOMAS;
  int c = tuned_choice(g, routine);
  return c == TGX_SCHEME_AUTO ? TGX_SCHEME_PCR : c;
}

extern "C" int tgx_ng_step(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                           void *work, void *stream) {
  // dynamics_ng.F90 order: oplus (:99-121), minors (:147-165), dt (:203),
  // duv (:210), comp_o2o + comp (:221-233)
  if (int rc = check(g, work)) return rc;
  if (!t_glb && g->scheme == TGX_SCHEME_AUTO && g->nk <= 128 && tuned_choice(g, R_OPLUS) == TGX_SCHEME_AUTO)
    if (int rc = tgx_ng_autotune(g, in, out, work, stream)) return rc;
  Fork *F = nullptr;
  if (int rc = get_fork(&F)) return rc;
  hipStream_t s0 = (hipStream_t)stream, sa = F->s[0], sb = F->s[1], sc = F->s[2];
  // Fork: every stream starts after what the caller queued before the step.
  // From here on every return joins the three fork streams back into s0, so
  // that a failed call still leaves a graph capture joined and no forked
  // kernel running past the caller's stream.
  TGX_HIP(hipEventRecord(F->ev[5], s0));
  for (hipStream_t st : {sa, sb, sc}) TGX_HIP(hipStreamWaitEvent(st, F->ev[5], 0));
  struct Join {
    Fork *F;
    hipStream_t s0;
    ~Join() {
      (void)join(0, 1);
      (void)join(1, 2);
      (void)join(2, 3);
    }
    hipError_t join(int stream, int ev) {
      hipError_t e = hipEventRecord(F->ev[ev], F->s[stream]);
      return e == hipSuccess ? hipStreamWaitEvent(s0, F->ev[ev], 0) : e;
    }
  } join{F, s0};
  // The pointwise chemistry of all routines in one pass on s0 (ev[0]: done),
  // beside the minor species' stencil prelude on stream b, which does not
  // read it; the rest of every routine starts after it.
  if (int rc = run_chem(g, in, out, work, s0)) return rc;
  TGX_HIP(hipEventRecord(F->ev[0], s0));
  TGX_HIP(hipStreamWaitEvent(sa, F->ev[0], 0));
  TGX_HIP(hipStreamWaitEvent(sc, F->ev[0], 0));
  // stream a; ev[4]: its Fe/Fn are written (the last sub-step's stage 1)
  if (int rc = run_oplus(g, in, out, work, sa, true, g->nstep_sub, F->ev[4])) return rc;
  if (int rc = run_minors(g, in, out, work, sb, true, F->ev[0])) return rc;   // stream b
  if (int rc = run_comp(g, in, out, work, sc, true)) return rc;     // stream c
  if (int rc = tgx_ng_dt(g, in, out, work, s0)) return rc;       // caller's stream
  // duv reads Fe/Fn of oplus and no other output or workspace region of it,
  // so it starts while the rest of the O+ sub-step (its column solve, ring
  // filter, post pass) still runs
  TGX_HIP(hipStreamWaitEvent(s0, F->ev[4], 0));
  return tgx_ng_duv(g, in, out, work, s0);  // the fork streams joined on return
}

// The same step strictly in the reference's order on one stream.
extern "C" int tgx_ng_step_serial(const tgx_ng_grid *g, const tgx_ng_fields *in,
                                  const tgx_ng_out *out, void *work, void *stream) {
  if (int rc = check(g, work)) return rc;
  if (!t_glb && g->scheme == TGX_SCHEME_AUTO && g->nk <= 128 && tuned_choice(g, R_OPLUS) == TGX_SCHEME_AUTO)
    if (int rc = tgx_ng_autotune(g, in, out, work, stream)) return rc;
  if (int rc = run_chem(g, in, out, work, (hipStream_t)stream)) return rc;
  if (int rc = run_oplus(g, in, out, work, stream, true, g->nstep_sub)) return rc;
  if (int rc = run_minors(g, in, out, work, stream, true)) return rc;
  if (int rc = tgx_ng_dt(g, in, out, work, stream)) return rc;
  if (int rc = tgx_ng_duv(g, in, out, work, stream)) return rc;
  return run_comp(g, in, out, work, stream, true);
}

// tgx_ng_step with comp in its precision form.
extern "C" int tgx_ng_step_mixed(const tgx_ng_grid *g, const tgx_ng_fields *in, const tgx_ng_out *out,
                                 void *work, void *stream, int nsweep) {
  if (int rc = check(g, work)) return rc;
  if (int rc = check_mixed(g, nsweep, "tgx_ng_step_mixed")) return rc;
  struct Scope {
    int prev = t_comp_nsweep;
    ~Scope() { t_comp_nsweep = prev; }
  } scope;
  t_comp_nsweep = nsweep;
  return tgx_ng_step(g, in, out, work, stream);
}

// ---------------------------------------------------------------------------
// The global grid (include/tgx_global.h): the routines above on a grid whose
// ghost columns are periodic points and whose ghost rows lie across the poles.
namespace {
int check_glb(const tgx_glb_grid *gg) {
  if (!gg) return tgx_set_error(TGX_ERR_ARG, "tgx_glb: null grid");
  const tgx_ng_grid &g = gg->g;
  if (g.nlon < 4 || (g.nlon & 1) || g.nlon > TGX_RING_MAX_NLON)
    return tgx_set_error(TGX_ERR_ARG, "tgx_glb: nlon %d must be even, 4..%d", g.nlon, TGX_RING_MAX_NLON);
  if (gg->nlat_filter < 0 || 2 * gg->nlat_filter > g.nlat || gg->nlat_filter > TGX_RING_MAX_ROWS / 2 ||
      (gg->nlat_filter > 0 && !gg->chunk))
    return tgx_set_error(TGX_ERR_ARG, "tgx_glb: nlat_filter %d (0..%d, at most half of nlat %d) or chunk table",
                         gg->nlat_filter, TGX_RING_MAX_ROWS / 2, g.nlat);
  for (int l = 0; l < gg->nlat_filter; ++l)
    if (gg->chunk[l] < 1 || gg->chunk[l] > g.nlon || g.nlon % gg->chunk[l] != 0)
      return tgx_set_error(TGX_ERR_ARG, "tgx_glb: chunk count %d does not divide nlon %d", gg->chunk[l], g.nlon);
  return 0;
}
// Sets this thread's global grid for the duration of one entry point.
struct GlbScope {
  const tgx_glb_grid *prev;
  const double *prev_tab;
  GlbScope(const tgx_glb_grid *gg, const double *tab) : prev(t_glb), prev_tab(t_glb_tab) {
    t_glb = gg;
    t_glb_tab = tab;
    tgx_ng_set_global_mode(1, gg->nlat_filter);
  }
  ~GlbScope() {
    t_glb = prev;
    t_glb_tab = prev_tab;
    tgx_ng_set_global_mode(prev != nullptr, prev ? prev->nlat_filter : 0);
  }
};
// The ring filters' table follows the nest form's workspace.
double *glb_table(const tgx_ng_grid *g, void *work) {
  return (double *)((char *)work + tgx_ng_workspace_bytes(g));
}
int glb_call(RoutineFn fn, const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
             void *stream) {
  if (int rc = check_glb(gg)) return rc;
  if (!work) return tgx_set_error(TGX_ERR_ARG, "tgx_glb: null workspace");
  double *tab = glb_table(&gg->g, work);
  if (int rc = tgx_ring_table_dev(tab, gg->g.nlon, kGlbTableOrder, gg->g.dlamda, (hipStream_t)stream)) return rc;
  GlbScope scope(gg, tab);
  return fn(&gg->g, in, out, work, stream);
}
}  // namespace

extern "C" unsigned long long tgx_glb_workspace_bytes(const tgx_glb_grid *gg) {
  if (!gg) return 0;
  return tgx_ng_workspace_bytes(&gg->g) + sizeof(double) * 2ull * kGlbTableOrder * (unsigned long long)gg->g.nlon;
}
extern "C" int tgx_glb_oplus(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                             void *stream) {
  return glb_call(tgx_ng_oplus, gg, in, out, work, stream);
}
extern "C" int tgx_glb_minors(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out,
                              void *work, void *stream) {
  return glb_call(tgx_ng_minors, gg, in, out, work, stream);
}
extern "C" int tgx_glb_dt(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                          void *stream) {
  return glb_call(tgx_ng_dt, gg, in, out, work, stream);
}
extern "C" int tgx_glb_duv(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                           void *stream) {
  return glb_call(tgx_ng_duv, gg, in, out, work, stream);
}
extern "C" int tgx_glb_comp(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                            void *stream) {
  return glb_call(tgx_ng_comp, gg, in, out, work, stream);
}
extern "C" int tgx_glb_step(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                            void *stream) {
  return glb_call(tgx_ng_step, gg, in, out, work, stream);
}
extern "C" int tgx_glb_step_serial(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out,
                                   void *work, void *stream) {
  return glb_call(tgx_ng_step_serial, gg, in, out, work, stream);
}
