// Host-layout boundary of the fused step: a context that holds one nest
// domain's fields on the device in the library's level planes, filled and
// read back by name from arrays in the model's own layout (Fortran, level
// fastest: 3-D f(nk, lond0:lond1, latd0:latd1), 2-D f(lond0:lond1,
// latd0:latd1); fields_ng.F90:226-261).  The Fortran drop-ins with the
// reference's routine signatures (fortran/ng_shims.F90) are written on it.
// 3-D fields go through the LDS transpose of pack.hip; see include/tgx_step.h.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tgx.h"
#include "../../include/tgx_step.h"
#include "tgx_common.h"

// nrun O+ sub-steps (step_api.cpp)
int tgx_ng_oplus_substeps(const tgx_ng_grid *, const tgx_ng_fields *, const tgx_ng_out *, void *, void *, int);

namespace {

enum Kind { K3 = 3, K2 = 2, KLONB = 4, KLATB = 5, KLONB_OP = 6, KLATB_OP = 7 };
struct Ref {
  const char *name;
  size_t off;
  int kind;
};

#define I3(n) {#n, offsetof(tgx_ng_fields, n), K3}
#define I2(n) {#n, offsetof(tgx_ng_fields, n), K2}
#define IL(n) {"lonb_" #n, offsetof(tgx_ng_fields, lonb_##n), KLONB}, {"latb_" #n, offsetof(tgx_ng_fields, latb_##n), KLATB}
const Ref kInputs[] = {
    I3(tn), I3(tn_nm), I3(un), I3(vn), I3(un_nm), I3(vn_nm), I3(o2), I3(o2_nm), I3(o1), I3(o1_nm), I3(he),
    I3(he_nm), I3(op), I3(op_nm), I3(n2d), I3(n2d_nm), I3(n4s), I3(n4s_nm), I3(no), I3(no_nm), I3(ar),
    I3(ar_nm), I3(ti), I3(te), I3(ne), I3(o2p), I3(w), I3(z), I3(mbar), I3(barm), I3(xnmbar), I3(xnmbari),
    I3(scht), I3(schti), I3(n2), I3(w_upd), I3(cp), I3(kt), I3(km), I3(hdt), I3(hdu), I3(hdv), I3(hdo2),
    I3(hdo1), I3(hdhe), I3(qji_tn), I3(cool_imp), I3(cool_exp), I3(qtotal), I3(rkm12), I3(ui), I3(vi), I3(wi),
    I3(lxx), I3(lyy), I3(lxy), I3(lyx), I3(qop2p), I3(qop2d), I3(qop), I3(qo2p), I3(qtef), I3(rj), I3(rk1),
    I3(rk2), I3(rk3), I3(rk19), I3(rk20), I3(rk25), I3(beta1), I3(beta3), I3(beta5), I3(beta8), I3(beta9),
    I3(beta9n), I3(beta17), I3(ra1), I3(ra2), I3(ra3), I3(n2p), I3(nplus), I3(nop),
    I2(t_lbc), I2(u_lbc), I2(v_lbc), I2(tlbc), I2(ulbc), I2(vlbc), I2(flx_he), I2(bx), I2(by), I2(bz),
    I2(bmod2), I2(dipmag), I2(sndec), I2(csdec), I2(chi), I2(rlatm),
    IL(tn), IL(un), IL(vn), IL(o2), IL(o1), IL(he), IL(n2d), IL(n4s), IL(no), IL(ar),
    {"lonb_op", offsetof(tgx_ng_fields, lonb_op), KLONB_OP}, {"latb_op", offsetof(tgx_ng_fields, latb_op), KLATB_OP},
};
#undef I3
#undef I2
#undef IL
#define O3(n) {#n, offsetof(tgx_ng_out, n), K3}
#define O2(n) {#n, offsetof(tgx_ng_out, n), K2}
const Ref kOutputs[] = {
    O3(tn), O3(tn_nm), O3(un), O3(un_nm), O3(vn), O3(vn_nm), O3(n2d), O3(n2d_nm), O3(n4s), O3(n4s_nm), O3(no),
    O3(no_nm), O3(ar), O3(ar_nm), O3(op), O3(op_nm), O3(xiop2p), O3(xiop2d), O3(Fe), O3(Fn), O3(o2), O3(o2_nm),
    O3(o1), O3(o1_nm), O3(he), O3(he_nm),
    O2(tlbc), O2(tlbc_nm), O2(ulbc), O2(vlbc), O2(ulbc_nm), O2(vlbc_nm), O2(flx_he),
};
#undef O3
#undef O2

struct Ctx {
  tgx_ng_grid g;
  tgx_ng_fields in;
  tgx_ng_out out;
  std::vector<void *> allocs;
  void *work = nullptr;
  double *stage = nullptr;  // one 3-D field in the model's layout
  double *lonb_op1 = nullptr, *latb_op1 = nullptr;  // O+ lateral data of one sub-step
  hipStream_t s = nullptr;
  int dev = 0;
  int nlond = 0, nlatd = 0;
};

size_t elems(const Ctx &c, int kind) {
  const size_t nk = c.g.nk, ld = c.g.ld, nlatd = c.nlatd, nsub = c.g.nstep_sub;
  switch (kind) {
    case K3: return nk * ld * nlatd;
    case K2: return ld * nlatd;
    case KLONB: return nk * 2 * nlatd;
    case KLATB: return nk * 2 * ld;
    case KLONB_OP: return nsub * nk * 2 * nlatd;
    default: return nsub * nk * 2 * ld;
  }
}

int dev_alloc(Ctx &c, size_t n, double **p) {
  void *q = nullptr;
  TGX_HIP(hipMalloc(&q, n * sizeof(double)));
  c.allocs.push_back(q);
  TGX_HIP(hipMemsetAsync(q, 0, n * sizeof(double), c.s));
  *p = (double *)q;
  return 0;
}

int dev_copy_1d(Ctx &c, const double *host, size_t n, const double **slot) {
  if (!host) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_create: a level or latitude array is NULL");
  double *d = nullptr;
  if (int rc = dev_alloc(c, n, &d)) return rc;
  TGX_HIP(hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, c.s));
  *slot = d;
  return 0;
}

const Ref *find(const Ref *tab, size_t n, const char *name) {
  for (size_t i = 0; i < n; ++i)
    if (std::strcmp(tab[i].name, name) == 0) return &tab[i];
  return nullptr;
}

double **slot_of(void *base, const Ref *r) { return (double **)((char *)base + r->off); }

void release(Ctx *c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->work) (void)hipFree(c->work);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

extern "C" int tgx_ng_pack_levels(const double *, double *, int, int, int, int, void *);
extern "C" int tgx_ng_unpack_levels(const double *, double *, int, int, int, int, void *);

// Host (model layout) -> device planes for a field of kind `kind`.
int put_any(Ctx &c, int kind, const double *src, double *dst, bool src_on_device) {
  const int nk = c.g.nk, nlond = c.nlond, nlatd = c.nlatd, ld = c.g.ld;
  const hipMemcpyKind k = src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (kind == K3) {
    const double *from = src;
    if (!src_on_device) {
      TGX_HIP(hipMemcpyAsync(c.stage, src, sizeof(double) * nk * (size_t)nlond * nlatd, k, c.s));
      from = c.stage;
    }
    return tgx_ng_pack_levels(from, dst, nk, nlond, nlatd, ld, c.s);
  }
  if (kind == K2) {
    TGX_HIP(hipMemcpy2DAsync(dst, sizeof(double) * ld, src, sizeof(double) * nlond, sizeof(double) * nlond, nlatd,
                             k, c.s));
    return 0;
  }
  // lateral data: the model's lon_b(nk, 2, latd) and lat_b(nk, lond, 2), one
  // block per O+ sub-step for O+, reordered on the host into the library's
  // lonb[(k*2 + side)*nlatd + j] and latb[(k*2 + side)*ld + i]
  if (src_on_device) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put_dev: lateral data are taken from the host");
  const bool lon = kind == KLONB || kind == KLONB_OP;
  const int nblk = (kind == KLONB_OP || kind == KLATB_OP) ? c.g.nstep_sub : 1;
  const size_t per = lon ? (size_t)nk * 2 * nlatd : (size_t)nk * 2 * ld;
  std::vector<double> h(per * nblk, 0.);
  for (int b = 0; b < nblk; ++b)
    for (int k = 0; k < nk; ++k)
      for (int side = 0; side < 2; ++side) {
        if (lon)
          for (int j = 0; j < nlatd; ++j)
            h[b * per + ((size_t)k * 2 + side) * nlatd + j] = src[(size_t)b * nk * 2 * nlatd + k + (size_t)nk * (side + 2 * (size_t)j)];
        else
          for (int i = 0; i < nlond; ++i)
            h[b * per + ((size_t)k * 2 + side) * ld + i] = src[(size_t)b * nk * nlond * 2 + k + (size_t)nk * (i + (size_t)nlond * side)];
      }
  TGX_HIP(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c.s));
  TGX_HIP(hipStreamSynchronize(c.s));  // h goes out of scope
  return 0;
}

}  // namespace

extern "C" void *tgx_ng_ctx_create(const tgx_ng_grid *gh) {
  tgx_clear_error();
  if (!gh || gh->nk < 4 || gh->nlon < 2 || gh->nlat < 2 || gh->nstep_sub < 1) {
    tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_create: bad grid");
    return nullptr;
  }
  Ctx *c = new Ctx();
  c->g = *gh;
  c->nlond = gh->nlon + 4;
  c->nlatd = gh->nlat + 4;
  c->g.ld = (c->nlond + 15) / 16 * 16;
  c->g.capture = nullptr;
  auto fail = [&]() -> void * {
    release(c);
    return nullptr;
  };
  if (hipGetDevice(&c->dev) != hipSuccess) {
    tgx_set_error(TGX_ERR_NODEV, "tgx_ng_ctx_create: no HIP device");
    return fail();
  }
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    tgx_set_error(TGX_ERR_HIP, "tgx_ng_ctx_create: hipStreamCreate failed");
    return fail();
  }
  const size_t nk = gh->nk, nlatd = c->nlatd;
  if (dev_copy_1d(*c, gh->expz, nk, &c->g.expz) || dev_copy_1d(*c, gh->difk, nk, &c->g.difk) ||
      dev_copy_1d(*c, gh->dift, nk, &c->g.dift) || dev_copy_1d(*c, gh->xmue, nk, &c->g.xmue) ||
      dev_copy_1d(*c, gh->zpint, nk, &c->g.zpint) || dev_copy_1d(*c, gh->zpmid, nk, &c->g.zpmid) ||
      dev_copy_1d(*c, gh->racs, nlatd, &c->g.racs) || dev_copy_1d(*c, gh->cs, nlatd, &c->g.cs) ||
      dev_copy_1d(*c, gh->tanphi, nlatd, &c->g.tanphi) || dev_copy_1d(*c, gh->cor, nlatd, &c->g.cor) ||
      dev_copy_1d(*c, gh->glat, nlatd, &c->g.glat))
    return fail();
  for (const Ref &r : kInputs)
    if (dev_alloc(*c, elems(*c, r.kind), slot_of(&c->in, &r))) return fail();
  for (const Ref &r : kOutputs)
    if (dev_alloc(*c, elems(*c, r.kind), slot_of(&c->out, &r))) return fail();
  if (dev_alloc(*c, elems(*c, K3), &c->stage) || dev_alloc(*c, elems(*c, KLONB), &c->lonb_op1) ||
      dev_alloc(*c, elems(*c, KLATB), &c->latb_op1))
    return fail();
  const unsigned long long wb = tgx_ng_workspace_bytes(&c->g);
  if (hipMalloc(&c->work, wb) != hipSuccess) {
    tgx_set_error(TGX_ERR_HIP, "tgx_ng_ctx_create: hipMalloc of %llu bytes of workspace", wb);
    return fail();
  }
  if (hipStreamSynchronize(c->s) != hipSuccess) {
    tgx_set_error(TGX_ERR_HIP, "tgx_ng_ctx_create: set-up failed");
    return fail();
  }
  return c;
}

extern "C" int tgx_ng_ctx_destroy(void *ctx) {
  release((Ctx *)ctx);
  return 0;
}

extern "C" int tgx_ng_ctx_put(void *ctx, const char *name, const double *host) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !name || !host) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put: null argument");
  int kind = 0;
  double *dst = nullptr;
  if (const Ref *r = find(kInputs, sizeof kInputs / sizeof *kInputs, name)) {
    kind = r->kind;
    dst = *slot_of(&c->in, r);
  } else if (std::strcmp(name, "lonb_op_sub") == 0) {
    kind = KLONB;
    dst = c->lonb_op1;
  } else if (std::strcmp(name, "latb_op_sub") == 0) {
    kind = KLATB;
    dst = c->latb_op1;
  } else {
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put: no input field named '%s'", name);
  }
  if (int rc = put_any(*c, kind, host, dst, false)) return rc;
  TGX_HIP(hipStreamSynchronize(c->s));  // the host array may change after return
  return 0;
}

// A result of an earlier routine that a later one reads (duv: this step's TN
// from dt, Fe and Fn from oplus), set from the model's copy.
extern "C" int tgx_ng_ctx_put_out(void *ctx, const char *name, const double *host) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !name || !host) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put_out: null argument");
  const Ref *r = find(kOutputs, sizeof kOutputs / sizeof *kOutputs, name);
  if (!r) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put_out: no output field named '%s'", name);
  if (int rc = put_any(*c, r->kind, host, *slot_of(&c->out, r), false)) return rc;
  TGX_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int tgx_ng_ctx_put_dev(void *ctx, const char *name, const double *dev, void *stream) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !name || !dev) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put_dev: null argument");
  const Ref *r = find(kInputs, sizeof kInputs / sizeof *kInputs, name);
  if (!r || (r->kind != K3 && r->kind != K2))
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_put_dev: no 3-D or 2-D input field named '%s'", name);
  // order the context's stream after the caller's work on the source
  hipEvent_t e;
  TGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t er = hipEventRecord(e, (hipStream_t)stream);
  if (er == hipSuccess) er = hipStreamWaitEvent(c->s, e, 0);
  (void)hipEventDestroy(e);
  if (er != hipSuccess) return tgx_set_error(TGX_ERR_HIP, "tgx_ng_ctx_put_dev: %s", hipGetErrorString(er));
  if (int rc = put_any(*c, r->kind, dev, *slot_of(&c->in, r), true)) return rc;
  TGX_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int tgx_ng_ctx_get(void *ctx, const char *name, double *host) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !name || !host) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_get: null argument");
  const Ref *r = find(kOutputs, sizeof kOutputs / sizeof *kOutputs, name);
  const double *src = nullptr;
  int kind = 0;
  if (r) {
    src = *slot_of(&c->out, r);
    kind = r->kind;
  } else if ((r = find(kInputs, sizeof kInputs / sizeof *kInputs, name)) && (r->kind == K3 || r->kind == K2)) {
    src = *slot_of(&c->in, r);
    kind = r->kind;
  } else {
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_get: no 3-D or 2-D field named '%s'", name);
  }
  const int nk = c->g.nk, nlond = c->nlond, nlatd = c->nlatd, ld = c->g.ld;
  if (kind == K3) {
    if (int rc = tgx_ng_unpack_levels(src, c->stage, nk, nlond, nlatd, ld, c->s)) return rc;
    TGX_HIP(hipMemcpyAsync(host, c->stage, sizeof(double) * nk * (size_t)nlond * nlatd, hipMemcpyDeviceToHost, c->s));
  } else {
    TGX_HIP(hipMemcpy2DAsync(host, sizeof(double) * nlond, src, sizeof(double) * ld, sizeof(double) * nlond, nlatd,
                             hipMemcpyDeviceToHost, c->s));
  }
  TGX_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int tgx_ng_ctx_run(void *ctx, const char *routine) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !routine) return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_run: null argument");
  typedef int (*Fn)(const tgx_ng_grid *, const tgx_ng_fields *, const tgx_ng_out *, void *, void *);
  struct {
    const char *name;
    Fn fn;
  } const tab[] = {{"oplus", tgx_ng_oplus}, {"minors", tgx_ng_minors}, {"dt", tgx_ng_dt},
                   {"duv", tgx_ng_duv},     {"comp", tgx_ng_comp},     {"step", tgx_ng_step}};
  for (const auto &t : tab)
    if (std::strcmp(t.name, routine) == 0) {
      if (int rc = t.fn(&c->g, &c->in, &c->out, c->work, c->s)) return rc;
      TGX_HIP(hipStreamSynchronize(c->s));
      return 0;
    }
  if (std::strcmp(routine, "oplus_sub") == 0) {  // one sub-step, lateral data lonb_op_sub / latb_op_sub
    tgx_ng_fields in1 = c->in;
    in1.lonb_op = c->lonb_op1;
    in1.latb_op = c->latb_op1;
    if (int rc = tgx_ng_oplus_substeps(&c->g, &in1, &c->out, c->work, c->s, 1)) return rc;
    TGX_HIP(hipStreamSynchronize(c->s));
    return 0;
  }
  return tgx_set_error(TGX_ERR_ARG, "tgx_ng_ctx_run: no routine named '%s'", routine);
}
