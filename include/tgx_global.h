/*
 * libtgx — the fused implicit column step on the global grid, device resident.
 *
 * Replaces the assembly+solve of the global model's routines: dt
 * (dt.F:2-4), duv (duv.F:2-4), minor with comp_{n2d,n4s,no,ar} (minor.F:2-4
 * and the comp_*.F wrappers), oplus with its nstep_sub sub-cycles (oplus.F:22-24,
 * dynamics.F:319-359), and comp_o2o + comp (comp_o2o.F:2-4, comp.F:2-4), each
 * with its ring filter (filter_tn, filter_uv, filter_minor, filter_op,
 * filter_o2o: ringfilter.F) between the column solve and the time smoothing.
 *
 * Layout: the level planes of tgx_step.h with nlon, nlat the global grid's
 * longitudes and latitudes.  The two ghost columns on each side hold the
 * periodic points (i = 0, 1 repeat i = nlon, nlon+1; i = nlon+2, nlon+3
 * repeat i = 2, 3; i = longitude index + 1 of the reference's 1..nlon+4), the
 * two ghost rows on each side the latitudes across the pole (row j =
 * latitude + 1).  The caller fills both in every input before a call, as the
 * model does before a step (mk_polelat.F, mp_periodic_f4d); tgx_glb_halos
 * does it on the device.  The grid's latitude arrays (cs, racs, tanphi, cor,
 * glat) have nlat+4 entries including the ghost rows, where cos(latitude)
 * changes sign (cons.F:314-317).  Lateral data (lonb_*, latb_*) are not read.
 *
 * Outputs: the interior rows (all nlon+4 columns, the periodic points copied
 * from the interior as mp_periodic_f3d does); the top level is spval (1e36)
 * for TN, UN, VN and their n-1 fields (dt.F:407-408, duv.F:385-386,
 * :512-513) and zero for O+ (oplus.F:922-923).  Outputs on the ghost rows are
 * not defined; call tgx_glb_halos on the fields a next step reads.  Between
 * O+ sub-cycles the ghost rows of O+ keep the step input's (the reference
 * exchanges none there, oplus.F:942).
 *
 * The global forms run the Thomas scheme (g.scheme is ignored).  All pointers
 * except `chunk` are device pointers; every entry point is stream-ordered on
 * the hipStream_t passed as void* and returns 0 or a TGX_ERR_* code.
 */
#ifndef TGX_GLOBAL_H
#define TGX_GLOBAL_H

#include "tgx_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  tgx_ng_grid g;      /* as for a nest domain; nlon even */
  int nlat_filter;    /* latitudes filtered next to each pole (cons.F:331-355) */
  const int *chunk;   /* HOST array [nlat_filter]: chunk counts, pole first;
                         each divides nlon */
} tgx_glb_grid;

/* Workspace bytes: those of tgx_ng_workspace_bytes(&gg->g), then the ring
 * filters' table of cos/sin values (6 nlon doubles). */
unsigned long long tgx_glb_workspace_bytes(const tgx_glb_grid *gg);

/* The routines on their own, in the model's order (dynamics.F:319-1000):
 * oplus, minors (N2D, N4S, NO, Ar), dt, duv (needs this step's tn from dt and
 * Fe, Fn from oplus), comp (comp_o2o's fs formed inside). */
int tgx_glb_oplus(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                  void *stream);
int tgx_glb_minors(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                   void *stream);
int tgx_glb_dt(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
               void *stream);
int tgx_glb_duv(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                void *stream);
int tgx_glb_comp(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                 void *stream);
/* The whole column step (dynamics.F:319-1000): the independent routines on
 * library streams forked from and joined into `stream`, results identical to
 * tgx_glb_step_serial (the reference's order on one stream). */
int tgx_glb_step(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out, void *work,
                 void *stream);
int tgx_glb_step_serial(const tgx_glb_grid *gg, const tgx_ng_fields *in, const tgx_ng_out *out,
                        void *work, void *stream);

/* The halos of one global field in level planes (nk levels, row pitch ld):
 * the periodic points of the interior rows from the interior longitudes
 * (mp_periodic_f4d), and the two rows across each pole from the interior
 * rows next to it shifted by 180 degrees of longitude, times `sign` (-1 for
 * the wind components, fields.F:548-553; mk_polelat.F:2-57).  nk = 1 for a
 * 2-D field. */
int tgx_glb_halos(double *f, int nk, int nlon, int nlat, int ld, double sign, void *stream);

#ifdef __cplusplus
}
#endif
#endif
