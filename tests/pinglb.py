"""Independent checks that pin the global-grid forms of the oracle's routines
(dt.F, duv.F, minor.F with the comp_{no,n4s,n2d,ar}.F wrappers, oplus.F,
comp_o2o.F, comp.F), CPU only.

The oracle runs each routine's global form as its nest form with the global
differences of the reference (tgx_oracle.h, tgx_ogrid.glb).  The nest forms
are pinned by tests/pinlib.py; these checks hold the global forms to:

* periodicity: a global grid has no preferred longitude, so rotating the
  state by whole filter chunks rotates the solution (this fails if a lateral
  boundary is applied or the periodic points are taken from the wrong
  longitudes);
* the pinned nest form: with the ring filter off, the global solution equals
  the nest form's wherever the nest form's one-sided edge stencils and
  lateral data do not reach, except for the global differences themselves,
  which are checked separately:
  - comp.F's top interface (temperature from the midpoint below it, mean
    mass averaged over the two levels beside it): the global solution does
    not depend on the top midpoint level's temperature at all, and equals
    the nest form's once that temperature equals the level below and the
    upper-boundary He flux is rescaled by the ratio of the two interface
    masses;
  - minor.F's phi_x(N2) in Blanc's sum (minor.F:359-362): the molecular
    diffusion coefficient recovered from the rows against the model
    description's formula (pinlib's check, without the nest form's quirk);
  - the top levels: spval for TN, UN, VN (dt.F:407-408, duv.F:385-386,
    :512-513), zero for O+ (oplus.F:922-923);
* the ring filter's place: on the latitudes the chunk table filters, the
  filtered solution is the ring filter (pinned by test_oracle_ringfilter.py)
  of the solution computed with the filter off, and the n-1 field is the
  time smoothing of the filtered solution (the reference filters between
  the solve and the time smoothing);
* the periodic points of every output repeat the interior longitudes.

`CHECKS`, `TOL` and `MUTATIONS` are used as in pinlib; the planted mistakes
are the MUT(801..809) sites of oracle/column.c.
"""
import numpy as np

import helpers
import pinlib

DTSMOOTH = 0.95
DTSMOOTH_DIV2 = 0.5 * (1. - DTSMOOTH)
SPVAL = 1.e36


def pkg():
    return helpers.load_pkg()


def oracle():
    return helpers.load_oracle()


def global_state(name="small", nstep_sub=1, **kw):
    """(grid, Fortran-layout numpy state) of a configuration run as a global grid."""
    p = pkg()
    g = p.synth.make_grid(name, glb=True, nstep_sub=nstep_sub, **kw)
    Fo = p.synth.to_fortran_layout(p.synth.make_global_state(g))
    return g, Fo


def fill_halos(g, Fo):
    """The global halos of a Fortran-layout state (synth.fill_global_halos on
    (nlat+4, nlon+4[, nk]) arrays)."""
    p = pkg()
    nlon, nlat = g["nlon"], g["nlat"]
    nlond, nlatd = nlon + 4, nlat + 4
    src = [((i - 2 + nlon // 2) % nlon) + 2 for i in range(nlond)]
    for n, a in Fo.items():
        if n.startswith(("lonb_", "latb_")) or a.shape[:2] != (nlatd, nlond):
            continue
        a[:, 0] = a[:, nlon]
        a[:, 1] = a[:, nlon + 1]
        a[:, nlon + 2] = a[:, 2]
        a[:, nlon + 3] = a[:, 3]
        sign = -1. if n in p.synth.POLE_ODD else 1.
        for jnew, jold in ((1, 2), (0, 3), (nlat + 2, nlat + 1), (nlat + 3, nlat)):
            a[jnew] = sign * a[jold][src]
    return Fo


def rotate(g, Fo, s):
    """The state rotated by s longitudes (interior columns rolled, halos refilled)."""
    nlon, nlond, nlatd = g["nlon"], g["nlon"] + 4, g["nlat"] + 4
    out = {}
    for n, a in Fo.items():
        if n.startswith(("lonb_", "latb_")) or a.shape[:2] != (nlatd, nlond):
            out[n] = a.copy()
            continue
        b = a.copy()
        b[:, 2:nlon + 2] = np.roll(a[:, 2:nlon + 2], s, axis=1)
        out[n] = b
    return fill_halos(g, out)


def rot_out(g, a, s):
    nlon = g["nlon"]
    b = a.copy()
    b[:, 2:nlon + 2] = np.roll(a[:, 2:nlon + 2], s, axis=1)
    return b


def no_filter(g):
    return pinlib.grid_with(g, chunk=[])


def run(g, Fo, routine):
    """The routine's outputs (dict of Fortran-layout arrays) on the oracle."""
    st = oracle().OracleState(g, Fo)
    if routine == "dt":
        return st.dt_ng(Fo["tlbc"])
    if routine == "duv":
        return st.duv_ng(Fo["tn"] * (1. + 1e-3), 1e-4 * Fo["un"], 1e-4 * Fo["vn"], Fo["tlbc"], Fo["ulbc"],
                         Fo["vlbc"])
    if routine == "minor":
        return st.minors(Fo["tlbc"])
    if routine == "oplus":
        return st.oplus_subcycled(Fo["lonb_op"], Fo["latb_op"])
    if routine == "comp_o2o":
        fs = st.comp_o2o_ng()
        return {"fs%d%d" % (m, s): fs[m, s] for m in range(4) for s in range(3)}
    if routine == "comp":
        return st.comp_ng(st.comp_o2o_ng(), Fo["tlbc"], Fo["flx_he"])
    raise ValueError(routine)


SOLVED = {"dt": ("tn",), "duv": ("un", "vn"), "minor": ("n2d", "n4s", "no", "ar"), "oplus": ("op",),
          "comp": ("o2", "o1", "he")}
NM = {"tn": "tn_nm", "un": "un_nm", "vn": "vn_nm", "n2d": "n2d_nm", "n4s": "n4s_nm", "no": "no_nm",
      "ar": "ar_nm", "op": "op_nm", "o2": "o2_nm", "o1": "o1_nm", "he": "he_nm"}
# levels the routine solves (and filters): all but the top for dt, duv, oplus
NLEV_CUT = {"dt": 1, "duv": 1, "minor": 0, "oplus": 1, "comp": 0}
FILTER = {"dt": (3, False), "duv": (3, False), "minor": (3, False), "oplus": (3, True), "comp": (1, False)}


def _interior(g, a, jm=0, im=0):
    nlon, nlat = g["nlon"], g["nlat"]
    return a[2 + jm:nlat + 2 - jm, 2 + im:nlon + 2 - im]


def _frel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.))


# ---- periodicity ------------------------------------------------------------------
def rotation(routine):
    g, Fo = global_state()
    s = g["nlon"] // min(g["chunk"])     # whole chunks of every filtered latitude
    out = run(g, Fo, routine)
    rot = run(g, rotate(g, Fo, s), routine)
    cut = NLEV_CUT.get(routine, 1)
    e = 0.
    for n, a in out.items():
        if a.ndim != 3:
            continue
        nk = a.shape[2]
        e = max(e, _frel(_interior(g, rot[n])[..., :nk - cut], _interior(g, rot_out(g, a, s))[..., :nk - cut]))
    return {routine + "_rotation": e}


def periodic_points(routine):
    g, Fo = global_state()
    out = run(g, Fo, routine)
    nlon = g["nlon"]
    e = 0.
    for n, a in out.items():
        if a.ndim != 3 or routine == "comp_o2o":
            continue
        r = a[2:g["nlat"] + 2]
        for ip, isrc in ((0, nlon), (1, nlon + 1), (nlon + 2, 2), (nlon + 3, 3)):
            e = max(e, float(np.abs(r[:, ip] - r[:, isrc]).max()))
    return {routine + "_periodic": e}


# ---- the nest form, away from its edges -----------------------------------------------
def matches_nest_form(routine):
    g, Fo = global_state()
    gn = no_filter(g)
    out = run(gn, Fo, routine)
    ref = run(pinlib.grid_with(gn, glb=False), Fo, routine)
    cut = NLEV_CUT.get(routine, 1)
    e = 0.
    for n in SOLVED.get(routine, ()) + tuple(NM[x] for x in SOLVED.get(routine, ())):
        nk = out[n].shape[2]
        e = max(e, _frel(_interior(g, out[n], 2, 2)[..., :nk - cut], _interior(g, ref[n], 2, 2)[..., :nk - cut]))
    if routine == "comp_o2o":
        for n in out:
            nk = out[n].shape[2]
            e = max(e, _frel(_interior(g, out[n])[..., :nk - 1], _interior(g, ref[n])[..., :nk - 1]))
    return {routine + "_vs_nest": e}


def comp_top_interface():
    """comp.F's top interface: no dependence on the top level's TN; the nest
    form once TN(top) = TN(top-1) and the He flux is rescaled by the ratio of
    the interface masses comp_ng.F90 and comp.F:632-633 use."""
    g, Fo = global_state()
    gn = no_filter(g)
    nk = g["nk"]
    base = run(gn, Fo, "comp")
    F2 = {n: a.copy() for n, a in Fo.items()}
    F2["tn"][..., nk - 1] *= 1.3
    hot = run(gn, F2, "comp")
    e_inv = max(_frel(_interior(g, hot[n]), _interior(g, base[n])) for n in ("o2", "o1", "he", "o2_nm", "he_nm"))
    F3 = {n: a.copy() for n, a in Fo.items()}
    F3["tn"][..., nk - 1] = F3["tn"][..., nk - 2]
    glb = run(gn, F3, "comp")
    mb = F3["mbar"]
    F4 = {n: a.copy() for n, a in F3.items()}
    F4["flx_he"] = np.ascontiguousarray(F3["flx_he"] * (mb[..., nk - 3] + mb[..., nk - 2]) /
                                        (mb[..., nk - 2] + mb[..., nk - 1]))
    ref = run(pinlib.grid_with(gn, glb=False), F4, "comp")
    e_eq = max(_frel(_interior(g, glb[n], 2, 2), _interior(g, ref[n], 2, 2))
               for n in ("o2", "o1", "he", "o2_nm", "o1_nm", "he_nm"))
    return {"comp_top_tn_independent": e_inv, "comp_top_vs_nest": e_eq}


def minor_blanc():
    """pinlib's molecular-diffusion check of the minor species on the global
    form: Blanc's sum with phi_x(N2), the model description's (no quirk)."""
    errs = {}
    for sp in (0, 3):
        g, Fo, j, i, dmdz = pinlib._minor_state()
        g = pinlib.grid_with(g, glb=True, chunk=[])
        cap, _ = pinlib._minor_rows(g, Fo, sp)
        nk, dz = g["nk"], g["dz"]
        p, r = cap[0, j, i], cap[2, j, i]
        A, _ = pinlib._minor_AE(g, Fo, j, i, dmdz, sp, nest_quirk=False)
        k = np.arange(1, nk - 1)
        Ar = .5 * (p[k] + r[k - 1]) * dz ** 2
        errs["minor_glb_A_" + pinlib.MINOR[sp][0]] = pinlib.rel(Ar, A[k])
    return errs


# ---- the top levels ----------------------------------------------------------------
TOP = {"dt": {"tn": SPVAL, "tn_nm": SPVAL}, "duv": {"un": SPVAL, "vn": SPVAL, "un_nm": SPVAL, "vn_nm": SPVAL},
       "oplus": {"op": 0., "op_nm": 0.}}


def top_levels(routine):
    g, Fo = global_state()
    out = run(g, Fo, routine)
    e = 0.
    for n, v in TOP[routine].items():
        t = _interior(g, out[n])[..., -1]
        e = max(e, float(np.abs(t - v).max() / (v if v else 1.)))
    return {routine + "_top": e}


# ---- the ring filter's place -------------------------------------------------------------
def filter_place(routine):
    g, Fo = global_state()
    o = oracle()
    out = run(g, Fo, routine)
    raw = run(no_filter(g), Fo, routine)
    order, opl = FILTER[routine]
    nlon, nlat, nk = g["nlon"], g["nlat"], g["nk"]
    nlev = nk - NLEV_CUT[routine]
    if routine == "comp":
        # the levels below the N2 extrapolation, where the normalisation that
        # follows the filter is the identity up to rounding (comp.F:975-1022)
        nlev = nk - 3
    e_f = e_nm = 0.
    for lat in range(1, nlat + 1):
        nch = o.ringfilter_chunks(lat, nlat, g["chunk"])
        if nch == 0:
            continue
        j = lat + 1
        # NO and N4S take this step's (filtered) N2D in their sources, so
        # only the species solved first compare with the unfiltered run
        for n in (("n2d", "ar") if routine == "minor" else SOLVED[routine]):
            row = np.ascontiguousarray(raw[n][j, :, :nlev].T)          # (nlev, nlon+4)
            o.ringfilter_row(row, nlon, order, opl, nch, g["dlamda"])
            e_f = max(e_f, _frel(out[n][j, :, :nlev], row.T))
            nm = DTSMOOTH_DIV2 * (Fo[NM[n]][j, :, :nlev] + out[n][j, :, :nlev]) + DTSMOOTH * Fo[n][j, :, :nlev]
            e_nm = max(e_nm, _frel(out[NM[n]][j, :, :nlev], nm))
    return {routine + "_filtered": e_f, routine + "_nm_of_filtered": e_nm}


def _checks(routine, extra=()):
    fns = [lambda r=routine: rotation(r), lambda r=routine: periodic_points(r)]
    # minor: phi_x(N2) changes every row (minor_blanc checks it); comp: the top
    # interface changes every column (comp_top_interface checks it)
    if routine not in ("minor", "comp"):
        fns.append(lambda r=routine: matches_nest_form(r))
    if routine in TOP:
        fns.append(lambda r=routine: top_levels(r))
    if routine in FILTER:
        fns.append(lambda r=routine: filter_place(r))
    return tuple(fns) + tuple(extra)


CHECKS = {
    "dt": _checks("dt"),
    "duv": _checks("duv"),
    "minor": _checks("minor", (minor_blanc,)),
    "oplus": _checks("oplus"),
    "comp_o2o": _checks("comp_o2o"),
    "comp": _checks("comp", (comp_top_interface,)),
}

TOL = {}
for _r in CHECKS:
    TOL[_r + "_rotation"] = 1e-12          # the filter's Fourier sums see the longitudes in another order
    TOL[_r + "_periodic"] = 0.0
    TOL[_r + "_vs_nest"] = 1e-14
    TOL[_r + "_top"] = 0.0
    TOL[_r + "_filtered"] = 1e-13
    TOL[_r + "_nm_of_filtered"] = 1e-13
TOL["comp_vs_nest"] = 1e-12                # the top interface differs (checked below)
TOL.update({"comp_top_tn_independent": 0.0, "comp_top_vs_nest": 1e-12,
            "minor_glb_A_no": 1e-13, "minor_glb_A_ar": 1e-13})

MUTATIONS = {
    "dt": {802: "TN not ring-filtered", 803: "periodic point 0 from the wrong longitude",
           806: "top level extrapolated instead of spval"},
    "minor": {801: "Blanc's sum with phi_x(He) (the nest form's)", 804: "minor species not ring-filtered",
              805: "n-1 level smoothed from the unfiltered solution"},
    "oplus": {807: "top level held instead of zero"},
    "comp": {808: "top interface temperature averaged with the top level",
             809: "He flux normalised by the interface below the top"},
}


def run_checks(routine):
    errs = {}
    for fn in CHECKS[routine]:
        errs.update(fn())
    return errs


def failures(errs):
    return {n: e for n, e in errs.items() if not (e <= TOL[n])}
