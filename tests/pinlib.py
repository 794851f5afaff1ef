"""Independent checks that pin the oracle's column-step routines (CPU only).

Each routine of oracle/column.c is a restatement of the reference's Fortran;
these checks hold it to things the restatement does not itself contain:

* the partial differential equations the reference discretises, as written in
  the reference's model description (doc/description/*.tex): interface
  coefficients recovered from the rows the oracle hands its column solve
  (the row-capture hook) are compared with the coefficients those equations
  give from the physical inputs;
* conservation: the vertical transport operators are in flux form, so every
  unknown's coefficients sum to zero over the rows it appears in; the
  chemistry of comp_o2o conserves mass reaction by reaction;
* symmetry: the longitude mirror image of a state (zonal components and the
  other odd fields negated) gives the mirror image of the solution;
* closed forms of degenerate cases (pointwise limits, decoupled components).

`CHECKS[routine]` lists the checks of a routine; each returns a dict of named
errors and `TOL` gives the bound of each name.  tests/test_oracle_pins.py runs
them on the oracle and on oracle copies with one deliberate mistake planted
(oracle.set_mutation), and shows each planted mistake fails at least one.
"""
import contextlib
import math

import numpy as np

import helpers

# cons.F:25-46 (cgs), as the model description states them
GRAV = 870.0          # cm/s^2
P0 = 5.0e-4           # reference pressure, microbar
GASK = 8.314e7        # gas constant, erg/K/mol
RE = 6.37122e8        # Earth radius, cm
AVO = 6.023e23
TSURPLUS = 5.11 * 1.602e-12   # f_T of the O recombination heating (dt.tex)


def pkg():
    return helpers.load_pkg()


def oracle():
    return helpers.load_oracle()


@contextlib.contextmanager
def mutation(mid):
    o = oracle()
    o.set_mutation(mid)
    try:
        yield
    finally:
        o.set_mutation(0)


# ---- states -------------------------------------------------------------------
def base_state(name="small", **grid_kw):
    """(grid, Fortran-layout numpy state) of a synthetic configuration."""
    p = pkg()
    g = p.synth.make_grid(name, **grid_kw)
    Fo = p.synth.to_fortran_layout(p.synth.make_state(g))
    return g, Fo


def uniform(g, Fo, at=None):
    """Every 3-D and 2-D field replaced by one column's values everywhere, so
    that horizontal advection vanishes and smoothing is the identity."""
    nlatd, nlond = g["nlat"] + 4, g["nlon"] + 4
    jc, ic = at if at else (nlatd // 2, nlond // 2)
    out = {}
    for n, a in Fo.items():
        if n.startswith("lonb_") or n.startswith("latb_"):
            out[n] = a.copy()
            continue
        if a.ndim == 3 and a.shape[:2] == (nlatd, nlond):
            out[n] = np.ascontiguousarray(np.broadcast_to(a[jc, ic], a.shape))
        elif a.ndim == 2 and a.shape == (nlatd, nlond):
            out[n] = np.full(a.shape, a[jc, ic])
        else:
            out[n] = a.copy()
    # lateral data equal to the uniform field itself
    for n in [k[5:] for k in out if k.startswith("lonb_") and k != "lonb_op"]:
        f = out[n]
        out["lonb_" + n] = np.ascontiguousarray(np.stack([f[:, 0, :], f[:, -1, :]], axis=1))
        out["latb_" + n] = np.ascontiguousarray(np.stack([f[0], f[-1]], axis=0))
    if "lonb_op" in out:
        f = out["op"]
        ns = out["lonb_op"].shape[0]
        out["lonb_op"] = np.ascontiguousarray(np.stack([np.stack([f[:, 0, :], f[:, -1, :]], axis=1)] * ns))
        out["latb_op"] = np.ascontiguousarray(np.stack([np.stack([f[0], f[-1]], axis=0)] * ns))
    return out


def set_levels(Fo, name, prof):
    """A 3-D field set to a vertical profile (nk,) in every column."""
    Fo[name] = np.ascontiguousarray(np.broadcast_to(np.asarray(prof, dtype=np.float64), Fo[name].shape))


def grid_with(g, **kw):
    g = dict(g)
    for k, v in kw.items():
        g[k] = v
    return g


# ---- running the oracle -----------------------------------------------------------
def capture(g, fn):
    """Runs fn() with the row-capture hook on; returns the planes as an array
    (68, nlatd, nlond, nk) and fn's result."""
    o = oracle()
    nlatd, nlond, nk = g["nlat"] + 4, g["nlon"] + 4, g["nk"]
    buf = np.zeros(o.CAPTURE_PLANES * nlatd * nlond * nk)
    o.set_capture(buf)
    try:
        res = fn()
    finally:
        o.set_capture(None)
    return buf.reshape(o.CAPTURE_PLANES, nlatd, nlond, nk), res


def state(g, Fo):
    return oracle().OracleState(g, Fo)


def rel(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


# ---- longitude mirror ---------------------------------------------------------------
# Fields odd under the mirror lon -> -lon: zonal components of vectors (winds,
# ion drift, the zonal magnetic component, the O+ zonal flux), the zonal
# horizontal diffusion, the off-diagonal ion-drag tensor (it couples the odd
# zonal and the even meridional wind), the sine of the declination and the
# Coriolis parameter (the planet's rotation reverses in a mirror).
ODD_3D = {"un", "un_nm", "ui", "hdu", "lxy", "lyx", "Fe"}
ODD_2D = {"u_lbc", "ulbc", "bx", "sndec"}


def mirror_lon(g, Fo):
    nlond = g["nlon"] + 4
    out = {}
    for n, a in Fo.items():
        if n.startswith("lonb_"):
            b = a[..., ::-1, :] if n != "lonb_op" else a[:, :, ::-1, :]
        elif n.startswith("latb_"):
            b = a[..., ::-1, :]
        elif a.ndim == 3 and a.shape[1] == nlond:
            b = a[:, ::-1, :]
        elif a.ndim == 2 and a.shape[1] == nlond:
            b = a[:, ::-1]
        else:
            b = a
        b = np.ascontiguousarray(b)
        base = n[5:] if n[:5] in ("lonb_", "latb_") else n
        if base in ODD_3D or base in ODD_2D:
            b = -b
        out[n] = b
    return out, grid_with(g, cor=[-c for c in g["cor"]])


def mirror_out(name, a):
    b = np.ascontiguousarray(a[:, ::-1] if a.ndim >= 2 else a)
    return -b if name in ("un", "un_nm", "Fe", "ulbc", "ulbc_nm") else b


# ================================================================================
# dt_ng (doc/description/dt.tex, eq. dt_1 and the implicit form that follows)
#
#   dT/dt = g e^z/(p0 Cp) d/dZ{ K_T/H dT/dZ + K_E H^2 Cp rho [g/Cp + 1/H dT/dZ] }
#           - v.grad T - W (dT/dZ + R* T/(Cp m)) + (Q - e^z L_exp)/Cp - L_imp T
#
# taken implicitly and multiplied by Cp e^-z (the code's expz = e^-Z at the
# midpoints): the row of midpoint k couples the interface fluxes at k and k+1,
#   G_k (T_k - T_k-1) + F_k (T_k + T_k-1) at interface k, with
#   G_k = g (K_T + K_E H^2 rho Cp) / (p0 H dz^2)
#   F_k = g^2 K_E H^2 rho / (2 dz p0 T_i)          (T_i = interface T at time n)
#   rho = p0 e^-Z_int / (g H)                       (ideal gas at the interface)
# so p_k = G_k - F_k, r_k-1 = G_k + F_k, and the remaining diagonal is
#   -e^-Z (Cp_mid (1/(2 dt) + L_imp) + W_mid R*/m).
def _dt_rows(g, Fo):
    st = state(g, Fo)
    cap, res = capture(g, lambda: st.dt_ng(Fo["tlbc"]))
    return cap[0:4], res


def _dt_physics(g, Fo, j, i):
    """Interface coefficients G_k, F_k (k = 0..nk-1) from the physical inputs."""
    nk, dz = g["nk"], g["dz"]
    H = Fo["schti"][j, i]
    zint = np.asarray(g["zpint"])
    rho = P0 * np.exp(-zint) / (GRAV * H)
    kt, cp, tn = Fo["kt"][j, i], Fo["cp"][j, i], Fo["tn"][j, i]
    ke = np.asarray(g["dift"])
    Ti = np.empty(nk)
    Ti[0] = Fo["tlbc"][j, i]
    Ti[1:] = .5 * (tn[:-1] + tn[1:])
    G = GRAV * (kt + ke * H ** 2 * rho * cp) / (P0 * H * dz ** 2)
    F = GRAV ** 2 * ke * H ** 2 * rho / (2 * dz * P0 * Ti)
    return G, F


def _dt_diag(g, Fo, j, i):
    """The diagonal's time, implicit-cooling and adiabatic part at midpoints."""
    cp, w = Fo["cp"][j, i], Fo["w_upd"][j, i]
    cpm = .5 * (cp[:-1] + cp[1:])
    wm = .5 * (w[:-1] + w[1:])
    ez = np.asarray(g["expz"])[:-1]
    return -ez * (cpm * (g["dtx2inv"] + Fo["cool_imp"][j, i][:-1]) + wm * GASK / Fo["mbar"][j, i][:-1])


def dt_coefficients():
    """Rows of a horizontally uniform state against the equation's interface
    coefficients, diagonal and lower and upper boundary rows."""
    g, Fo = base_state("small")
    Fo = uniform(g, Fo)
    cap, _ = _dt_rows(g, Fo)
    j, i = (g["nlat"] + 4) // 2, (g["nlon"] + 4) // 2
    p, q, r, f = (cap[n, j, i] for n in range(4))
    nk = g["nk"]
    G, F = _dt_physics(g, Fo, j, i)
    # interfaces 1..nk-2 from the rows: p_k = G - F, r_k-1 = G + F
    Gr = .5 * (p[1:nk - 1] + r[0:nk - 2])
    Fr = .5 * (r[0:nk - 2] - p[1:nk - 1])
    D = _dt_diag(g, Fo, j, i)
    k = np.arange(1, nk - 2)
    qint = -(G[k] + G[k + 1] + F[k] - F[k + 1]) + D[k]
    # lower boundary (Dirichlet T = T_lbc at interface 0, ghost T_-1 = 2 T_lbc - T_0)
    p0 = G[0] - F[0]
    q0 = -(G[0] + G[1] + F[0] - F[1]) + D[0] - p0
    # upper boundary: no flux through interface nk-1
    qtop = -(G[nk - 2] + F[nk - 2]) + D[nk - 2]
    return {
        "dt_G": rel(Gr, G[1:nk - 1]), "dt_F": rel(Fr, F[1:nk - 1]),
        "dt_diag": rel(q[1:nk - 2], qint),
        "dt_bottom_q": rel(q[0], q0, abs(q0)), "dt_bottom_p": abs(p[0]),
        "dt_top_q": rel(q[nk - 2], qtop, abs(qtop)), "dt_top_r": abs(r[nk - 2]),
    }


def dt_sources():
    """The right-hand side of a uniform state at rest: the smoothed n-1
    temperature's time term, the O recombination heating of dt.tex
    (Q_OR = f_T r_km N_A/m [p0 e^-z m/(k T) Psi_O/m_O]^2, with the state's
    p0 e^-z m/(k T)), the Joule heating and the explicit cooling, each
    isolated by changing it alone; the Dirichlet lower boundary's 2 p_0 T_lbc."""
    g, Fo = base_state("small")
    Fo = uniform(g, Fo)
    nk = g["nk"]
    j, i = (g["nlat"] + 4) // 2, (g["nlon"] + 4) // 2
    for n in ("w_upd", "qtotal", "hdt", "rkm12", "qji_tn", "cool_exp"):
        Fo[n] = np.zeros_like(Fo[n])
    # winds uniform with height and equal to their lower-boundary values: no
    # molecular-viscosity heating
    for n, lb in (("un", "ulbc"), ("vn", "vlbc")):
        set_levels(Fo, n, np.full(nk, Fo[n][j, i, 0]))
        Fo[lb] = np.full_like(Fo[lb], Fo[n][j, i, 0])
    cap0, _ = _dt_rows(g, Fo)
    f0 = cap0[3, j, i]
    cp = Fo["cp"][j, i]
    ez = np.asarray(g["expz"])
    exp_f = -ez[:-1] * .5 * (cp[:-1] + cp[1:]) * g["dtx2inv"] * Fo["tn_nm"][j, i][:-1]
    G, F = _dt_physics(g, Fo, j, i)
    exp_f[0] -= 2. * (G[0] - F[0]) * Fo["t_lbc"][j, i]
    errs = {"dt_rhs_time": rel(f0[:nk - 1], exp_f)}
    # one heating or cooling term at a time, each of the size of the time term
    d = np.linspace(1., 2., nk) * np.abs(f0).max()
    # Joule heating Q enters as -e^-Z Q, explicit cooling as +L_exp
    for n, val, expect in (("qji_tn", d / ez, -d), ("cool_exp", d, d)):
        F2 = dict(Fo)
        set_levels(F2, n, val)
        c2, _ = _dt_rows(g, F2)
        errs["dt_rhs_" + n] = rel((c2[3, j, i] - f0)[:nk - 1], expect[:nk - 1])
    F2 = dict(Fo)
    rk = 1e-33 * np.linspace(1., 2., nk)
    set_levels(F2, "rkm12", rk)
    c2, _ = _dt_rows(g, F2)
    xn = Fo["xnmbar"][j, i]
    qor = TSURPLUS * rk * AVO / Fo["mbar"][j, i] * (xn * Fo["o1"][j, i] / 16.) ** 2
    errs["dt_rhs_recomb"] = rel((c2[3, j, i] - f0)[:nk - 1], (-ez * qor)[:nk - 1])
    return errs


def dt_mirror():
    g, Fo = base_state("small")
    a = state(g, Fo).dt_ng(Fo["tlbc"])
    Fm, gm = mirror_lon(g, Fo)
    b = state(gm, Fm).dt_ng(Fm["tlbc"])
    return {"dt_mirror": max(rel(mirror_out(n, a[n]), b[n]) for n in ("tn", "tn_nm"))}


def dt_solution():
    """The solution solves the rows handed to the solve (rows 0..nk-2), the
    top level is the geometric extrapolation T(nk) = T(nk-1)^2/T(nk-2)
    (dt_ng.F90:167), and the n-1 output is the time filter
    (1-c)/2 (T_nm + T_new) + c T with c = 0.95 (cons.F dtsmooth)."""
    g, Fo = base_state("small")
    cap, res = _dt_rows(g, Fo)
    nk = g["nk"]
    p, q, r, f = cap[0], cap[1], cap[2], cap[3]
    # before the 100 K floor and the lateral data: interior columns only
    t = res["tn"][2:-2, 2:-2]
    pp, qq, rr, ff = (x[2:-2, 2:-2] for x in (p, q, r, f))
    lhs = qq[..., :nk - 1] * t[..., :nk - 1]
    lhs[..., 1:] += pp[..., 1:nk - 1] * t[..., :nk - 2]
    lhs[..., :-1] += rr[..., :nk - 2] * t[..., 1:nk - 1]
    top = t[..., nk - 2] ** 2 / t[..., nk - 3]
    c = 0.95
    tnm = (1 - c) / 2 * (Fo["tn_nm"] + res["tn"]) + c * Fo["tn"]
    return {"dt_residual": rel(lhs, ff[..., :nk - 1]), "dt_top_extrap": rel(t[..., nk - 1], top),
            "dt_time_filter": rel(res["tn_nm"], tnm)}


# ================================================================================
# duv_ng (doc/description/duv.tex): the zonal and meridional momentum equations
#   du/dt = g e^z/p0 d/dZ[mu/H du/dZ] + f v + l_xx (v_ExB,x - u) + l_xy (v_ExB,y - v)
#           - v.grad u + u v tan(phi)/R - 1/(R cos phi) dPhi/dlambda - W du/dZ + hd_u
#   dv/dt = ...       - f u + l_yy (v_ExB,y - v) - l_yx (v_ExB,x - u) ...
# taken implicitly and multiplied by e^-Z: the rows are 2x2 blocks with
#   P_k = -f_vis(k)/dz^2 I - e^-Z W/(2 dz) I,  R_k = -f_vis(k+1)/dz^2 I + e^-Z W/(2 dz) I,
#   Q_k = (f_vis(k)+f_vis(k+1))/dz^2 I + e^-Z [[1/(2dt) + l_xx, -(f + u tan/R - l_xy)],
#                                               [f + u tan/R - l_yx, 1/(2dt) + l_yy]]
#   (the description notes that the code's metric term uses u in both rows),
#   f_vis = g mu/(p0 H) at interfaces; ion-drag coefficients and drifts are
#   interface values averaged to the midpoints.  Right-hand side:
#   e^-Z (u_nm,smoothed/(2dt) + hd_u + F_e + l_xx v_ExB,x + l_xy v_ExB,y
#         - advection - g/(R cos phi) dz*/dlambda), and the meridional one alike.
def _duv_rows(g, Fo, tn_upd=None):
    st = state(g, Fo)
    nk = g["nk"]
    tn_upd = Fo["tn"] if tn_upd is None else tn_upd
    Fe = Fo.get("Fe_in", np.zeros_like(Fo["tn"]))
    Fn = Fo.get("Fn_in", np.zeros_like(Fo["tn"]))
    cap, res = capture(g, lambda: st.duv_ng(tn_upd, Fe, Fn, Fo["tlbc"], Fo["ulbc"], Fo["vlbc"]))
    return cap[24:38], res


def _mid(x):
    return .5 * (x[:-1] + x[1:])


def _duv_uniform(w_zero=True):
    g, Fo = base_state("small")
    j, i = (g["nlat"] + 4) // 4, (g["nlon"] + 4) // 2      # away from the equator
    Fo = uniform(g, Fo, at=(j, i))
    g = grid_with(g, xmue=[1e-40] * g["nk"])               # molecular viscosity only
    if w_zero:
        Fo["w_upd"] = np.zeros_like(Fo["w_upd"])
    # an ion-drag tensor with four different components
    Fo["lyy"] = np.ascontiguousarray(1.3 * Fo["lxx"])
    Fo["lyx"] = np.ascontiguousarray(0.6 * Fo["lxy"])
    nk = g["nk"]
    Fo["Fe_in"] = np.ascontiguousarray(np.broadcast_to(np.linspace(.1, .3, nk), Fo["tn"].shape))
    Fo["Fn_in"] = np.ascontiguousarray(np.broadcast_to(np.linspace(-.2, .2, nk), Fo["tn"].shape))
    return g, Fo, j, i


def duv_coefficients():
    g, Fo, j, i = _duv_uniform()
    cap, _ = _duv_rows(g, Fo)
    nk, dz = g["nk"], g["dz"]
    A, B, Cb = cap[0:4, j, i], cap[4:8, j, i], cap[8:12, j, i]      # row-major 2x2: 11 12 21 22
    fvis = GRAV * Fo["km"][j, i] / (P0 * Fo["schti"][j, i]) / dz ** 2  # f_vis/dz^2 at interfaces
    ez = np.asarray(g["expz"])
    lxx, lyy, lxy, lyx = (_mid(Fo[n][j, i]) for n in ("lxx", "lyy", "lxy", "lyx"))
    f, tanr = g["cor"][j], g["tanphi"][j] / RE
    u = Fo["un"][j, i]
    k = np.arange(1, nk - 2)
    errs = {
        "duv_P": rel(np.stack([A[0, k], A[3, k]]), -np.stack([fvis[k], fvis[k]])),
        "duv_R": rel(np.stack([Cb[0, k - 1], Cb[3, k - 1]]), -np.stack([fvis[k], fvis[k]])),
        "duv_PR_offdiag": float(np.abs(np.stack([A[1, k], A[2, k], Cb[1, k], Cb[2, k]])).max()),
    }
    q11 = fvis[k] + fvis[k + 1] + ez[k] * (g["dtx2inv"] + lxx[k])
    q22 = fvis[k] + fvis[k + 1] + ez[k] * (g["dtx2inv"] + lyy[k])
    q12 = -ez[k] * (f + u[k] * tanr - lxy[k])
    q21 = ez[k] * (f + u[k] * tanr - lyx[k])
    errs["duv_Q"] = rel(np.stack([B[0, k], B[3, k]]), np.stack([q11, q22]))
    errs["duv_Q_coriolis"] = rel(np.stack([B[1, k], B[2, k]]), np.stack([q12, q21]))
    # lower boundary: u = u_lbc, v = v_lbc at interface 0 (ghost 2 u_lbc - u_0)
    p0 = -fvis[0]
    errs["duv_bottom_Q"] = rel(np.array([B[0, 0], B[3, 0]]),
                               np.array([fvis[0] + fvis[1] + ez[0] * (g["dtx2inv"] + lxx[0]) - p0,
                                         fvis[0] + fvis[1] + ez[0] * (g["dtx2inv"] + lyy[0]) - p0]))
    errs["duv_bottom_P"] = float(np.abs(A[:, 0]).max())
    # upper boundary: zero gradient (u(nk) = u(nk-1)) folded into the diagonal
    kt = nk - 2
    errs["duv_top_Q"] = rel(np.array([B[0, kt], B[3, kt]]),
                            np.array([fvis[kt] + ez[kt] * (g["dtx2inv"] + lxx[kt]),
                                      fvis[kt] + ez[kt] * (g["dtx2inv"] + lyy[kt])]))
    errs["duv_top_R"] = float(np.abs(Cb[:, kt]).max())
    return errs


def duv_advection_and_sources():
    """Vertical advection W du/dZ centred between the neighbours; the
    right-hand side of a uniform state (smoothed n-1 wind, horizontal
    diffusion, O+ momentum forcing, ion drift) and the Dirichlet lower
    boundary; the pressure gradient of a geopotential height linear in
    longitude and latitude."""
    g, Fo, j, i = _duv_uniform()
    nk, dz = g["nk"], g["dz"]
    cap0, _ = _duv_rows(g, Fo)
    ez = np.asarray(g["expz"])
    F2 = dict(Fo)
    wprof = np.linspace(1e-5, -2e-5, nk)
    set_levels(F2, "w_upd", wprof)
    cap1, _ = _duv_rows(g, F2)
    wm = _mid(wprof)
    k = np.arange(1, nk - 2)
    adv = ez[k] * wm[k] / (2 * dz)
    errs = {"duv_w_P": rel(cap1[0, j, i, k] - cap0[0, j, i, k], -adv),
            "duv_w_R": rel(cap1[8 + 3, j, i, k] - cap0[8 + 3, j, i, k], adv)}
    # right-hand side of the uniform state
    S1, S2 = cap0[12, j, i], cap0[13, j, i]
    lxx, lyy, lxy, lyx, ui, vi = (_mid(Fo[n][j, i]) for n in ("lxx", "lyy", "lxy", "lyx", "ui", "vi"))
    s1 = ez[:-1] * (g["dtx2inv"] * Fo["un_nm"][j, i][:-1] + Fo["hdu"][j, i][:-1] + Fo["Fe_in"][j, i][:-1]
                    + lxx * ui + lxy * vi)
    s2 = ez[:-1] * (g["dtx2inv"] * Fo["vn_nm"][j, i][:-1] + Fo["hdv"][j, i][:-1] + Fo["Fn_in"][j, i][:-1]
                    + lyy * vi - lyx * ui)
    fvis0 = GRAV * Fo["km"][j, i][0] / (P0 * Fo["schti"][j, i][0]) / dz ** 2
    s1[0] -= 2 * (-fvis0) * Fo["u_lbc"][j, i]
    s2[0] -= 2 * (-fvis0) * Fo["v_lbc"][j, i]
    errs["duv_rhs"] = max(rel(S1[1:nk - 1], s1[1:]), rel(S2[1:nk - 1], s2[1:]))
    errs["duv_rhs_bottom"] = max(rel(S1[0], s1[0], abs(s1[0])), rel(S2[0], s2[0], abs(s2[0])))
    # pressure gradient: z = z0 + ax*i + ay*j (cm per grid step)
    ax, ay = 3.0e3, -2.0e3
    F3 = dict(Fo)
    nlatd, nlond = g["nlat"] + 4, g["nlon"] + 4
    J, I = np.meshgrid(np.arange(nlatd), np.arange(nlond), indexing="ij")
    F3["z"] = np.ascontiguousarray(Fo["z"] + (ax * I + ay * J)[..., None])
    cap3, _ = _duv_rows(g, F3)
    d1 = cap3[12, j, i] - S1
    d2 = cap3[13, j, i] - S2
    e1 = -ez[:-1] * GRAV * (ax / g["dlamda"]) * g["racs"][j]
    e2 = -ez[:-1] * GRAV * (ay / g["dphi"]) / RE
    errs["duv_pressure_grad"] = max(rel(d1[1:nk - 1], e1[1:]), rel(d2[1:nk - 1], e2[1:]))
    return errs


def duv_mirror():
    g, Fo = base_state("small")
    nk = g["nk"]
    tn_upd = Fo["tn"] * 1.001
    z = np.zeros_like(Fo["tn"])
    st = state(g, Fo)
    a = st.duv_ng(tn_upd, Fo["lxx"] * 1e3, Fo["lyy"] * 1e3, Fo["tlbc"], Fo["ulbc"], Fo["vlbc"])
    Fm, gm = mirror_lon(g, Fo)
    stm = state(gm, Fm)
    # Fe is odd (zonal), Fn even
    b = stm.duv_ng(np.ascontiguousarray(tn_upd[:, ::-1]), -np.ascontiguousarray((Fo["lxx"] * 1e3)[:, ::-1]),
                   np.ascontiguousarray((Fo["lyy"] * 1e3)[:, ::-1]), np.ascontiguousarray(Fo["tlbc"][:, ::-1]),
                   Fm["ulbc"], Fm["vlbc"])
    return {"duv_mirror": max(rel(mirror_out(n, a[n]), b[n]) for n in ("un", "vn", "un_nm", "vn_nm"))}


# ================================================================================
# minor_ng + comp_{no,n4s,n2d,ar}_ng (doc/description/minor.tex, eq. minor_goveq)
#   dPsi/dt = -e^z d/dZ[A (d/dZ - E) Psi] + S Psi - R - (v.grad Psi + W dPsi/dZ)
#             + e^z d/dZ[e^-z K_E (d/dZ + 1/m dm/dZ) Psi] + H_sub
# multiplied by e^-Z.  With the molecular flux at interface k discretised as
#   A_k [(Psi_k - Psi_k-1)/dz - E_k (Psi_k + Psi_k-1)/2],
# row k has p = A_k (1/dz + E_k/2)/dz, r_k-1 = A_k (1/dz - E_k/2)/dz, and the
# diagonal closes the flux form.  The description gives
#   A = -m(z)/(tau m_N2) (T0/T)^0.25 / (sum_n phi_xn Psi_n)   (Blanc's law)
#   E = 1 - (m_x + dm/dZ)/m - alpha_x (dT/dZ)/T + (friction with the majors)
# and the friction term vanishes when the major species are in diffusive
# equilibrium, dPsi_n/dZ = (1 - (m_n + dm/dZ)/m) Psi_n.
MINOR = {  # species index: (name, m_x, phi_x (O2, O, He, N2), thermal diffusion alpha_x)
    0: ("no", 30., (0.814, 0.866, 0.662, 0.926), 0.),
    1: ("n4s", 14., (0.651, 0.731, 0.627, 0.741), 0.),
    2: ("n2d", 14., (0.651, 0.731, 0.627, 0.741), 0.),
    3: ("ar", 40., (1.042, 1.509, 0.732, 1.176), 0.17),
}
TAU, T00 = 1.86e3, 273.


def _minor_rows(g, Fo, sp, n2d=None):
    st = state(g, Fo)
    n2d = np.zeros_like(Fo["tn"]) if n2d is None else n2d
    cap, res = capture(g, lambda: st.minor(sp, n2d, Fo["tlbc"]))
    return cap[8 + 4 * sp: 12 + 4 * sp], res


def _equilibrium_majors(g, Fo):
    """O2, O, He in discrete diffusive equilibrium over a smooth mean mass."""
    nk, dz = g["nk"], g["dz"]
    z = np.asarray(g["zpmid"])
    mbar = 26. - 4. * (1. / (1. + np.exp(-(z - 1.) / 2.)))
    barm = 26.2 - 4.1 * (1. / (1. + np.exp(-(z - 1.2) / 2.)))
    set_levels(Fo, "mbar", mbar)
    set_levels(Fo, "barm", barm)
    dmdz = np.zeros(nk)
    dmdz[1:] = (mbar[1:] - mbar[:-1]) / dz
    for n, m, psi0 in (("o2", 32., 0.2), ("o1", 16., 1e-3), ("he", 4., 1e-8)):
        psi = np.empty(nk)
        psi[0] = psi0
        for k in range(1, nk):
            c = 1. - (m + dmdz[k]) / barm[k]
            psi[k] = psi[k - 1] * (1. + c * dz / 2.) / (1. - c * dz / 2.)
        set_levels(Fo, n, psi)
    return dmdz


def _minor_state(w_zero=True, difk_zero=True):
    g, Fo = base_state("small", difhor=0)
    j, i = (g["nlat"] + 4) // 2, (g["nlon"] + 4) // 2
    Fo = uniform(g, Fo, at=(j, i))
    if difk_zero:
        g = grid_with(g, difk=[0.] * g["nk"])
    if w_zero:
        Fo["w_upd"] = np.zeros_like(Fo["w_upd"])
    dmdz = _equilibrium_majors(g, Fo)
    return g, Fo, j, i, dmdz


def _minor_AE(g, Fo, j, i, dmdz, sp, nest_quirk=True):
    """A_k and E_k of the description at interfaces k = 0..nk-1 (k >= 1 valid)."""
    _, mx, phi, alfx = MINOR[sp]
    nk, dz = g["nk"], g["dz"]
    tn = Fo["tn"][j, i]
    T = np.empty(nk)
    T[0] = Fo["tlbc"][j, i]
    T[-1] = tn[-2]
    T[1:-1] = .5 * (tn[1:-1] + tn[:-2])
    psi = {}
    for n in ("o2", "o1", "he"):
        f = Fo[n][j, i]
        psi[n] = np.concatenate([[np.nan], .5 * (f[1:] + f[:-1])])
    n2 = 1. - psi["o2"] - psi["o1"] - psi["he"]
    blanc = phi[0] * psi["o2"] + phi[1] * psi["o1"] + phi[2] * psi["he"] + phi[3] * n2
    # minor_ng.F90:136 puts phi_x(He) where the description's phi_x(N2) belongs
    # (the global minor.F:359-362 has phi_x(N2)); the nest form is reproduced
    if nest_quirk:
        blanc = blanc + (phi[2] - phi[3])
    barm = Fo["barm"][j, i]
    A = -barm / (TAU * 28.) * (T00 / T) ** 0.25 / blanc
    dTdz = np.empty(nk)
    dTdz[1:-1] = (T[2:] - T[:-2]) / (2. * dz)
    dTdz[0] = (T[1] - T[0]) / dz
    dTdz[-1] = (T[-1] - T[-2]) / dz
    E = 1. - (mx + dmdz) / barm - alfx * dTdz / T
    return A, E


def minor_molecular():
    errs = {}
    for sp in (0, 3):   # NO (no thermal diffusion) and Ar (alpha_x = 0.17)
        g, Fo, j, i, dmdz = _minor_state()
        cap, _ = _minor_rows(g, Fo, sp)
        nk, dz = g["nk"], g["dz"]
        p, q, r = cap[0, j, i], cap[1, j, i], cap[2, j, i]
        A, E = _minor_AE(g, Fo, j, i, dmdz, sp)
        k = np.arange(1, nk - 1)
        Ar = .5 * (p[k] + r[k - 1]) * dz ** 2
        Er = (p[k] - r[k - 1]) * dz / Ar
        nm = MINOR[sp][0]
        errs["minor_A_" + nm] = rel(Ar, A[k])
        errs["minor_E_" + nm] = rel(Er, E[k])
        # zero flux through the top interface: the top row is
        # -dz [(Psi_top - Psi_top-1)/dz - E (Psi_top + Psi_top-1)/2] = 0
        errs["minor_top_" + nm] = max(rel(p[nk - 1], 1. + .5 * dz * E[nk - 1], 1.),
                                      rel(q[nk - 1], -1. + .5 * dz * E[nk - 1], 1.), abs(r[nk - 1]))
        if sp == 3:
            # flux form (Ar has no chemistry): each unknown's coefficients sum
            # to zero over its rows once the time term e^-Z/(2 dt) is taken
            # from the diagonal
            ez = np.asarray(g["expz"])
            qt = q - ez * g["dtx2inv"]
            k = np.arange(1, nk - 2)
            errs["minor_flux_form"] = rel(r[k - 1] + qt[k] + p[k + 1], 0 * k, np.abs(qt[k]).max())
    return errs


def minor_eddy_and_advection():
    """Eddy diffusion (the difference of rows with and without K_E) and
    vertical advection (with and without W) against the description."""
    errs = {}
    sp = 3
    g, Fo, j, i, dmdz = _minor_state()
    nk, dz = g["nk"], g["dz"]
    c0, _ = _minor_rows(g, Fo, sp)
    ke = np.exp(-np.asarray(g["zpint"])) * 2e-6
    g2 = grid_with(g, difk=list(ke))
    c1, _ = _minor_rows(g2, Fo, sp)
    barm = Fo["barm"][j, i]
    dm = dmdz / barm
    ezi = np.exp(-np.asarray(g["zpint"]))            # e^-Z at interfaces
    k = np.arange(1, nk - 1)
    # eddy flux at interface k: e^-Z K_E [(Psi_k - Psi_k-1)/dz + dm_k (Psi_k + Psi_k-1)/2]
    pe = -ezi[k] * ke[k] * (1. / dz - .5 * dm[k]) / dz
    re = -ezi[k] * ke[k] * (1. / dz + .5 * dm[k]) / dz
    errs["minor_eddy_P"] = rel(c1[0, j, i, k] - c0[0, j, i, k], pe)
    errs["minor_eddy_R"] = rel(c1[2, j, i, k - 1] - c0[2, j, i, k - 1], re)
    kq = np.arange(1, nk - 2)
    qe = (ezi[kq] * ke[kq] * (1. / dz + .5 * dm[kq]) + ezi[kq + 1] * ke[kq + 1] * (1. / dz - .5 * dm[kq + 1])) / dz
    errs["minor_eddy_Q"] = rel(c1[1, j, i, kq] - c0[1, j, i, kq], qe)
    F2 = dict(Fo)
    wprof = np.linspace(1e-5, -2e-5, nk)
    set_levels(F2, "w_upd", wprof)
    c2, _ = _minor_rows(g, F2, sp)
    ez = np.asarray(g["expz"])
    kk = np.arange(1, nk - 1)
    adv = ez[kk] * _mid(wprof)[kk] / (2 * dz)
    errs["minor_w_P"] = rel(c2[0, j, i, kk] - c0[0, j, i, kk], -adv)
    errs["minor_w_R"] = rel(c2[2, j, i, kk] - c0[2, j, i, kk], adv)
    return errs


def _ndens(Fo, name, m, j, i):
    """Number density of a species from its mass mixing ratio."""
    return Fo["xnmbar"][j, i] * Fo[name][j, i] / m


def minor_chemistry():
    """Each species' production and loss from the rate law of one reaction,
    changing its rate coefficient alone: production P (cm^-3 s^-1) enters the
    right-hand side as e^-Z P m_x/(p0 e^-Z m/(k T)) and a loss frequency L as
    -e^-Z L on the diagonal (minor.tex: s0prod, -e^-z S)."""
    errs = {}
    g, Fo = base_state("small")
    j, i = (g["nlat"] + 4) // 2, (g["nlon"] + 4) // 2
    Fo = uniform(g, Fo, at=(j, i))
    nk = g["nk"]
    ez = np.asarray(g["expz"])
    xn = Fo["xnmbar"][j, i]
    n2d_new = Fo["n2d"] * 1.3
    nn = lambda nm, m: _ndens(Fo, nm, m, j, i)
    # (species, rate field, change, expected change of production or loss frequency)
    cases = [
        # NO: N(4S) + O2 -> NO + O (beta1) produces NO; N(4S) + NO -> N2 + O (beta3) destroys it
        (0, "beta1", "prod", 1e-11, lambda d: d * nn("n4s", 14.) * nn("o2", 32.)),
        (0, "beta3", "loss", 1e-11, lambda d: -d * nn("n4s", 14.)),
        # N(4S): N(4S) + O2 (beta1) destroys it
        (1, "beta1", "loss", 1e-11, lambda d: -d * nn("o2", 32.)),
        # N(2D): N(2D) + e (beta5) destroys it (electron density at the midpoint
        # as the geometric mean of the neighbouring levels)
        (2, "beta5", "loss", 1e-7,
         lambda d: -d * np.sqrt(Fo["ne"][j, i] * np.concatenate([Fo["ne"][j, i][1:], [np.nan]]))),
    ]
    for sp, field, kind, d, expect in cases:
        c0, _ = _minor_rows(g, Fo, sp, n2d_new)
        F2 = dict(Fo)
        F2[field] = np.ascontiguousarray(Fo[field] + d)
        c1, _ = _minor_rows(g, F2, sp, n2d_new)
        nm = MINOR[sp][0]
        kk = np.arange(0, nk - 2) if field != "beta5" else np.arange(0, nk - 1)
        kk = kk[1:]
        if kind == "prod":
            got = (c1[3, j, i] - c0[3, j, i])[kk]
            want = ez[kk] * expect(d)[kk] * MINOR[sp][1] / xn[kk]
        else:
            got = (c1[1, j, i] - c0[1, j, i])[kk]
            want = -ez[kk] * expect(d)[kk]
        errs["minor_chem_%s_%s" % (nm, field)] = rel(got, want)
    # Ar has no chemistry
    c0, _ = _minor_rows(g, Fo, 3)
    F2 = dict(Fo)
    for fld in ("beta1", "beta3", "beta5", "rk3", "ne"):
        F2[fld] = np.ascontiguousarray(Fo[fld] * 2.)
    c1, _ = _minor_rows(g, F2, 3)
    errs["minor_chem_ar_none"] = float(np.abs(c1[:, j, i, :nk - 1] - c0[:, j, i, :nk - 1]).max())
    return errs


def minor_lower_boundary():
    """Ar's Dirichlet lower boundary Psi = sqrt(glbm1 glbm2) at interface 0
    (ghost 2 Psi_b - Psi_0): the right-hand side of row 0 changes by
    -2 p_0 dPsi_b, with p_0 within the smooth variation of the next
    interfaces' coefficients (p_0 against 2 p_1 - p_2)."""
    g, Fo, j, i, dmdz = _minor_state()
    c0, _ = _minor_rows(g, Fo, 3)
    g2 = grid_with(g, ar_glbm=[g["ar_glbm"][0] * 1.21, g["ar_glbm"][1]])
    c1, _ = _minor_rows(g2, Fo, 3)
    dpsi = math.sqrt(g2["ar_glbm"][0] * g2["ar_glbm"][1]) - math.sqrt(g["ar_glbm"][0] * g["ar_glbm"][1])
    p0 = -(c1[3, j, i, 0] - c0[3, j, i, 0]) / (2. * dpsi)
    p_lin = 2. * c0[0, j, i, 1] - c0[0, j, i, 2]
    return {"minor_bottom_dirichlet": abs(p0 / p_lin - 1.)}


def minor_mirror():
    g, Fo = base_state("small")
    st = state(g, Fo)
    a = st.minors(Fo["tlbc"])
    Fm, gm = mirror_lon(g, Fo)
    b = state(gm, Fm).minors(Fm["tlbc"])
    return {"minor_mirror": max(rel(mirror_out(n, a[n]), b[n]) for n in a)}


# ================================================================================
# oplus_ng (doc/description/oplus.tex, eq. oplus_optrans), one sub-step of
# length dt/nstep_sub.  With the vertical ambipolar flux at interface k
#   Phi_k = K b_z^2 [ ((2 Tp n)_k - (2 Tp n)_k-1) / (H_i dz) + (m g/R) (n_k + n_k-1)/2 ]
# (K = D_A at the interface, m = 16 for O+, Tp = (Te + Ti)/2), the operator
# (b_z/H d/dZ + div b) Phi is taken as
#   (1/(H dz) + div b/(2 b_z)) Phi_k+1 - (1/(H dz) - div b/(2 b_z)) Phi_k
# on the side of the equation where the time term is -n/(2 dt_sub); the loss
# L n and production Q enter as -L on the diagonal and -Q on the right; the
# neutral wind as -(b_z/H d/dZ + div b)(b.v n); the vertical ExB drift as
# -v_z/H dn/dZ - 6 v_z n / R_E (B^2 of a dipole falls as r^-6).  For a
# horizontally uniform b, div b = (1/(R cos phi)) d(cos phi b_y)/dphi + 2 b_z/R
# (oplus.tex, subroutine divb), the latitude derivative taken with the
# model's horizontal derivative (dffm_ng df, pinned by the stencil tests).
RMOP = 16.


def _divb_uniform(g, Fo, j, i):
    o = oracle()
    nlatd, nlond = g["nlat"] + 4, g["nlon"] + 4
    csby = np.ascontiguousarray(np.asarray(g["cs"])[:, None] * Fo["by"])
    d = np.zeros((nlatd, nlond))
    o.lib().tgx_o_df2d(o._p(csby), o._p(d), -1, nlond, nlatd)
    return d[j, i] / g["dphi"] / g["cs"][j] / RE + 2. * Fo["bz"][j, i] / RE


def _op_rows(g, Fo):
    st = state(g, Fo)
    cap, res = capture(g, lambda: st.oplus_ng(Fo["op"], Fo["op_nm"], Fo["lonb_op"][0], Fo["latb_op"][0]))
    return cap[4:8], res


def _op_state():
    g, Fo = base_state("small", nstep_sub=2)
    j, i = (g["nlat"] + 4) // 4, (g["nlon"] + 4) // 2
    Fo = uniform(g, Fo, at=(j, i))
    for n in ("un", "vn", "w", "ui", "vi", "wi"):
        Fo[n] = np.zeros_like(Fo[n])
    return g, Fo, j, i


def _op_DA(Fo, j, i):
    """D_A at the midpoints (oplus.tex, subroutine rrk), with the O-O+
    resonant charge-exchange temperature (Tn + Ti)/2 and the He term."""
    tr = .5 * (Fo["tn"][j, i] + Fo["ti"][j, i])
    den = (18.1 * Fo["o2"][j, i] / 32. + Fo["o1"][j, i] / 16. * np.sqrt(tr) * (1 - .064 * np.log10(tr)) ** 2 * 1.5
           + 3.6 * Fo["he"][j, i] / 4. + 18.6 * Fo["n2"][j, i] / 28.)
    return 1.42e17 / (Fo["xnmbar"][j, i] * den)


def _op_expected(g, Fo, j, i):
    """(p, r) of the ambipolar diffusion at rows 0..nk-2 from the equation."""
    nk, dz = g["nk"], g["dz"]
    bz = Fo["bz"][j, i]
    H = Fo["scht"][j, i]
    Tp = .5 * (Fo["ti"][j, i] + Fo["te"][j, i])
    hdz = 1. / (H * dz)
    hdzi = np.empty(nk)
    hdzi[1:-1] = .5 * (hdz[:-2] + hdz[1:-1])
    K = np.empty(nk)
    DA = _op_DA(Fo, j, i)
    K[1:-1] = .5 * (DA[:-2] + DA[1:-1])
    gm = GRAV * RMOP / (2 * GASK)
    divb = _divb_uniform(g, Fo, j, i)
    k = np.arange(1, nk - 2)
    # coefficient of n_k-1 in row k, and of n_k+1 in row k
    p = (hdz[k] - .5 * divb / bz) * bz ** 2 * K[k] * (2 * Tp[k - 1] * hdzi[k] - gm)
    r = (hdz[k] + .5 * divb / bz) * bz ** 2 * K[k + 1] * (2 * Tp[k + 1] * hdzi[k + 1] + gm)
    return k, p, r, K, hdz, hdzi, Tp, gm, divb


def oplus_diffusion():
    g, Fo, j, i = _op_state()
    cap, _ = _op_rows(g, Fo)
    k, p, r, K, hdz, hdzi, Tp, gm, divb = _op_expected(g, Fo, j, i)
    nk = g["nk"]
    bz = Fo["bz"][j, i]
    errs = {"oplus_P": rel(cap[0, j, i, k], p), "oplus_R": rel(cap[2, j, i, k], r)}
    # the top row: the flux through the top interface is the specified
    # plasmaspheric flux, so the diagonal keeps only the lower interface's
    # n_top term; the loss and time terms as in the interior
    kt = nk - 2
    L = Fo["xnmbar"][j, i] * (Fo["rk1"][j, i] * Fo["o2"][j, i] / 32. + Fo["rk2"][j, i] * Fo["n2"][j, i] / 28.
                              + 1.3e-10 * Fo["n2d"][j, i] / 14.)
    qtop = (-(hdz[kt] - .5 * divb / bz) * bz ** 2 * K[kt] * (2 * Tp[kt] * hdzi[kt] + gm)
            - g["dtx2inv"] * g["nstep_sub"] - L[kt])
    errs["oplus_top_q"] = rel(cap[1, j, i, kt], qtop, abs(qtop))
    errs["oplus_top_r"] = abs(cap[2, j, i, kt])
    return errs


def oplus_chemistry():
    """With the diffusion coefficient capped to nothing (opdiffcap) and no
    winds or drifts, the rows are the photochemical balance
    -(1/(2 dt_sub) + L) n = -(n_nm/(2 dt_sub) + Q): L from O+ + O2, N2, N(2D)
    (rk1, rk2, 1.3e-10), and Q's photoionisation part from qop."""
    g, Fo, j, i = _op_state()
    g = grid_with(g, opdiffcap=1e-200)
    nk = g["nk"]
    cap, _ = _op_rows(g, Fo)
    L = Fo["xnmbar"][j, i] * (Fo["rk1"][j, i] * Fo["o2"][j, i] / 32. + Fo["rk2"][j, i] * Fo["n2"][j, i] / 28.
                              + 1.3e-10 * Fo["n2d"][j, i] / 14.)
    k = np.arange(1, nk - 2)
    dt2 = g["dtx2inv"] * g["nstep_sub"]
    errs = {"oplus_chem_q": rel(cap[1, j, i, k], -(dt2 + L[k])),
            "oplus_chem_pr": float(np.abs(cap[0, j, i, k]).max() + np.abs(cap[2, j, i, k]).max())}
    F2 = dict(Fo)
    dq = np.linspace(10., 30., nk)
    set_levels(F2, "qop", Fo["qop"][j, i] + dq)
    c2, _ = _op_rows(g, F2)
    errs["oplus_chem_prod"] = rel((c2[3, j, i] - cap[3, j, i])[k], -_mid(dq)[k])
    return errs


def oplus_transport_terms():
    """Vertical ExB drift and neutral wind, each switched on alone (the row
    differences), and the smoothing of n at n-1 (Shapiro constant over
    nstep_sub, checked through the stencil-pinned smoother)."""
    g, Fo, j, i = _op_state()
    nk, dz = g["nk"], g["dz"]
    cap0, _ = _op_rows(g, Fo)
    H = Fo["scht"][j, i]
    hdz = 1. / (H * dz)
    k = np.arange(1, nk - 2)
    errs = {}
    F2 = dict(Fo)
    wprof = np.linspace(1e3, -3e3, nk)
    set_levels(F2, "wi", wprof)
    c2, _ = _op_rows(g, F2)
    vz = _mid(wprof)
    errs["oplus_drift"] = max(rel((c2[0] - cap0[0])[j, i, k], vz[k] * .5 * hdz[k]),
                              rel((c2[2] - cap0[2])[j, i, k], -vz[k] * .5 * hdz[k]),
                              rel((c2[1] - cap0[1])[j, i, k], -vz[k] * 6. / RE))
    # neutral wind along b: b.v = bx u (+ by v + bz W H)
    F3 = dict(Fo)
    uprof = np.linspace(3e4, -5e4, nk)
    set_levels(F3, "un", uprof)
    c3, _ = _op_rows(g, F3)
    bz, bx = Fo["bz"][j, i], Fo["bx"][j, i]
    bu = bx * uprof
    divb = _divb_uniform(g, Fo, j, i)
    errs["oplus_wind"] = max(rel((c3[0] - cap0[0])[j, i, k], bz * bu[k - 1] * .5 * hdz[k]),
                             rel((c3[2] - cap0[2])[j, i, k], -bz * bu[k + 1] * .5 * hdz[k]),
                             rel((c3[1] - cap0[1])[j, i, k], -bu[k] * divb))
    # n at n-1: -smooth(n_nm; shapiro/nsub) / (2 dt_sub) on the right
    F4 = dict(Fo)
    nlatd, nlond = g["nlat"] + 4, g["nlon"] + 4
    J, I = np.meshgrid(np.arange(nlatd), np.arange(nlond), indexing="ij")
    bump = (np.cos(2.1 * I) * np.sin(1.3 * J + .2))[..., None] * np.linspace(1., 2., nk) * 1e3
    F4["op_nm"] = np.ascontiguousarray(Fo["op_nm"] + bump)
    c4, _ = _op_rows(g, F4)
    sm = state(g, Fo).smooth(np.ascontiguousarray(np.broadcast_to(bump, Fo["op"].shape)),
                             g["shapiro"] / g["nstep_sub"])
    dt2 = g["dtx2inv"] * g["nstep_sub"]
    errs["oplus_smooth_nm"] = rel((c4[3] - cap0[3])[2:-2, 2:-2, 1:nk - 2], -(sm * dt2)[2:-2, 2:-2, 1:nk - 2])
    return errs


def oplus_affine_and_forcing():
    """The O+ matrix does not depend on O+, and the right-hand side is affine
    in (n, n at n-1); the O+ momentum forcing of a uniform n with uniform
    plasma temperature is the gravity term alone:
    Fe = (16 * 3.53e-11 * 1.42e17 / N m) (m g/R) n sinI cosI sinD."""
    g, Fo = base_state("small", nstep_sub=2)
    nk = g["nk"]
    ca, _ = _op_rows(g, Fo)
    F2 = dict(Fo)
    F2["op"] = np.ascontiguousarray(Fo["op"] * 1.7 + 50.)
    F2["op_nm"] = np.ascontiguousarray(Fo["op_nm"] * .6 + 20.)
    cb, _ = _op_rows(g, F2)
    F3 = dict(Fo)
    F3["op"] = np.ascontiguousarray(.25 * Fo["op"] + .75 * F2["op"])
    F3["op_nm"] = np.ascontiguousarray(.25 * Fo["op_nm"] + .75 * F2["op_nm"])
    cc, _ = _op_rows(g, F3)
    errs = {"oplus_matrix_independent": float(max(np.abs(ca[n] - cb[n]).max() for n in range(3))),
            "oplus_rhs_affine": rel(cc[3], .25 * ca[3] + .75 * cb[3], np.abs(cc[3]).max())}
    gu, Fu, j, i = _op_state()
    set_levels(Fu, "op", np.full(nk, 1e5))
    set_levels(Fu, "te", Fu["te"][j, i, 5] * np.ones(nk))
    set_levels(Fu, "ti", Fu["ti"][j, i, 5] * np.ones(nk))
    res = state(gu, Fu).oplus_ng(Fu["op"], Fu["op_nm"], Fu["lonb_op"][0], Fu["latb_op"][0])
    dip = Fu["dipmag"][j, i]
    fe = (16 * 3.53e-11 * 1.42e17 / Fu["xnmbar"][j, i]) * (RMOP * GRAV / GASK) * 1e5 * np.sin(dip) * np.cos(dip) * \
        Fu["sndec"][j, i]
    k = np.arange(1, nk - 2)
    errs["oplus_Fe_gravity"] = rel(res["Fe"][j, i, k], fe[k])
    return errs


def oplus_mirror():
    g, Fo = base_state("small", nstep_sub=2)
    st = state(g, Fo)
    a = st.oplus_subcycled(Fo["lonb_op"], Fo["latb_op"])
    Fm, gm = mirror_lon(g, Fo)
    b = state(gm, Fm).oplus_subcycled(Fm["lonb_op"], Fm["latb_op"])
    return {"oplus_mirror": max(rel(mirror_out(n, a[n]), b[n]) for n in a)}


# ================================================================================
# comp_o2o_ng: fs(isp, m) are the chemical tendencies of the mass mixing ratios
# of O2, O, He: d Psi_isp/dt = fs(isp,0) + fs(isp,1) Psi_O2 + fs(isp,2) Psi_O
# (+ fs(isp,3) Psi_He).  O2 + hv -> 2 O (rate rj) and O + O + M -> O2 + M
# (rkm12) move mass between O2 and O only, so with every other reactant absent
# the mass of O2 + O is conserved term by term; He takes part in no reaction.
_NO_CHEM = ("no", "n4s", "n2d", "o2p", "op", "nplus", "n2p", "nop", "qo2p", "qop", "rj", "rkm12")


def _fs(g, Fo):
    return state(g, Fo).comp_o2o_ng()     # (4, 3, nlatd, nlond, nk): [m, isp]


def o2o_conservation():
    g, Fo = base_state("small")
    errs = {}
    for keep in ("rj", "rkm12"):
        F2 = dict(Fo)
        for n in _NO_CHEM:
            if n != keep:
                F2[n] = np.zeros_like(Fo[n])
        fs = _fs(g, F2)
        scale = np.abs(fs[1:3, 0:2]).max()
        errs["o2o_mass_" + keep] = max(rel(fs[m, 0] + fs[m, 1], 0 * fs[m, 0], scale) for m in range(3))
        errs["o2o_active_" + keep] = float(scale == 0)
    # photodissociation removes O2 at the rate J (midpoint average of rj)
    F2 = dict(Fo)
    for n in _NO_CHEM:
        if n != "rj":
            F2[n] = np.zeros_like(Fo[n])
    fs = _fs(g, F2)
    rj = Fo["rj"]
    rjm = np.concatenate([.5 * (rj[..., :-1] + rj[..., 1:]), 1.5 * rj[..., -1:] - .5 * rj[..., -2:-1]], axis=-1)
    errs["o2o_photodissociation"] = rel(fs[1, 0], -rjm)
    fs = _fs(g, Fo)
    errs["o2o_helium_inert"] = float(np.abs(fs[:, 2]).max() + np.abs(fs[3]).max())
    return errs


# ================================================================================
# comp_ng: O2, O, He as one 3x3 block system (comp_o2o.tex, minor.tex
# notation): d Psi/dt = -e^z d/dZ[A (d/dZ - E) Psi] + eddy diffusion + chemistry,
# A = (m/(tau m_N2)) (T0/T)^0.25 alpha^-1 with alpha the Stefan-Maxwell
# friction matrix, E_i = 1 - (m_i + dm/dZ)/m (diffusive-equilibrium exponent of
# species i) and for He also the thermal diffusion -alpha_T (dT/dZ)/T with the
# thermal diffusion factor of helium alpha_T = -0.38.  Row k holds
#   P_k = A_k (1/dz + E_k/2)/dz,  R_k-1 = A_k (1/dz - E_k/2)/dz   (E per column)
# A trace species (He -> 0) decouples and its diffusion coefficient follows
# Blanc's law, 1/D_He = sum_n Psi_n / D_He,n, with the binary coefficients
# D_He,n varying as T^s_n relative to T^1.75, s = 1.710, 1.749, 1.718 for
# O2, O, N2 (comp_ng.F90:22-25), and phi_He,n the table of comp_ng.F90:23.
PHI_HE = (0.270, 0.404, 0.322)   # He with O2, O, N2
SS_HE = (1.710, 1.749, 1.718)


def _comp_rows(g, Fo, fs=None):
    st = state(g, Fo)
    fs = _fs(g, Fo) if fs is None else fs
    cap, res = capture(g, lambda: st.comp_ng(fs, Fo["tlbc"], Fo["flx_he"]))
    c = cap[38:68]
    return c[0:9], c[9:18], c[18:27], c[27:30], res


def _blk(x, j, i, k):
    return x[:, j, i, k].reshape(3, 3)


def _comp_state(he=None):
    g, Fo = base_state("small", difhor=0)
    j, i = (g["nlat"] + 4) // 2, (g["nlon"] + 4) // 2
    Fo = uniform(g, Fo, at=(j, i))
    for n in _NO_CHEM:
        Fo[n] = np.zeros_like(Fo[n])
    Fo["w"] = np.zeros_like(Fo["w"])
    if he is not None:
        Fo["he"] = np.ascontiguousarray(Fo["he"] * 0 + he)
    return g, Fo, j, i


def comp_molecular():
    errs = {}
    g, Fo, j, i = _comp_state(he=1e-14)
    nk, dz = g["nk"], g["dz"]
    g0 = grid_with(g, difk=[0.] * nk)
    P, Q, R, F, _ = _comp_rows(g0, Fo)
    tn, mb = Fo["tn"][j, i], Fo["mbar"][j, i]
    mass = (32., 16., 4.)
    e_err = a_err = dec = 0.
    for k in range(1, nk - 2):
        Pk, Rk = _blk(P, j, i, k), _blk(R, j, i, k - 1)
        A = (Pk + Rk) * dz ** 2 / 2.
        Ti = .5 * (tn[k] + tn[k - 1])
        mi = .5 * (mb[k] + mb[k - 1])
        dm = (mb[k] - mb[k - 1]) / dz
        dT = (tn[k] - tn[k - 1]) / dz
        E = np.array([1. - (m + dm) / mi for m in mass])
        E[2] += 0.38 * dT / Ti
        for m in range(3):
            col = (Pk[:, m] - Rk[:, m]) * dz
            big = np.abs(A[:, m]) > 1e-6 * np.abs(A).max()
            e_err = max(e_err, rel(col[big] / A[big, m], np.full(big.sum(), E[m]), abs(E[m]) + 1.))
        o2, o1 = .5 * (Fo["o2"][j, i][k] + Fo["o2"][j, i][k - 1]), .5 * (Fo["o1"][j, i][k] + Fo["o1"][j, i][k - 1])
        psi = (o2, o1, 1. - o2 - o1)
        blanc = sum(PHI_HE[n] * (Ti / T00) ** (1.75 - SS_HE[n]) * psi[n] for n in range(3))
        a_he = -(mi / (TAU * 28.)) * (T00 / Ti) ** 0.25 / blanc
        a_err = max(a_err, rel(A[2, 2], a_he, abs(a_he)))
        dec = max(dec, abs(A[2, 0]) / abs(A[2, 2]), abs(A[2, 1]) / abs(A[2, 2]))
    errs["comp_E"] = e_err
    errs["comp_A_trace_he"] = a_err
    errs["comp_trace_he_decoupled"] = dec
    # flux form (no chemistry, no vertical wind): every unknown's coefficients
    # sum to zero over its rows once the time term e^-Z/(2 dt) is removed
    ez = np.asarray(g["expz"])
    ff = 0.
    for k in range(1, nk - 3):
        S = _blk(R, j, i, k - 1) + _blk(Q, j, i, k) - ez[k] * g["dtx2inv"] * np.eye(3) + _blk(P, j, i, k + 1)
        ff = max(ff, np.abs(S).max() / np.abs(_blk(Q, j, i, k)).max())
    errs["comp_flux_form"] = ff
    return errs


def comp_eddy_and_sources():
    g, Fo, j, i = _comp_state()
    nk, dz = g["nk"], g["dz"]
    g0 = grid_with(g, difk=[0.] * nk)
    P0_, Q0_, R0_, F0_, _ = _comp_rows(g0, Fo)
    ke = np.exp(-np.asarray(g["zpint"])) * 2e-6
    P1_, Q1_, R1_, F1_, _ = _comp_rows(grid_with(g, difk=list(ke)), Fo)
    mb = Fo["mbar"][j, i]
    ezi = np.exp(-np.asarray(g["zpint"]))
    k = np.arange(1, nk - 2)
    dmm = (mb[k] - mb[k - 1]) / dz / (.5 * (mb[k] + mb[k - 1]))
    pe = -ezi[k] * ke[k] * (1. / dz - .5 * dmm) / dz
    re = -ezi[k] * ke[k] * (1. / dz + .5 * dmm) / dz
    errs = {}
    dP = (P1_ - P0_)[:, j, i, :]
    dR = (R1_ - R0_)[:, j, i, :]
    errs["comp_eddy_P"] = max(rel(dP[4 * d, k], pe) for d in range(3))
    errs["comp_eddy_R"] = max(rel(dR[4 * d, k - 1], re) for d in range(3))
    # diagonal: the coefficient of Psi_k in -(Phi_k+1 - Phi_k)/dz
    kq = np.arange(1, nk - 3)
    dmq = (mb[kq + 1] - mb[kq]) / dz / (.5 * (mb[kq + 1] + mb[kq]))
    dmk = (mb[kq] - mb[kq - 1]) / dz / (.5 * (mb[kq] + mb[kq - 1]))
    qe = (ezi[kq] * ke[kq] * (1. / dz + .5 * dmk) + ezi[kq + 1] * ke[kq + 1] * (1. / dz - .5 * dmq)) / dz
    dQ = (Q1_ - Q0_)[:, j, i, :]
    errs["comp_eddy_Q"] = max(rel(dQ[4 * d, kq], qe) for d in range(3))
    errs["comp_eddy_diagonal_only"] = float(max(np.abs(dP[e, k]).max() for e in (1, 2, 3, 5, 6, 7)) / np.abs(pe).max())
    # horizontal diffusion tendencies enter the right-hand side as e^-Z hd
    ez = np.asarray(g["expz"])
    _, _, _, Fb, _ = _comp_rows(g, Fo)
    errs["comp_rhs_hd"] = 0.
    for n, row in (("hdo2", 0), ("hdo1", 1), ("hdhe", 2)):
        F2 = dict(Fo)
        d = np.linspace(1e-8, 3e-8, nk)
        set_levels(F2, n, Fo[n][j, i] + d)
        _, _, _, Fx, _ = _comp_rows(g, F2)
        errs["comp_rhs_hd"] = max(errs["comp_rhs_hd"],
                                  rel((Fx[row, j, i] - Fb[row, j, i])[1:nk - 2], (ez * d)[1:nk - 2]))
    return errs


def comp_lower_boundary():
    """Ghost values of the majors below the lower boundary: Psi_-1 = b Psi_0 + fb
    (minor.tex, lower boundary of O2 and O; b, fb from bndcmp): row 0 takes
    P_0 Psi_-1 into its diagonal (+P_0 b) and right-hand side (-P_0 fb).  The
    ghost values also set the bottom interface's state, so (b, fb) are varied
    at a fixed ghost value b Psi_0 + fb: (0, g0) against (I, g0 - Psi_0) gives
    P_0, and the model's (b, fb) against (0, g0) must then differ by P_0 b on
    the diagonal and by P_0 b Psi_0 on the right-hand side."""
    g, Fo, j, i = _comp_state()
    b = np.asarray(g["b"], dtype=np.float64)
    fb = np.asarray(g["fb"], dtype=np.float64)
    psi0 = np.array([Fo["o2"][j, i, 0], Fo["o1"][j, i, 0], Fo["he"][j, i, 0]])
    ghost = b @ psi0 + fb
    _, QA, _, FA, _ = _comp_rows(g, Fo)
    _, QB, _, FB, _ = _comp_rows(grid_with(g, b=[[0.] * 3] * 3, fb=list(ghost)), Fo)
    PC, QC, _, FC, _ = _comp_rows(grid_with(g, b=np.eye(3).tolist(), fb=list(ghost - psi0)), Fo)
    P0 = _blk(QC, j, i, 0) - _blk(QB, j, i, 0)
    dQ = _blk(QA, j, i, 0) - _blk(QB, j, i, 0)
    dF = FA[:, j, i, 0] - FB[:, j, i, 0]
    return {"comp_bottom_b": rel(dQ, P0 @ b), "comp_bottom_fb": rel(dF, P0 @ b @ psi0),
            "comp_bottom_p0": float(np.abs(PC[:, j, i, 0]).max())}


def comp_mirror():
    g, Fo = base_state("small")
    st = state(g, Fo)
    a = st.comp_ng(st.comp_o2o_ng(), Fo["tlbc"], Fo["flx_he"])
    Fm, gm = mirror_lon(g, Fo)
    sm = state(gm, Fm)
    b = sm.comp_ng(sm.comp_o2o_ng(), Fm["tlbc"], Fm["flx_he"])
    return {"comp_mirror": max(rel(mirror_out(n, a[n]), b[n]) for n in a if n != "flx_he")}


TOL = {
    "dt_G": 1e-13, "dt_F": 1e-12, "dt_diag": 1e-12, "dt_bottom_q": 1e-12, "dt_bottom_p": 0.0,
    "dt_top_q": 1e-12, "dt_top_r": 0.0, "dt_rhs_time": 1e-13, "dt_rhs_qji_tn": 1e-13,
    "dt_rhs_cool_exp": 1e-9, "dt_rhs_recomb": 1e-9, "dt_mirror": 1e-14,
    "dt_residual": 1e-12, "dt_top_extrap": 1e-14, "dt_time_filter": 1e-15,
    "duv_P": 1e-13, "duv_R": 1e-13, "duv_PR_offdiag": 0.0, "duv_Q": 1e-13, "duv_Q_coriolis": 1e-13,
    "duv_bottom_Q": 1e-13, "duv_bottom_P": 0.0, "duv_top_Q": 1e-13, "duv_top_R": 0.0,
    "duv_w_P": 1e-9, "duv_w_R": 1e-9, "duv_rhs": 1e-13, "duv_rhs_bottom": 1e-13, "duv_pressure_grad": 1e-9,
    "duv_mirror": 1e-13,
    "minor_A_no": 1e-13, "minor_E_no": 1e-12, "minor_top_no": 1e-13,
    "minor_A_ar": 1e-13, "minor_E_ar": 1e-12, "minor_top_ar": 1e-13, "minor_flux_form": 1e-12,
    "minor_eddy_P": 1e-9, "minor_eddy_R": 1e-9, "minor_eddy_Q": 1e-9, "minor_w_P": 1e-9, "minor_w_R": 1e-9,
    "minor_chem_no_beta1": 1e-9, "minor_chem_no_beta3": 1e-9, "minor_chem_n4s_beta1": 1e-9,
    "minor_chem_n2d_beta5": 1e-9, "minor_chem_ar_none": 0.0, "minor_bottom_dirichlet": 0.25,
    "minor_mirror": 1e-13,
    "oplus_P": 1e-12, "oplus_R": 1e-12, "oplus_top_q": 1e-12, "oplus_top_r": 0.0,
    "oplus_chem_q": 1e-12, "oplus_chem_pr": 1e-30, "oplus_chem_prod": 1e-9,
    "oplus_drift": 1e-9, "oplus_wind": 1e-9, "oplus_smooth_nm": 1e-9,
    "oplus_matrix_independent": 0.0, "oplus_rhs_affine": 1e-13, "oplus_Fe_gravity": 1e-12,
    "oplus_mirror": 1e-13,
    "o2o_mass_rj": 1e-15, "o2o_active_rj": 0.0, "o2o_mass_rkm12": 1e-15, "o2o_active_rkm12": 0.0,
    "o2o_photodissociation": 1e-15, "o2o_helium_inert": 0.0,
    "comp_E": 1e-12, "comp_A_trace_he": 1e-12, "comp_trace_he_decoupled": 1e-9, "comp_flux_form": 1e-12,
    "comp_eddy_P": 1e-9, "comp_eddy_R": 1e-9, "comp_eddy_Q": 1e-9, "comp_eddy_diagonal_only": 0.0, "comp_rhs_hd": 1e-9,
    "comp_bottom_b": 1e-9, "comp_bottom_fb": 1e-9, "comp_bottom_p0": 0.0, "comp_mirror": 1e-13,
}

CHECKS = {
    "dt_ng": (dt_coefficients, dt_sources, dt_mirror, dt_solution),
    "duv_ng": (duv_coefficients, duv_advection_and_sources, duv_mirror),
    "minor_ng": (minor_molecular, minor_eddy_and_advection, minor_chemistry, minor_lower_boundary, minor_mirror),
    "oplus_ng": (oplus_diffusion, oplus_chemistry, oplus_transport_terms, oplus_affine_and_forcing, oplus_mirror),
    "comp_o2o_ng": (o2o_conservation,),
    "comp_ng": (comp_molecular, comp_eddy_and_sources, comp_lower_boundary, comp_mirror),
}

# planted mistakes of oracle/column.c, by routine (each must fail a check)
MUTATIONS = {
    "dt_ng": {101: "eddy term F without its diffusion coefficient", 102: "sign of F(k+1) in the diagonal",
              103: "adiabatic W R/m term dropped", 104: "density from exp(-Z) at the wrong level",
              105: "explicit cooling with the wrong sign", 106: "interface temperature one level off",
              107: "advection with the wind components exchanged", 108: "lower boundary ghost factor 1 for 2"},
    "duv_ng": {201: "metric term u tan(phi)/R with the wrong sign", 202: "off-diagonal ion drag l_xy, l_yx exchanged",
               203: "vertical advection dropped from P", 204: "viscosity of the upper interface one level off",
               205: "zonal pressure gradient without 1/(R cos phi)", 206: "lower boundary ghost factor 1 for 2",
               207: "zonal n-1 wind smoothed from the meridional one"},
    "minor_ng": {301: "phi_x(N2) in Blanc's sum where the nest form has phi_x(He)",
                 302: "eddy term of the lower interface with dm/dZ of the wrong sign",
                 303: "production multiplied by p0 e^-z m/(kT) instead of divided",
                 304: "molecular coefficient of the upper neighbour one level off",
                 305: "zero-flux top row with the thermal term of the wrong sign",
                 306: "lower boundary value term halved"},
    "oplus_ng": {401: "6 v_z n/R_E drift term with the wrong sign", 402: "upper neighbour's wind taken at its own level",
                 403: "loss frequency doubled", 404: "top flux condition with the ratio inverted",
                 405: "gravity term of the ambipolar flux with the wrong sign",
                 406: "n-1 smoothing with the full Shapiro constant in every sub-step"},
    "comp_o2o_ng": {501: "O loss by recombination at half its rate", 502: "O gain from photodissociation halved",
                    503: "O gain from O2 without the mass ratio m_O/m_O2",
                    504: "O2 gain by recombination with one factor 1/m_O"},
    "comp_ng": {601: "eddy diagonal with dm/dZ of the wrong sign at the lower interface",
                602: "diffusion coefficient with (T0/T)^0.5", 603: "lower boundary matrix b transposed",
                604: "thermal diffusion of He with the wrong sign", 605: "horizontal diffusion of O2 dropped",
                606: "upper interface's exponent E taken from the lower interface"},
}


def run_checks(routine):
    errs = {}
    for fn in CHECKS[routine]:
        errs.update(fn())
    return errs


def failures(errs):
    return {n: e for n, e in errs.items() if not (e <= TOL[n])}
