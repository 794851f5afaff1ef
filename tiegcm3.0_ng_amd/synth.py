"""Seeded synthetic model state for the column step.

Fields have the magnitudes of a TIE-GCM thermosphere (cgs units, winds in
cm/s): an exponential TN profile 200-1500 K, O/O2/He mass mixing ratios with
O rising and O2 falling with height, N2 = 1-O-O2-He, eddy diffusivities
decaying as exp(-z), winds of tens of m/s, and so on.  Every field is a smooth
horizontal pattern times a vertical profile times (1 + small seeded noise), so
horizontal stencils and vertical solves see non-trivial data.

Generation runs in torch (CPU for tests, the GPU for the large benchmark
grids) in float64 and returns arrays in the *logical* shape (nk, nlatd, nlond)
for 3-D fields and (nlatd, nlond) for 2-D fields, where nlatd = nlat+4 and
nlond = nlon+4 include the two ghost rows/columns on every side
(fields_ng.F90:137-140).
"""
import math

import torch

# cons.F:25-46, :320-323 and set_const_fields_ng.F90:21
RE = 6.37122e8
P0 = 5.0e-4
GASK = 8.314e7
GRAV = 870.0
BOLTZ = 1.38e-16
OMEGA = 7.292e-5
ZIBOT = -7.0
SMOOTH_FAC = 3.e-3

# name: (nlon, nlat, nk, dlev, step_s, default_step_s, region)
# region None = global-sized domain; else (glat0, glon0, dlat, dlon) of a nest.
CONFIGS = {
    "tiny": (8, 6, 12, 1.0, 60., 60., None),
    "small": (16, 12, 29, 0.5, 60., 60., None),
    "tall": (12, 8, 113, 0.125, 10., 10., None),         # few columns, all 113 levels
    "g5": (72, 36, 29, 0.5, 60., 60., None),            # 5 deg, 29 levels
    "g2p5": (144, 72, 57, 0.25, 30., 30., None),        # 2.5 deg, 57 levels
    "g1p25": (288, 144, 113, 0.125, 10., 10., None),    # 1.25 deg, 113 levels
    "g0p625": (576, 288, 113, 0.125, 5., 5., None),     # 0.625 deg, 113 levels
    "nest1": (120, 60, 57, 0.25, 15., 30., (20.0, 60.0, 0.5, 0.5)),
    "nest2": (100, 50, 57, 0.25, 3., 30., (30.0, 80.0, 0.1, 0.1)),
    # a small nested pair inside "small" for the coupling tests: a 2.5 deg nest
    # on the same levels and a 0.5 deg nest inside it on finer levels
    "nest_s1": (20, 12, 29, 0.5, 30., 60., (-10.0, 40.0, 2.5, 2.5)),
    "nest_s2": (10, 8, 57, 0.25, 10., 60., (-4.0, 45.0, 0.5, 0.5)),
}
SEED_BASE = 0x7E6C0000


def bndcmp(dz):
    """lbc.F:527-594: lower-boundary matrix b and vector fb of comp and minor.

    Host-side set-up used by the synthetic state; the same formula is
    computed by the library (tgx_bndcmp) and by the oracle."""
    alfa, pshelb = 0.22 + 0.014, 0.1154E-5
    e = [[0., 0., 0.], [0., 1., 0.], [0., 0., 0.]]
    f = [[1., 1., 0.], [0., -1., 0.], [0., 0., 1.]]
    wm1 = [[e[l][m] / dz - f[l][m] / 2. for m in range(3)] for l in range(3)]
    wm2 = [[e[l][m] / dz + f[l][m] / 2. for m in range(3)] for l in range(3)]
    A = wm1
    det = (A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2]
           + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0])
    adj = [[A[1][1] * A[2][2] - A[1][2] * A[2][1], -(A[0][1] * A[2][2] - A[0][2] * A[2][1]), A[0][1] * A[1][2] - A[0][2] * A[1][1]],
           [-(A[1][0] * A[2][2] - A[1][2] * A[2][0]), A[0][0] * A[2][2] - A[0][2] * A[2][0], -(A[0][0] * A[1][2] - A[0][2] * A[1][0])],
           [A[1][0] * A[2][1] - A[1][1] * A[2][0], -(A[0][0] * A[2][1] - A[0][1] * A[2][0]), A[0][0] * A[1][1] - A[0][1] * A[1][0]]]
    wm3 = [[adj[l][m] / det for m in range(3)] for l in range(3)]
    g = [-alfa, -0., -pshelb]
    b = [[wm3[l][0] * wm2[0][m] + wm3[l][1] * wm2[1][m] + wm3[l][2] * wm2[2][m] for m in range(3)] for l in range(3)]
    fb = [wm3[l][0] * g[0] + wm3[l][1] * g[1] + wm3[l][2] * g[2] for l in range(3)]
    return b, fb


# Ring-filter chunk tables of the small test grids run as global grids (the
# model resolutions' tables are tgx.RING_CHUNKS, cons.F:331-355): chunk counts
# dividing nlon, from the pole equatorward.
TEST_RING_CHUNKS = {(8, 6): [2, 4], (16, 12): [4, 8, 16], (12, 8): [3, 6]}


def make_grid(name, nstep_sub=1, calc_helium=1, colfac=1.5, opdiffcap=0., difhor=1, n2_top=None, glb=False):
    """Grid constants of configuration `name`.  n2_top: if set, make_state
    shifts O in the upper levels so that N2 = 1-O-O2-He falls to n2_top there
    (the comp post pass then extrapolates N2 above its 0.01 level).  glb: the
    configuration as a global grid (dt.F ... comp.F): nlat latitude cells
    from pole to pole (cell centres -90+(j+0.5)dphi, cons.F:304, dphi =
    pi/nlat), periodic longitudes, the ghost rows on the latitudes across
    the poles (where cos(lat) changes sign, cons.F:314-317), and the ring
    filter's chunk table."""
    nlon, nlat, nk, dlev, step, default_step, region = CONFIGS[name]
    nlond, nlatd = nlon + 4, nlat + 4
    dtr = math.pi / 180.
    if glb:
        assert region is None, "a nest domain is not a global grid"
        dlat = 180. / nlat
        glat = [-90. + (j - 2 + .5) * dlat for j in range(nlatd)]
        dlon = 360. / nlon
        glon = [-180. + (i - 2) * dlon for i in range(nlond)]
    elif region is None:
        dlat = 180. / (nlatd + 1)         # keep ghost rows off the poles
        glat = [-90. + (j + 1) * dlat for j in range(nlatd)]
        dlon = 360. / nlon
        glon = [(i - 2) * dlon for i in range(nlond)]
    else:
        glat0, glon0, dlat, dlon = region
        glat = [glat0 + (j - 2) * dlat for j in range(nlatd)]
        glon = [glon0 + (i - 2) * dlon for i in range(nlond)]
    expdz = math.exp(-dlev)
    expz, difk, dift = [math.exp(-ZIBOT - .5 * dlev)], [5.0e-6], [5.0e-6]
    for k in range(1, nk):                # set_const_fields_ng.F90:57-66
        expz.append(expz[-1] * expdz)
        difk.append(difk[-1] * expdz)
        dift.append(dift[-1] * expdz)
    xmue = [difk[k] if k != nk - 1 else difk[k - 1] for k in range(nk)]
    expzmid = math.exp(-.5 * dlev)
    b, fb = bndcmp(dlev)
    cs = [math.cos(x * dtr) for x in glat]
    return dict(
        name=name, nk=nk, nlon=nlon, nlat=nlat, dz=dlev, dlamda=dlon * dtr, dphi=dlat * dtr,
        dtx2inv=1. / (2. * step), shapiro=step / default_step * SMOOTH_FAC,
        expzmid=expzmid, expzmid_inv=1. / expzmid,
        expz=expz, difk=difk, dift=dift, xmue=xmue,
        zpint=[ZIBOT + k * dlev for k in range(nk)],
        zpmid=[ZIBOT + .5 * dlev + k * dlev for k in range(nk)],
        glat=glat, glon=glon, cs=cs, tanphi=[math.tan(x * dtr) for x in glat],
        cor=[2. * OMEGA * math.sin(x * dtr) for x in glat], racs=[1. / (RE * c) for c in cs],
        nstep_sub=nstep_sub, calc_helium=calc_helium, colfac=colfac, opdiffcap=opdiffcap,
        difhor=difhor, b=b, fb=fb, pshelb=0.1154E-5, ar_glbm=[1.2e-3, 1.1e-3],
        step=step, seed=SEED_BASE + sorted(CONFIGS).index(name), n2_top=n2_top,
        glb=bool(glb), chunk=_ring_chunks(nlon, nlat) if glb else [],
    )


def _ring_chunks(nlon, nlat):
    from . import tgx
    t = tgx.RING_CHUNKS.get((nlon, nlat)) or TEST_RING_CHUNKS.get((nlon, nlat))
    if t is None:
        raise ValueError("no ring-filter chunk table for a %dx%d global grid" % (nlon, nlat))
    return list(t)


# Fields whose values across the poles change sign (fields.F:548-553: the
# horizontal wind components and their n-1 levels; the other components of
# vectors are treated alike in the synthetic state).
POLE_ODD = ("un", "vn", "un_nm", "vn_nm", "ui", "vi", "u_lbc", "v_lbc", "ulbc", "vlbc", "bx", "by")


def fill_global_halos(F, grid):
    """The halos of a global state, as the model fills them before a step:
    the periodic points from the interior longitudes (mp_periodic_f4d:
    i = 0, 1 <- nlon, nlon+1 and nlon+2, nlon+3 <- 2, 3), then the two rows
    across each pole from the interior rows next to it, shifted by 180
    degrees in longitude, with the sign of POLE_ODD (mk_polelat.F:2-57:
    latitude 0 <- 1, -1 <- 2, nlat+1 <- nlat, nlat+2 <- nlat-1).  F: dict of
    logical-layout tensors (3-D (nk, nlat+4, nlon+4), 2-D (nlat+4, nlon+4)),
    changed in place and returned."""
    nlon, nlat = grid["nlon"], grid["nlat"]
    nlond, nlatd = nlon + 4, nlat + 4
    src = [((i - 2 + nlon // 2) % nlon) + 2 for i in range(nlond)]
    for n, t in F.items():
        if n.startswith(("lonb_", "latb_")) or t.shape[-2:] != (nlatd, nlond):
            continue
        t[..., 0] = t[..., nlon]
        t[..., 1] = t[..., nlon + 1]
        t[..., nlon + 2] = t[..., 2]
        t[..., nlon + 3] = t[..., 3]
        sign = -1. if n in POLE_ODD else 1.
        for jnew, jold in ((1, 2), (0, 3), (nlat + 2, nlat + 1), (nlat + 3, nlat)):
            t[..., jnew, :] = sign * t[..., jold, src]
    return F


def make_global_state(grid, device="cpu", seed=None):
    """make_state for a global grid, its halos filled by fill_global_halos."""
    assert grid.get("glb"), "make_global_state needs a grid made with glb=True"
    return fill_global_halos(make_state(grid, device=device, seed=seed), grid)


def make_state(grid, device="cpu", seed=None):
    """Returns dict name -> float64 tensor in logical layout (see module doc)."""
    nk, nlat, nlon = grid["nk"], grid["nlat"], grid["nlon"]
    nlatd, nlond = nlat + 4, nlon + 4
    gen = torch.Generator(device=device)
    gen.manual_seed(int(grid["seed"] if seed is None else seed))
    dt = torch.float64
    kw = dict(device=device, dtype=dt)

    def noise(*shape):
        return torch.randn(*shape, generator=gen, **kw)

    zmid = torch.tensor(grid["zpmid"], **kw).view(nk, 1, 1)
    zint = torch.tensor(grid["zpint"], **kw).view(nk, 1, 1)
    s = zmid - ZIBOT
    if grid.get("glb"):   # periodic in longitude over the nlon interior points
        X = (2 * math.pi * (torch.arange(nlond, **kw) - 2) / nlon).view(1, 1, nlond)
    else:
        X = (2 * math.pi * torch.arange(nlond, **kw) / nlond).view(1, 1, nlond)
    Y = (math.pi * (torch.arange(nlatd, **kw) + .5) / nlatd).view(1, nlatd, 1)
    h1 = torch.sin(2 * X + .3) * torch.cos(Y)
    h2 = torch.cos(3 * X - .2) * torch.sin(2 * Y + .1)
    sig = torch.sigmoid(zmid / 1.5)

    def n3(a):
        return a * noise(nk, nlatd, nlond)

    F = {}
    tn = (200. + 1300. * (1. - torch.exp(-s / 2.5))) * (1. + .05 * h1) * (1. + n3(.002))
    F["tn"] = tn
    F["tn_nm"] = tn * (1. + n3(1e-3))
    F["un"] = 5000. * (h1 + .5 * h2) * (.3 + s / 14.) + n3(300.)
    F["vn"] = 5000. * (h2 - .3 * h1) * (.3 + s / 14.) + n3(300.)
    F["un_nm"] = F["un"] + n3(50.)
    F["vn_nm"] = F["vn"] + n3(50.)
    o1 = (.01 + .92 / (1. + torch.exp(-zmid / 1.2))) * (1. + .02 * h2) * (1. + n3(1e-3))
    o2 = (.22 * torch.exp(-s / 3.) * (1. + .03 * h1) + .002) * (1. + n3(1e-3))
    he = 5e-6 * torch.exp(s / 1.6) * (1. + .1 * h1) * (1. + n3(1e-3))
    if grid.get("n2_top") is not None:  # O takes up the N2 of the upper levels
        # (and He's share too when He is not solved: comp then sets He to its
        # lower-boundary value, so N2 = 1-O-O2 there)
        top = torch.sigmoid((zmid - 3.5) / 0.5)
        he_share = he if grid["calc_helium"] != 0 else 0.
        o1 = o1 + top * (1. - o2 - o1 - he_share - grid["n2_top"])
    n2 = torch.clamp(1. - o2 - o1 - he, min=5e-3)
    F["o1"], F["o2"], F["he"], F["n2"] = o1, o2, he, n2
    for nm in ("o1", "o2", "he"):
        F[nm + "_nm"] = F[nm] * (1. + n3(1e-3))
    mbar = 1. / (o2 / 32. + o1 / 16. + he / 4. + n2 / 28.)
    F["mbar"] = mbar
    F["barm"] = mbar * (1. + n3(1e-4))
    F["xnmbar"] = P0 * torch.exp(-zmid) * mbar / (BOLTZ * tn)
    F["xnmbari"] = P0 * torch.exp(-zint) * mbar / (BOLTZ * tn)
    scht = GASK * tn / (mbar * GRAV)
    F["scht"] = scht
    F["schti"] = scht * (1. - .01 * torch.tanh(s - 7.)) * (1. + n3(1e-4))
    z0 = 9.6e6 + 1e5 * h1[0]
    F["z"] = z0.unsqueeze(0) + torch.cumsum(scht * grid["dz"], dim=0) - scht[0:1] * grid["dz"]
    w = 1e-5 * (h2 + .3 * noise(nk, nlatd, nlond)) * (1. + s / 7.)
    F["w"] = w
    F["w_upd"] = w + n3(2e-6)
    ti = tn * (1. + .3 * sig)
    F["ti"] = ti
    F["te"] = ti * (1. + .5 * sig) * (1. + n3(1e-3))
    ne = (1e4 + 1e6 * torch.exp(-((zmid - 1.) / 2.) ** 2)) * (1. + .2 * h1) * (1. + n3(1e-3))
    F["ne"] = ne
    F["op"] = .9 * ne * torch.sigmoid(zmid) + 10.
    F["op_nm"] = F["op"] * (1. + n3(1e-3))
    F["o2p"] = 1e4 * torch.exp(-((zmid + 3.) / 2.) ** 2) * (1. + .1 * h2) + 10.
    F["n2p"] = 1e2 * torch.exp(-((zmid + 1.) / 2.) ** 2) * (1. + .1 * h1) + 1.
    F["nplus"] = 3e3 * torch.exp(-((zmid - 2.) / 2.) ** 2) * (1. + .1 * h2) + 1.
    F["nop"] = 1e5 * torch.exp(-((zmid + 4.) / 2.) ** 2) * (1. + .1 * h1) + 10.
    F["n4s"] = 2e-6 * torch.exp(s / 4.) * (1. + .05 * h2) * (1. + n3(1e-3))
    F["n2d"] = 1e-8 * torch.exp(s / 3.) * (1. + .05 * h1) * (1. + n3(1e-3))
    F["no"] = 3e-7 * torch.exp(-((zmid + 3.) / 3.) ** 2) * (1. + .05 * h1) + 1e-9
    F["ar"] = 1.2e-3 * torch.exp(-s / 3.) * (1. + .02 * h2) + 1e-7
    for nm in ("n4s", "n2d", "no", "ar"):
        F[nm + "_nm"] = F[nm] * (1. + n3(1e-3))
    F["cp"] = 1.0e7 * (1. + .2 * o1) * (1. + n3(1e-4))
    F["kt"] = 56. * tn ** .69 * (1. + .3 * o1)
    F["km"] = 3.34e-4 * tn ** .71 * (1. + .1 * o1)
    for nm in ("hdt", "hdu", "hdv"):
        F[nm] = n3(1e-2) * (1. + s)
    for nm in ("hdo2", "hdo1", "hdhe"):
        F[nm] = n3(1e-9)
    F["qji_tn"] = 50. * torch.exp(-((zmid - 1.) / 3.) ** 2) * (1. + .5 * h1 ** 2)
    F["cool_imp"] = 2e-5 * torch.exp(-((zmid + 2.) / 3.) ** 2) + 1e-7
    F["cool_exp"] = 5e2 * torch.exp(-((zmid + 4.) / 2.) ** 2) * (1. + .1 * h2)
    F["qtotal"] = 3e2 * torch.exp(-((zmid - 0.) / 3.) ** 2) * (1. + .3 * h1)
    F["rkm12"] = 1e-38 * (300. / tn)
    for nm, a in (("ui", 3e3), ("vi", 3e3), ("wi", 1e2)):
        F[nm] = a * (h2 + h1) * sig + n3(.1 * a)
    for nm, a in (("lxx", 1e-4), ("lyy", 1e-4), ("lxy", 3e-5), ("lyx", 3e-5)):
        F[nm] = a * torch.exp(-((zmid - 1.) / 2.5) ** 2) * (1. + .2 * h1) + a * 1e-2
    F["qop2p"] = 50. * torch.exp(-((zmid + 1.) / 2.) ** 2) + .1
    F["qop2d"] = 80. * torch.exp(-((zmid + 1.) / 2.) ** 2) + .1
    F["qop"] = 200. * torch.exp(-((zmid + 1.) / 2.) ** 2) * (1. + .2 * h1) + 1.
    F["qo2p"] = 300. * torch.exp(-((zmid + 3.) / 2.) ** 2) + 1.
    F["qtef"] = 100. * torch.exp(-((zmid + 1.) / 2.5) ** 2) * (1. + .2 * h2) + .5
    F["rj"] = 3e-6 * torch.sigmoid(zmid + 2.)
    F["rk1"] = 2e-11 * (300. / ti) ** .5
    F["rk2"] = 1e-12 * (300. / ti) ** .4
    F["rk3"] = 1.4e-10 * (300. / ti) ** .44
    F["rk19"] = 4e-8 * (300. / F["te"]) ** .5
    F["rk20"] = 1.5e-7 * (300. / F["te"]) ** .5
    F["rk25"] = 6.6e-8 * (300. / F["te"]) ** .5
    F["beta1"] = 1.5e-11 * torch.exp(-3600. / tn)
    F["beta3"] = 3.4e-11 * (tn / 300.) ** .1
    F["beta5"] = 3.6e-10 * (F["te"] / 300.) ** .5
    F["beta8"] = 4e-6 * torch.sigmoid(zmid)
    F["beta9"] = 2e-6 * torch.sigmoid(zmid + 1.)
    F["beta9n"] = 1e-6 * torch.sigmoid(zmid + 2.)
    F["beta17"] = 1e-32 * (300. / tn) ** .5
    F["ra1"] = 4.2e-7 * (300. / F["te"]) ** .85
    F["ra2"] = 1.9e-7 * (300. / F["te"]) ** .5
    F["ra3"] = 1.8e-7 * (300. / F["te"]) ** .39
    for nm in list(F):
        F[nm] = F[nm].expand(nk, nlatd, nlond).contiguous()
    # 2-D fields
    lat = torch.tensor(grid["glat"], **kw).view(nlatd, 1) * (math.pi / 180.)
    lon = (torch.arange(nlond, **kw) * grid["dlamda"]).view(1, nlond)
    H1, H2 = h1[0], h2[0]

    def n2d_(a):
        return a * noise(nlatd, nlond)

    F["t_lbc"] = 181. + 3. * H1 + n2d_(.1)
    F["u_lbc"] = 500. * H2 + n2d_(10.)
    F["v_lbc"] = 400. * H1 + n2d_(10.)
    F["ulbc"] = F["u_lbc"] + n2d_(5.)
    F["vlbc"] = F["v_lbc"] + n2d_(5.)
    F["tlbc"] = F["t_lbc"] + n2d_(.05)       # flds%tlbc before dt_ng of this step
    F["flx_he"] = 2e-13 * (1. + .5 * H1)   # g/cm^2/s, size of comp.F:106-121
    rlatm = lat + .1 * torch.sin(lon)
    dip = torch.atan(2. * torch.tan(rlatm))
    dec = .15 * torch.cos(lon + lat)
    bz = -torch.sin(dip)
    dipmin = math.sin(2. * grid["dphi"])
    bz = torch.where(bz.abs() < dipmin, torch.where(bz < 0, -dipmin, dipmin) + 0. * bz, bz)
    F["bz"] = bz
    F["bx"] = torch.cos(dip) * torch.sin(dec)
    F["by"] = torch.cos(dip) * torch.cos(dec)
    F["bmod2"] = .3 + .25 * torch.sin(rlatm) ** 2
    F["dipmag"] = dip
    F["sndec"], F["csdec"] = torch.sin(dec), torch.cos(dec)
    F["chi"] = torch.arccos(torch.clamp(torch.cos(lat) * torch.cos(lon - 1.), -1., 1.))
    F["rlatm"] = rlatm
    for nm in ("t_lbc", "u_lbc", "v_lbc", "ulbc", "vlbc", "tlbc", "flx_he", "bz", "bx", "by", "bmod2",
               "dipmag", "sndec", "csdec", "chi", "rlatm"):
        F[nm] = F[nm].expand(nlatd, nlond).contiguous()
    # lateral boundary data: the field's own ghost values, slightly perturbed
    for nm in ("tn", "un", "vn", "o2", "o1", "he", "n2d", "n4s", "no", "ar"):
        f = F[nm]
        F["lonb_" + nm] = torch.stack([f[:, :, 0], f[:, :, nlond - 1]], dim=1) * (1. + 1e-3 * noise(nk, 2, nlatd))
        F["latb_" + nm] = torch.stack([f[:, 0, :], f[:, nlatd - 1, :]], dim=1) * (1. + 1e-3 * noise(nk, 2, nlond))
    nsub = int(grid["nstep_sub"])
    op = F["op"]
    F["lonb_op"] = torch.stack([torch.stack([op[:, :, 0], op[:, :, nlond - 1]], dim=1)
                                * (1. + 1e-3 * noise(nk, 2, nlatd)) for _ in range(nsub)])
    F["latb_op"] = torch.stack([torch.stack([op[:, 0, :], op[:, nlatd - 1, :]], dim=1)
                                * (1. + 1e-3 * noise(nk, 2, nlond)) for _ in range(nsub)])
    return F


def to_fortran_layout(F):
    """Logical layout -> the reference's Fortran layout as numpy C-order arrays:
    3-D (nk,nlatd,nlond) -> (nlatd,nlond,nk); 2-D unchanged; lateral data:
    lonb (nk,2,nlatd) -> (nlatd,2,nk), latb (nk,2,nlond) -> (2,nlond,nk);
    O+ per-sub-step lateral data gains a leading sub-step axis."""
    out = {}
    for nm, t in F.items():
        if nm == "lonb_op":
            out[nm] = t.permute(0, 3, 2, 1).contiguous().cpu().numpy()
        elif nm == "latb_op":
            out[nm] = t.permute(0, 2, 3, 1).contiguous().cpu().numpy()
        elif nm.startswith("lonb_"):
            out[nm] = t.permute(2, 1, 0).contiguous().cpu().numpy()
        elif nm.startswith("latb_") or t.dim() == 3:
            out[nm] = t.permute(1, 2, 0).contiguous().cpu().numpy()
        else:
            out[nm] = t.contiguous().cpu().numpy()
    return out


def _lin_axis(a, xp, x, axis):
    """Linear interpolation of array a along `axis` from points xp to x
    (inside the range of xp)."""
    import numpy as np
    i1 = np.clip(np.searchsorted(xp, x, side="right"), 1, len(xp) - 1)
    i0 = i1 - 1
    w = (x - xp[i0]) / (xp[i1] - xp[i0])
    shape = [1] * a.ndim
    shape[axis] = len(x)
    w = w.reshape(shape)
    return np.take(a, i0, axis=axis) * (1. - w) + np.take(a, i1, axis=axis) * w


def nested_states(grids, device="cpu"):
    """Synthetic states for a nest chain that agree with each other: the
    outermost domain's state as make_state gives it, and every nest's 3-D and
    2-D fields interpolated (linearly in z, latitude and longitude) from the
    outermost state, as a model initialises its nests from the global state
    (init_fields_ng.F90).  The nests' lateral data are those of make_state
    (a coupled run replaces them).  Returns one dict per grid."""
    import numpy as np
    outer = make_state(grids[0], device="cpu")
    out = [outer]
    g0 = grids[0]
    z0, y0, x0 = (np.asarray(g0[k], dtype=np.float64) for k in ("zpint", "glat", "glon"))
    for g in grids[1:]:
        F = make_state(g, device="cpu")
        z, y, x = (np.asarray(g[k], dtype=np.float64) for k in ("zpint", "glat", "glon"))
        for n, t in outer.items():
            if n.startswith(("lonb_", "latb_")) or n not in F:
                continue
            a = t.numpy()
            if a.ndim == 3:
                a = _lin_axis(_lin_axis(_lin_axis(a, z0, z, 0), y0, y, 1), x0, x, 2)
            elif a.ndim == 2:
                a = _lin_axis(_lin_axis(a, y0, y, 0), x0, x, 1)
            else:
                continue
            F[n] = torch.from_numpy(np.ascontiguousarray(a))
        out.append(F)
    if device != "cpu":
        out = [{k: v.to(device) for k, v in F.items()} for F in out]
    return out
