"""ctypes loader for the C oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker.  Arrays follow the reference's Fortran
layout: a 3-D nest field f(nk, nlond, nlatd) is held in numpy as shape
(nlatd, nlond, nk) C-order (identical bytes), 2-D fields as (nlatd, nlond).
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GRID_FIELDS = [
    ("nk", C.c_int), ("nlon", C.c_int), ("nlat", C.c_int),
    ("dz", C.c_double), ("dlamda", C.c_double), ("dphi", C.c_double),
    ("dtx2inv", C.c_double), ("shapiro", C.c_double),
    ("expzmid", C.c_double), ("expzmid_inv", C.c_double),
    ("expz", C.c_void_p), ("difk", C.c_void_p), ("dift", C.c_void_p), ("xmue", C.c_void_p),
    ("zpint", C.c_void_p), ("zpmid", C.c_void_p),
    ("racs", C.c_void_p), ("cs", C.c_void_p), ("tanphi", C.c_void_p), ("cor", C.c_void_p),
    ("glat", C.c_void_p),
    ("nstep_sub", C.c_int), ("calc_helium", C.c_int),
    ("colfac", C.c_double), ("opdiffcap", C.c_double), ("difhor", C.c_double),
    ("b", C.c_double * 9), ("fb", C.c_double * 3), ("pshelb", C.c_double),
    ("glb", C.c_int), ("nlat_filter", C.c_int), ("chunk", C.c_void_p),
]

# Order must match tgx_ofields in tgx_oracle.h.
FIELD_NAMES_3D = (
    "tn tn_nm un vn un_nm vn_nm o2 o2_nm o1 o1_nm he he_nm "
    "op op_nm n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm ti te ne o2p w z "
    "mbar barm xnmbar xnmbari scht schti n2 w_upd "
    "cp kt km hdt hdu hdv hdo2 hdo1 hdhe qji_tn cool_imp cool_exp qtotal rkm12 "
    "ui vi wi lxx lyy lxy lyx qop2p qop2d qop qo2p qtef rj rk1 rk2 rk3 rk19 rk20 rk25 "
    "beta1 beta3 beta5 beta8 beta9 beta9n beta17 ra1 ra2 ra3 n2p nplus nop").split()
FIELD_NAMES_2D = ("t_lbc u_lbc v_lbc ulbc vlbc bx by bz bmod2 dipmag sndec csdec chi rlatm").split()
LATERAL = "tn un vn o2 o1 he n2d n4s no ar op".split()

_FIELDS = ([(n, C.c_void_p) for n in FIELD_NAMES_3D] + [(n, C.c_void_p) for n in FIELD_NAMES_2D]
           + sum([[("lonb_" + n, C.c_void_p), ("latb_" + n, C.c_void_p)] for n in LATERAL], [])
           + [("ar_glbm1", C.c_double), ("ar_glbm2", C.c_double)])


class OGrid(C.Structure):
    _fields_ = GRID_FIELDS


class OFields(C.Structure):
    _fields_ = _FIELDS


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = C.CDLL(path)
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(C.c_void_p)


class OracleState:
    """Holds ctypes views of a synthetic state (dict of numpy arrays)."""

    def __init__(self, grid, fields):
        self.grid_np = grid
        self.fields_np = fields
        g = OGrid()
        g.nk, g.nlon, g.nlat = grid["nk"], grid["nlon"], grid["nlat"]
        for n in ("dz", "dlamda", "dphi", "dtx2inv", "shapiro", "expzmid", "expzmid_inv",
                  "colfac", "opdiffcap", "difhor", "pshelb"):
            setattr(g, n, float(grid[n]))
        g.nstep_sub, g.calc_helium = int(grid["nstep_sub"]), int(grid["calc_helium"])
        self._keep = []
        for n in ("expz", "difk", "dift", "xmue", "zpint", "zpmid", "racs", "cs", "tanphi", "cor", "glat"):
            a = np.ascontiguousarray(grid[n], dtype=np.float64)
            self._keep.append(a)
            setattr(g, n, a.ctypes.data)
        bb = np.asarray(grid["b"], dtype=np.float64).reshape(3, 3)  # b[l][m] = b(l+1,m+1)
        for i in range(9):
            g.b[i] = float(bb.reshape(-1)[i])
        for i in range(3):
            g.fb[i] = float(grid["fb"][i])
        g.glb = int(bool(grid.get("glb", False)))
        chunk = np.ascontiguousarray(grid.get("chunk", []) if g.glb else [], dtype=np.int32)
        self._keep.append(chunk)
        g.nlat_filter = len(chunk)
        g.chunk = chunk.ctypes.data if len(chunk) else None
        self.g = g
        f = OFields()
        for n in FIELD_NAMES_3D + FIELD_NAMES_2D:
            a = fields[n]
            self._keep.append(a)
            setattr(f, n, _p(a).value)
        for n in LATERAL:
            setattr(f, "lonb_" + n, _p(fields["lonb_" + n]).value)
            setattr(f, "latb_" + n, _p(fields["latb_" + n]).value)
        f.ar_glbm1, f.ar_glbm2 = float(grid["ar_glbm"][0]), float(grid["ar_glbm"][1])
        self.f = f
        self.shape3 = (grid["nlat"] + 4, grid["nlon"] + 4, grid["nk"])
        self.shape2 = self.shape3[:2]

    def z3(self):
        return np.zeros(self.shape3)

    def z2(self):
        return np.zeros(self.shape2)

    # ---- column-step routines --------------------------------------------
    def dt_ng(self, tlbc):
        tlbc = tlbc.copy()
        tlbc_nm, tn, tn_nm = self.z2(), self.z3(), self.z3()
        lib().tgx_o_dt_ng(C.byref(self.g), C.byref(self.f), _p(tlbc), _p(tlbc_nm), _p(tn), _p(tn_nm))
        return dict(tn=tn, tn_nm=tn_nm, tlbc=tlbc, tlbc_nm=tlbc_nm)

    def duv_ng(self, tn_upd, Fe, Fn, tlbc_nm, ulbc, vlbc):
        ulbc, vlbc = ulbc.copy(), vlbc.copy()
        ulbc_nm, vlbc_nm = self.z2(), self.z2()
        un, unm, vn, vnm = self.z3(), self.z3(), self.z3(), self.z3()
        lib().tgx_o_duv_ng(C.byref(self.g), C.byref(self.f), _p(tn_upd), _p(Fe), _p(Fn), _p(tlbc_nm),
                           _p(ulbc), _p(vlbc), _p(ulbc_nm), _p(vlbc_nm), _p(un), _p(unm), _p(vn), _p(vnm))
        return dict(un=un, un_nm=unm, vn=vn, vn_nm=vnm, ulbc=ulbc, vlbc=vlbc, ulbc_nm=ulbc_nm, vlbc_nm=vlbc_nm)

    def minor(self, species, n2d_upd, tlbc):
        out, out_nm = self.z3(), self.z3()
        lib().tgx_o_minor_species(C.byref(self.g), C.byref(self.f), C.c_int(species), _p(n2d_upd),
                                  _p(tlbc), _p(out), _p(out_nm))
        return out, out_nm

    def minors(self, tlbc):
        """N2D first, then N4S and NO with the new N2D, then AR (dynamics_ng.F90:147-165)."""
        n2d, n2d_nm = self.minor(2, np.zeros(self.shape3), tlbc)
        n4s, n4s_nm = self.minor(1, n2d, tlbc)
        no, no_nm = self.minor(0, n2d, tlbc)
        ar, ar_nm = self.minor(3, n2d, tlbc)
        return dict(n2d=n2d, n2d_nm=n2d_nm, n4s=n4s, n4s_nm=n4s_nm, no=no, no_nm=no_nm, ar=ar, ar_nm=ar_nm)

    def oplus_ng(self, op, optm1, lonb, latb):
        outs = [self.z3() for _ in range(6)]
        lib().tgx_o_oplus_ng(C.byref(self.g), C.byref(self.f), _p(op), _p(optm1), _p(lonb), _p(latb),
                             *[_p(o) for o in outs])
        return dict(zip(("op", "op_nm", "xiop2p", "xiop2d", "Fe", "Fn"), outs))

    def oplus_subcycled(self, lonb_sub, latb_sub):
        """dynamics_ng.F90:99-121 (dynamics.F:319-359 on a global grid):
        nstep_sub sub-cycles, itp/itc swapped between them.  On a global grid
        a sub-step's O+ keeps the rows across the poles of the step's input
        (the reference exchanges no latitude halo there, oplus.F:942, so they
        hold what the model's buffer held before the sub-cycles)."""
        op, optm1 = self.fields_np["op"], self.fields_np["op_nm"]
        op0, optm10 = op, optm1
        out = None
        for s in range(int(self.grid_np["nstep_sub"])):
            out = self.oplus_ng(op, optm1, lonb_sub[s], latb_sub[s])
            op, optm1 = out["op"], out["op_nm"]
            if self.g.glb:
                op, optm1 = op.copy(), optm1.copy()
                for a, a0 in ((op, op0), (optm1, optm10)):
                    a[:2] = a0[:2]
                    a[-2:] = a0[-2:]
        return out

    def comp_o2o_ng(self):
        fs = np.zeros((4, 3) + self.shape3)
        lib().tgx_o_comp_o2o_ng(C.byref(self.g), C.byref(self.f), _p(fs))
        return fs

    def comp_ng(self, fs, tlbc, flx_he):
        flx_he = flx_he.copy()
        outs = [self.z3() for _ in range(6)]
        lib().tgx_o_comp_ng(C.byref(self.g), C.byref(self.f), _p(fs), _p(tlbc), _p(flx_he),
                            *[_p(o) for o in outs])
        d = dict(zip(("o2", "o2_nm", "o1", "o1_nm", "he", "he_nm"), outs))
        d["flx_he"] = flx_he
        return d


# ---- test hooks --------------------------------------------------------------
CAPTURE_PLANES = 68


def set_capture(buf):
    """Rows handed to each column solve are written into buf, a float64 array
    of CAPTURE_PLANES*ncol*nk (layout of the library's capture mode), or not
    at all with None.  The caller keeps buf alive while it is set."""
    if buf is None:
        lib().tgx_o_set_capture(None)
    else:
        lib().tgx_o_set_capture(_p(buf))


def ringfilter_row(f, nlon, order, oplus, nchnk, dlamda):
    """ringfilter.F on one row, in place: f is (nlev, nlon+4) float64, C order
    (the reference's f(nlonp4, lev0:lev1))."""
    assert f.shape[1] == nlon + 4
    rc = lib().tgx_o_ringfilter_row(_p(f), int(nlon), int(f.shape[0]), int(order), int(bool(oplus)),
                                    int(nchnk), C.c_double(dlamda))
    if rc != 0:
        raise ValueError("ringfilter: order %d >= nlon %d" % (order, nlon))
    return f


def ringfilter_chunks(lat, nlat, chunk):
    arr = (C.c_int * len(chunk))(*[int(c) for c in chunk])
    return lib().tgx_o_ringfilter_chunks(int(lat), int(nlat), len(chunk), arr)


# cons.F:331-355: latitudes filtered next to each pole and their chunk counts
RING_CHUNKS = {
    (72, 36): [3, 9, 18, 36, 36],
    (144, 72): [3, 3, 9, 9, 18, 18, 36, 36, 72, 72],
    (288, 144): [3, 3, 3, 9, 9, 9, 18, 18, 18, 36, 36, 36, 72, 72, 72, 72, 144, 144, 144, 144, 144],
    (576, 288): [3, 3, 3, 3, 3, 9, 9, 9, 9, 9, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36,
                 72, 72, 72, 72, 72, 72, 144, 144, 144, 144, 144, 144, 288, 288, 288, 288, 288, 288],
}


def set_mutation(mid):
    """Plants the deliberate mistake `mid` (see the MUT sites of column.c);
    0 restores the oracle."""
    lib().tgx_o_set_mutation(C.c_int(int(mid)))


# ---- plain solvers ----------------------------------------------------------
def trsolv(a, b, c, f, x, lev0, lev1, k1, k2, lon0, lon1, lonmax):
    lib().tgx_o_trsolv(_p(a), _p(b), _p(c), _p(f), _p(x), lev0, lev1, k1, k2, lon0, lon1, lonmax)
    return x


def trsolv_ng(a, b, c, f, x, nk, ncol, k2):
    lib().tgx_o_trsolv_ng(_p(a), _p(b), _p(c), _p(f), _p(x), nk, ncol, k2)
    return x


def blktri(a, b, c, f, x, lev0, lev1, lon0, lon1, lonmax):
    lib().tgx_o_blktri(_p(a), _p(b), _p(c), _p(f), _p(x), lev0, lev1, lon0, lon1, lonmax)
    return x


def blktri_ng(pp, qq, rr, ss, xx, nk, ncol):
    """duv_ng.F90:218-248 on assembled 2x2 blocks (2,2,nk,ncol), rows 1..nk-1."""
    lib().tgx_o_blktri_ng(_p(pp), _p(qq), _p(rr), _p(ss), _p(xx), nk, ncol)
    return xx


def blktri3(p, q, r, f, x, nk, ncol, k2):
    lib().tgx_o_blktri3(_p(p), _p(q), _p(r), _p(f), _p(x), nk, ncol, k2)
    return x


def matinv3(a):
    """a: (3,3) numpy matrix; returns its inverse via the oracle (adjugate/det)."""
    af = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    out = np.zeros(9)
    lib().tgx_o_matinv3(_p(af), _p(out))
    return out.reshape(3, 3, order="F")


def bndcmp(dz):
    b = np.zeros(9)
    fb = np.zeros(3)
    lib().tgx_o_bndcmp(C.c_double(dz), _p(b), _p(fb))
    return b.reshape(3, 3), fb


def df3d(f, order):
    """dffm_ng df_3d on a (nlatd, nlond, nk) array."""
    out = np.zeros_like(f)
    lib().tgx_o_df3d(_p(f), _p(out), int(order), f.shape[2], f.shape[1], f.shape[0])
    return out


def fm3d(f, order):
    out = np.zeros_like(f)
    lib().tgx_o_fm3d(_p(f), _p(out), int(order), f.shape[2], f.shape[1], f.shape[0])
    return out


def _methods():
    def advec(self, f):
        out = self.z3()
        lib().tgx_o_advec(C.byref(self.g), _p(f), _p(self.fields_np["un"]), _p(self.fields_np["vn"]), _p(out))
        return out

    def smooth(self, f, shapiro=None):
        out = self.z3()
        s = self.grid_np["shapiro"] if shapiro is None else shapiro
        lib().tgx_o_smooth(C.byref(self.g), _p(f), _p(out), C.c_double(s))
        return out

    def div(self, f, direction):
        out = self.z3()
        lib().tgx_o_div(C.byref(self.g), _p(f), _p(out), int(direction))
        return out
    OracleState.advec, OracleState.smooth, OracleState.div = advec, smooth, div


_methods()


# ---- interpolation between nest levels (interp.F90, interp_fields_ng.F90) ---
def _v(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def interp_find(xp, x0):
    """find (interp.F90:105-126): 1-based bracket, 0 below, len(xp) at or above."""
    xp = _v(xp)
    lib().tgx_o_interp_find.restype = C.c_int
    return lib().tgx_o_interp_find(_p(xp), len(xp), C.c_double(x0))


def interp1d(x, xp, fp):
    x, xp, fp = _v(x), _v(xp), _v(fp)
    f = np.empty(len(x))
    lib().tgx_o_interp1d(_p(x), len(x), _p(xp), len(xp), _p(fp), _p(f))
    return f


def interp2d(x, y, xp, yp, fp):
    """fp: C-order (nyp, nxp) (the reference's fp(nxp, nyp)); returns (ny, nx)."""
    x, y, xp, yp, fp = _v(x), _v(y), _v(xp), _v(yp), _v(fp)
    assert fp.shape == (len(yp), len(xp))
    f = np.empty((len(y), len(x)))
    lib().tgx_o_interp2d(_p(x), len(x), _p(y), len(y), _p(xp), len(xp), _p(yp), len(yp), _p(fp), _p(f))
    return f


def interp3d(z, x, y, zp, xp, yp, fp, zlog):
    """fp: C-order (nyp, nxp, nzp) (the reference's fp(nzp, nxp, nyp));
    returns (ny, nx, nz)."""
    z, x, y, zp, xp, yp, fp = _v(z), _v(x), _v(y), _v(zp), _v(xp), _v(yp), _v(fp)
    assert fp.shape == (len(yp), len(xp), len(zp))
    f = np.empty((len(y), len(x), len(z)))
    lib().tgx_o_interp3d(_p(z), len(z), _p(x), len(x), _p(y), len(y), _p(zp), len(zp), _p(xp), len(xp), _p(yp),
                         len(yp), _p(fp), int(bool(zlog)), _p(f))
    return f


def time_interp(f0, f1, t1, ta, tb):
    """interp_fields_ng.F90: values at t = ta..tb between f0 (t = 0) and f1
    (t = t1); returns (tb-ta+1,) + f0.shape."""
    f0, f1 = _v(f0), _v(f1)
    f = np.empty((tb - ta + 1,) + f0.shape)
    lib().tgx_o_time_interp(_p(f0), _p(f1), C.c_long(f0.size), int(t1), int(ta), int(tb), _p(f))
    return f
