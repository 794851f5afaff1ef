"""Host side of the fused column step over libtgx (include/tgx_step.h).

`NestState` holds one nest domain's model state on the GPU in the library's
level-plane layout (longitude fastest, row pitch `ld`), the outputs and the
workspace, and calls the C-ABI entry points that mirror the reference's
routines: oplus_ng (with its nstep_sub sub-cycles), the minor species,
dt_ng, duv_ng and comp_o2o_ng + comp_ng, or the whole step in the
reference's order (dynamics_ng.F90:99-233).
"""
import ctypes as C

import torch

from . import synth as _synth
from . import tgx as _tgx

_P = C.c_void_p

GRID_FIELDS = [
    ("nk", C.c_int), ("nlon", C.c_int), ("nlat", C.c_int), ("ld", C.c_int),
    ("nstep_sub", C.c_int), ("calc_helium", C.c_int), ("scheme", C.c_int),
    ("dz", C.c_double), ("dlamda", C.c_double), ("dphi", C.c_double), ("dtx2inv", C.c_double),
    ("shapiro", C.c_double), ("expzmid", C.c_double), ("expzmid_inv", C.c_double),
    ("colfac", C.c_double), ("opdiffcap", C.c_double), ("difhor", C.c_double),
    ("pshelb", C.c_double), ("ar_glbm1", C.c_double), ("ar_glbm2", C.c_double),
    ("b", C.c_double * 9), ("fb", C.c_double * 3),
    ("expz", _P), ("difk", _P), ("dift", _P), ("xmue", _P), ("zpint", _P), ("zpmid", _P),
    ("racs", _P), ("cs", _P), ("tanphi", _P), ("cor", _P), ("glat", _P),
    ("capture", _P),
]
CAPTURE_PLANES = 68
IN_3D = ("tn tn_nm un vn un_nm vn_nm o2 o2_nm o1 o1_nm he he_nm "
         "op op_nm n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm ti te ne o2p w z "
         "mbar barm xnmbar xnmbari scht schti n2 w_upd "
         "cp kt km hdt hdu hdv hdo2 hdo1 hdhe qji_tn cool_imp cool_exp qtotal rkm12 "
         "ui vi wi lxx lyy lxy lyx qop2p qop2d qop qo2p qtef rj rk1 rk2 rk3 rk19 rk20 rk25 "
         "beta1 beta3 beta5 beta8 beta9 beta9n beta17 ra1 ra2 ra3 n2p nplus nop").split()
IN_2D = "t_lbc u_lbc v_lbc tlbc ulbc vlbc flx_he bx by bz bmod2 dipmag sndec csdec chi rlatm".split()
LATERAL = "tn un vn o2 o1 he n2d n4s no ar op".split()
OUT_3D = ("tn tn_nm un un_nm vn vn_nm n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm "
          "op op_nm xiop2p xiop2d Fe Fn o2 o2_nm o1 o1_nm he he_nm").split()
OUT_2D = "tlbc tlbc_nm ulbc vlbc ulbc_nm vlbc_nm flx_he".split()
# The model state a step advances: its outputs (itc) are the next step's
# inputs (itp) under the same names (the time levels swap between steps,
# dynamics_ng.F90 / advance_ng.F90).
PROGNOSTIC = ("tn tn_nm un un_nm vn vn_nm o2 o2_nm o1 o1_nm he he_nm op op_nm "
              "n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm tlbc ulbc vlbc flx_he").split()


class NgGrid(C.Structure):
    _fields_ = GRID_FIELDS


class NgFields(C.Structure):
    _fields_ = ([(n, _P) for n in IN_3D] + [(n, _P) for n in IN_2D]
                + sum([[("lonb_" + n, _P), ("latb_" + n, _P)] for n in LATERAL], []))


class NgOut(C.Structure):
    _fields_ = [(n, _P) for n in OUT_3D + OUT_2D]


_SIGS = {
    "tgx_ng_workspace_bytes": ([_P], C.c_ulonglong),
    "tgx_ng_oplus": ([_P] * 5, C.c_int),
    "tgx_ng_minors": ([_P] * 5, C.c_int),
    "tgx_ng_dt": ([_P] * 5, C.c_int),
    "tgx_ng_duv": ([_P] * 5, C.c_int),
    "tgx_ng_comp": ([_P] * 5, C.c_int),
    "tgx_ng_comp_mixed": ([_P] * 5 + [C.c_int], C.c_int),
    "tgx_ng_step_mixed": ([_P] * 5 + [C.c_int], C.c_int),
    "tgx_ng_step": ([_P] * 5, C.c_int),
    "tgx_ng_step_serial": ([_P] * 5, C.c_int),
    "tgx_ng_profile_begin": ([C.c_int], C.c_int),
    "tgx_ng_autotune": ([_P] * 5, C.c_int),
    "tgx_ng_scheme_choice": ([_P, C.c_int], C.c_int),
    "tgx_ng_wave_columns_choice": ([_P, C.c_int], C.c_int),
    "tgx_ng_force_wave_columns": ([C.c_int], None),
    "tgx_ng_set_tile_dma": ([C.c_int], None),
    "tgx_ng_profile_end": ([_P, _P], C.c_int),
    "tgx_glb_workspace_bytes": ([_P], C.c_ulonglong),
    "tgx_glb_oplus": ([_P] * 5, C.c_int),
    "tgx_glb_minors": ([_P] * 5, C.c_int),
    "tgx_glb_dt": ([_P] * 5, C.c_int),
    "tgx_glb_duv": ([_P] * 5, C.c_int),
    "tgx_glb_comp": ([_P] * 5, C.c_int),
    "tgx_glb_step": ([_P] * 5, C.c_int),
    "tgx_glb_step_serial": ([_P] * 5, C.c_int),
    "tgx_glb_halos": ([_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _P], C.c_int),
}

ROUTINES = ("oplus", "minors", "dt", "duv", "comp")
SCHEMES = {"auto": 0, "thomas": 1, "pcr": 2}


def profile_begin(max_calls):
    """Start recording a hipEvent pair around every routine call (on the
    stream the routine runs on), for at most max_calls calls per routine."""
    if _lib().tgx_ng_profile_begin(int(max_calls)) != 0:
        raise _tgx.TgxError(_tgx.last_error())


def profile_end():
    """Stop recording; returns ({routine: (summed ms, calls)}, {routine: (summed
    ms, calls) of its column kernel})."""
    ms = (C.c_double * 10)()
    n = (C.c_int * 10)()
    if _lib().tgx_ng_profile_end(C.cast(ms, _P), C.cast(n, _P)) != 0:
        raise _tgx.TgxError(_tgx.last_error())
    return ({r: (ms[i], n[i]) for i, r in enumerate(ROUTINES)},
            {r: (ms[5 + i], n[5 + i]) for i, r in enumerate(ROUTINES)})


def _lib():
    L = _tgx.lib()
    for name, (args, res) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def pitch(nlond, align=16):
    return (nlond + align - 1) // align * align


class NestState:
    """One nest domain on the GPU.

    grid:   dict from synth.make_grid
    fields: dict of float64 tensors in logical layout (synth.make_state),
            on the CPU or already on the GPU.
    scheme: column solve, "thomas" (the reference's operation order), "pcr"
            (one wavefront per column) or "auto" (timed once per shape).
    """

    def __init__(self, grid, fields, device="cuda", scheme="thomas", capture=False):
        self.grid = grid
        nk, nlat, nlon = grid["nk"], grid["nlat"], grid["nlon"]
        self.nk, self.nlatd, self.nlond = nk, nlat + 4, nlon + 4
        self.ld = pitch(self.nlond)
        self.dev = device
        kw = dict(dtype=torch.float64, device=device)
        self._keep = []

        def lvl(name):
            t = torch.tensor(grid[name], **kw)
            self._keep.append(t)
            return t.data_ptr()

        g = NgGrid()
        g.nk, g.nlon, g.nlat, g.ld = nk, nlon, nlat, self.ld
        g.nstep_sub, g.calc_helium = int(grid["nstep_sub"]), int(grid["calc_helium"])
        g.scheme = SCHEMES[scheme]
        for n in ("dz", "dlamda", "dphi", "dtx2inv", "shapiro", "expzmid", "expzmid_inv",
                  "colfac", "opdiffcap", "difhor", "pshelb"):
            setattr(g, n, float(grid[n]))
        g.ar_glbm1, g.ar_glbm2 = float(grid["ar_glbm"][0]), float(grid["ar_glbm"][1])
        bb = [float(x) for row in grid["b"] for x in row]
        for i in range(9):
            g.b[i] = bb[i]
        for i in range(3):
            g.fb[i] = float(grid["fb"][i])
        for n in ("expz", "difk", "dift", "xmue", "zpint", "zpmid", "racs", "cs", "tanphi", "cor", "glat"):
            setattr(g, n, lvl(n))
        self.g = g

        self.inp = {}
        f = NgFields()
        for n in IN_3D:
            self.inp[n] = self._put3(fields[n])
        for n in IN_2D:
            self.inp[n] = self._put2(fields[n])
        for n in LATERAL:
            lonb, latb = fields["lonb_" + n], fields["latb_" + n]
            self.inp["lonb_" + n] = lonb.to(**kw).contiguous()
            self.inp["latb_" + n] = self._padlast(latb.to(**kw))
        for n, t in self.inp.items():
            setattr(f, n, t.data_ptr())
        self.f = f

        self.out = {n: torch.zeros((nk, self.nlatd, self.ld), **kw) for n in OUT_3D}
        self.out.update({n: torch.zeros((self.nlatd, self.ld), **kw) for n in OUT_2D})
        o = NgOut()
        for n, t in self.out.items():
            setattr(o, n, t.data_ptr())
        self.o = o
        nbytes = _lib().tgx_ng_workspace_bytes(C.byref(self.g))
        self.work = torch.empty((nbytes + 7) // 8, **kw)
        self.capture = None
        if capture:  # the assembled column systems, written by the Thomas kernels
            ncol = self.nlond * self.nlatd
            self.capture = torch.zeros((CAPTURE_PLANES, ncol, nk), **kw)
            self.g.capture = self.capture.data_ptr()

    # ---- time stepping ------------------------------------------------------------
    def enable_feedback(self):
        """Makes successive steps advance the state: step n reads the
        prognostic fields step n-1 wrote.  Two buffer sets alternate (the
        inputs given at construction and the outputs), so no copies are made:
        phase 0 reads set A and writes set B, phase 1 reads B and writes A.
        current(name) returns the latest value of a prognostic field."""
        if getattr(self, "_phases", None):
            return
        f1 = NgFields()
        o1 = NgOut()
        for n, _ in NgFields._fields_:
            setattr(f1, n, getattr(self.f, n))
        for n, _ in NgOut._fields_:
            setattr(o1, n, getattr(self.o, n))
        for n in PROGNOSTIC:
            setattr(f1, n, self.out[n].data_ptr())
            setattr(o1, n, self.inp[n].data_ptr())
        self._phases = [(self.f, self.o), (f1, o1)]
        self.phase = 0

    def current(self, name):
        """Latest value of a field in the padded device layout."""
        if getattr(self, "_phases", None) and name in PROGNOSTIC and self.phase == 0:
            return self.inp[name]
        return self.out[name]

    # ---- layout helpers ------------------------------------------------------
    def _padlast(self, t):
        shp = t.shape[:-1] + (self.ld,)
        out = torch.zeros(shp, dtype=torch.float64, device=self.dev)
        out[..., :t.shape[-1]] = t
        return out

    def _put3(self, t):
        return self._padlast(t.to(dtype=torch.float64, device=self.dev))

    def _put2(self, t):
        return self._padlast(t.to(dtype=torch.float64, device=self.dev))

    def fortran(self, name):
        """An output in the reference's Fortran layout, numpy (nlatd, nlond, nk)
        or (nlatd, nlond); with feedback, the latest value of a prognostic field."""
        t = self.current(name)[..., :self.nlond]
        if t.dim() == 3:
            t = t.permute(1, 2, 0)
        return t.contiguous().cpu().numpy()

    # ---- entry points -----------------------------------------------------------
    _STEPS = ("tgx_ng_step", "tgx_ng_step_serial", "tgx_ng_step_mixed")

    def _call(self, name, stream=None, extra=()):
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        f, o = (self.f, self.o)
        phases = getattr(self, "_phases", None)
        if phases and name in self._STEPS:
            f, o = phases[self.phase]
        rc = getattr(_lib(), name)(C.byref(self.g), C.byref(f), C.byref(o),
                                   C.c_void_p(self.work.data_ptr()), C.c_void_p(s), *extra)
        if rc != 0:
            raise _tgx.TgxError(_tgx.last_error())
        if phases and name in self._STEPS:
            self.phase ^= 1

    def oplus(self, stream=None):
        self._call("tgx_ng_oplus", stream)

    def minors(self, stream=None):
        self._call("tgx_ng_minors", stream)

    def dt(self, stream=None):
        self._call("tgx_ng_dt", stream)

    def duv(self, stream=None):
        self._call("tgx_ng_duv", stream)

    def comp(self, stream=None):
        self._call("tgx_ng_comp", stream)

    def comp_mixed(self, nsweep, stream=None):
        """comp's precision form: FP32 assembly and elimination with FP64 block
        pivots and nsweep (0..2) FP64-residual refinement sweeps."""
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        rc = _lib().tgx_ng_comp_mixed(C.byref(self.g), C.byref(self.f), C.byref(self.o),
                                      C.c_void_p(self.work.data_ptr()), C.c_void_p(s), C.c_int(int(nsweep)))
        if rc != 0:
            raise _tgx.TgxError(_tgx.last_error())

    # comp's refinement sweeps in step() (None: the FP64 routine; 0..2: its
    # precision form, tgx_ng_step_mixed)
    comp_nsweep = None

    def step(self, stream=None):
        if self.comp_nsweep is None:
            self._call("tgx_ng_step", stream)
        else:
            self._call("tgx_ng_step_mixed", stream, (C.c_int(int(self.comp_nsweep)),))

    def step_serial(self, stream=None):
        self._call("tgx_ng_step_serial", stream)

    def autotune(self, stream=None):
        self._call("tgx_ng_autotune", stream)

    def schemes(self):
        """The column-solve scheme each routine runs with ("thomas" / "pcr")."""
        names = {1: "thomas", 2: "pcr"}
        return {r: names.get(_lib().tgx_ng_scheme_choice(C.byref(self.g), n), "?")
                for n, r in enumerate(ROUTINES)}

    def wave_columns(self):
        """Columns (wavefronts) per block of each routine that runs the
        wavefront-per-column scheme."""
        sch = self.schemes()
        return {r: int(_lib().tgx_ng_wave_columns_choice(C.byref(self.g), n))
                for n, r in enumerate(ROUTINES) if sch[r] == "pcr"}


class NestedStep:
    """One global step of the global grid and its nests in the reference's
    recursion (advance_ng.F90:57-199): nest level i runs nstep_ng[i] sub-steps
    per step of level i-1, so level i runs prod(nstep_ng[:i+1]) sub-steps per
    global step.  Each domain's column step is the fused step on its own
    device state, and each sub-step advances that state (feedback: the
    prognostic fields one sub-step writes are the next one's inputs; the
    other inputs and the lateral data stay those given, since the explicit
    physics and the map_in/set_bndry coupling are outside the column step).

    run() launches the sequence eagerly on the current stream; capture() records
    it into a HIP graph and replay() launches the whole outer step as one
    graph.  AUTO schemes are tuned by a first eager run before capture.  With
    feedback, a domain taking an odd number of sub-steps per outer step ends
    it on the other buffer set, so capture() records two graphs, one for each
    starting buffer set, and replay() alternates them.
    forked: each domain step is tgx_ng_step (routines on forked streams,
    joined before the next sub-step) rather than tgx_ng_step_serial.
    """

    def __init__(self, domains, nstep_ng, scheme="thomas", device="cuda", forked=True, feedback=True):
        self.states = [NestState(g, F, device=device, scheme=scheme) for g, F in domains]
        if feedback:
            for st in self.states:
                st.enable_feedback()
        self.nstep_ng = list(nstep_ng)
        self.forked = forked
        self.graph = None
        self.order = []  # domain index of each sub-step, in launch order

        def visit(level):
            for _ in range(self.nstep_ng[level]):
                self.order.append(level)
                if level + 1 < len(self.states):
                    visit(level + 1)
        visit(0)

    def substeps(self):
        return {i: self.order.count(i) for i in range(len(self.states))}

    def run(self):
        for i in self.order:
            if self.forked:
                self.states[i].step()
            else:
                self.states[i].step_serial()

    def _phases(self):
        return [getattr(st, "phase", 0) for st in self.states]

    def _set_phases(self, ph):
        for st, p in zip(self.states, ph):
            if getattr(st, "_phases", None):
                st.phase = p

    def capture(self):
        torch.cuda.synchronize()
        self.run()  # tunes AUTO schemes and settles allocations before capture
        torch.cuda.synchronize()
        start = self._phases()
        odd = any(getattr(st, "_phases", None) and self.order.count(i) % 2
                  for i, st in enumerate(self.states))
        self.graphs = []
        for _ in range(2 if odd else 1):   # nothing runs while capturing
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self.run()
            self.graphs.append((graph, self._phases()))
        self._set_phases(start)             # the data is still where the warm-up left it
        self._next = 0
        self.graph = self.graphs[0][0]
        torch.cuda.synchronize()

    def replay(self):
        graph, end = self.graphs[self._next]
        graph.replay()
        self._set_phases(end)
        self._next = (self._next + 1) % len(self.graphs)


# Algorithmic HBM traffic of each routine: the distinct 3-D fields its fused
# kernels must read and write once (8 bytes per column-level each).  2-D and
# 1-D inputs are negligible and not counted.
ALG_FIELDS = {
    "oplus": (  # oplus_ng.F90:35-75 inputs, :14-16 outputs (per sub-step)
        "tn te ti o2 o1 n2 he n2d ne un vn w ui vi wi xnmbar scht op op_nm qop2p qop2d qop "
        "rk1 rk2 rk19 rk20 rk25".split(),
        "op op_nm xiop2p xiop2d Fe Fn".split()),
    "minors": (  # minor_ng.F90:42-50 + comp_{n2d,n4s,no,ar}_ng.F90 inputs
        "tn o2 o1 he mbar barm xnmbar w_upd un vn n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm "
        "qtef ne rk3 n2p ra1 nop ra3 op beta8 beta5 rk2 n2 nplus beta1 beta3 beta17 o2p beta9 beta9n".split(),
        "n2d n2d_nm n4s n4s_nm no no_nm ar ar_nm".split()),
    "dt": (  # dt_ng.F90:27-43
        "tn tn_nm un vn o1 mbar xnmbar scht schti cp kt km hdt qji_tn cool_imp cool_exp w_upd "
        "qtotal rkm12".split(),
        "tn tn_nm".split()),
    "duv": (  # duv_ng.F90:33-55 (tn_upd = this step's TN)
        "tn tn_upd tn_nm un vn un_nm vn_nm w_upd mbar scht schti z hdu hdv ui vi lxx lyy lxy lyx "
        "km Fe Fn".split(),
        "un un_nm vn vn_nm".split()),
    "comp": (  # comp_ng.F90:47-59 + comp_o2o_ng.F90:17-40
        "tn o2 o2_nm o1 o1_nm he he_nm w mbar hdo2 hdo1 hdhe xnmbar un vn op no n4s n2d o2p ne "
        "n2p nplus nop rj qo2p qop ra1 ra2 beta3 beta8 rk1 beta1 rkm12 rk3".split(),
        "o2 o2_nm o1 o1_nm he he_nm".split()),
}
# unknowns (column-level solves) per column and level of each routine
UNKNOWNS = {"oplus": 1, "minors": 4, "dt": 1, "duv": 2, "comp": 3}


def alg_bytes_per_column_level(routine):
    ins, outs = ALG_FIELDS[routine]
    return 8 * (len(set(ins)) + len(set(outs)))


# The column kernel of each routine: its name per scheme and the distinct
# planes it must read and write once per launch (8 bytes per column-level
# each).  Chemistry terms and (for the species) the horizontal-stencil
# right-hand-side parts arrive as planes from the routine's level-parallel
# kernels; the dt, duv and O+ Thomas kernels evaluate their stencils (and
# duv its vertical integral) themselves from tiles of the fields, while their
# PCR kernels read the stencil planes (PCR_INS).  Thomas ratio / gamma planes
# are scratch, not algorithmic.
COLUMN_KERNELS = {
    "comp": (("comp_kernel", "comp_wave_kernel"),  # rhs_*: the prelude's smoothed/advected part
             "o2 o1 he mbar tn w fs01 fs02 fs11 fs12 rhs_o2 rhs_o1 rhs_he "
             "hdo2 hdo1 hdhe o2_nm o1_nm he_nm xnmbar".split(),
             "o2 o1 he o2_nm o1_nm he_nm".split()),
    "minors": (("minor_kernel<2, 3>", "minor_wave_kernel<2, 3>"),  # N2D + Ar
               "o1 o2 he mbar tn barm w_upd rhs_n2d rhs_ar sloss_n2d "
               "s0prod_n2d n2d n2d_nm ar ar_nm".split(),
               "n2d n2d_nm ar ar_nm".split()),
    "dt": (("dt_tile_kernel", "dt_wave_kernel"),
           "tn w_upd cp km un vn qtotal schti kt xnmbar o1 rkm12 mbar scht "
           "hdt qji_tn cool_imp cool_exp tn_nm".split(),
           "tn tn_nm".split()),
    "duv": (("duv_tile_kernel", "duv_wave_stage2"),  # duv_ng.F90:33-55 (tn_upd = this step's TN)
            "w_upd lxx lyy lxy lyx ui vi km schti scht Fe Fn hdu hdv un vn un_nm vn_nm "
            "z tn tn_nm tn_upd mbar".split(),
            "un vn un_nm vn_nm".split()),
    "oplus": (("oplus_tile_kernel", "oplus_wave_stage3"),  # per O+ sub-step
              "scht ti te dj bdotu wi bdopj diffj bvel opbm2 djbz optm1 ui vi "
              "oploss opprod op".split(),
              "op op_nm".split()),
}
PCR_INS = {
    "dt": "tn w_upd cp km un vn qtotal schti kt hadv_tn smooth_tn_nm xnmbar o1 rkm12 mbar scht "
          "hdt qji_tn cool_imp cool_exp tn_nm".split(),
    "duv": "w_upd lxx lyy lxy lyx ui vi divx_ztmp divy_z km schti scht Fe Fn unms vnms advu advv "
           "hdu hdv un un_nm vn_nm vn".split(),
    "oplus": "scht ti te dj bdotu bdopj wi bd_diff bd_diffj djbz vdotn bd_bvel opts2 bd_djbz "
             "oploss opprod op optm1".split(),
}


def column_kernel_bytes_per_column_level(routine, scheme="thomas"):
    _, ins, outs = COLUMN_KERNELS[routine]
    if scheme == "pcr":
        ins = PCR_INS.get(routine, ins)
    return 8 * (len(set(ins)) + len(set(outs)))


def run_oracle_steps(oracle, grid, Fo, nsteps):
    """The oracle's column step taken nsteps times, each step's prognostic
    outputs (PROGNOSTIC) the next step's inputs; returns the last step's
    outputs."""
    Fo = dict(Fo)
    res = None
    for _ in range(nsteps):
        res = run_oracle_step(oracle, grid, Fo)
        for n in PROGNOSTIC:
            Fo[n] = res[n]
    return res


def run_oracle_step(oracle, grid, Fo):
    """The oracle's column step in the reference's order (dynamics_ng.F90)."""
    st = oracle.OracleState(grid, Fo)
    res = {}
    opl = st.oplus_subcycled(Fo["lonb_op"], Fo["latb_op"])
    res.update(opl)
    res.update(st.minors(Fo["tlbc"]))
    d = st.dt_ng(Fo["tlbc"])
    res.update(dict(tn=d["tn"], tn_nm=d["tn_nm"], tlbc=d["tlbc"], tlbc_nm=d["tlbc_nm"]))
    uv = st.duv_ng(d["tn"], opl["Fe"], opl["Fn"], d["tlbc_nm"], Fo["ulbc"], Fo["vlbc"])
    res.update(uv)
    fs = st.comp_o2o_ng()
    c = st.comp_ng(fs, d["tlbc"], Fo["flx_he"])
    res.update(c)
    return res


# Top levels the global forms set to spval (TN, UN, VN) or zero (O+)
# (dt.F:407-408, duv.F:385-386, :512-513, oplus.F:922-923).
_FIXED_TOP = ("tn", "tn_nm", "un", "un_nm", "vn", "vn_nm", "op", "op_nm")


def smoke_check(oracle, name="tiny", tol=1e-12):
    """One small fused column step on cuda:0 against the oracle, as a nest
    domain (tgx_ng_step) and as a global grid (tgx_glb_step, with its ring
    filters), every output field within tol of the oracle relative to the
    field's largest magnitude (on the global grid: the interior latitudes,
    below the spval / zero top, which must match exactly)."""
    import numpy as np
    from . import synth
    worst = 0.
    for glb in (False, True):
        g = synth.make_grid(name, nstep_sub=2, glb=glb)
        F = synth.make_global_state(g) if glb else synth.make_state(g)
        Fo = synth.to_fortran_layout(F)
        ref = run_oracle_step(oracle, g, Fo)
        ns = GlobalState(g, F) if glb else NestState(g, F)
        ns.step()
        torch.cuda.synchronize()
        for n in OUT_3D:
            got, exp = ns.fortran(n), ref[n]
            if glb:
                got, exp = got[2:g["nlat"] + 2], exp[2:g["nlat"] + 2]
                if n in _FIXED_TOP:
                    assert np.array_equal(got[..., -1], exp[..., -1]), n
                    got, exp = got[..., :-1], exp[..., :-1]
            err = float(np.abs(got - exp).max() / max(np.abs(exp).max(), 1e-300))
            worst = max(worst, err)
            assert err <= tol, ("global" if glb else "nest", n, err)
    print("smoke ok: fused column step on %s as a nest domain and as a global grid, worst field-relative "
          "error %.3e" % (name, worst))
    return worst


class GlbGrid(C.Structure):
    _fields_ = [("g", NgGrid), ("nlat_filter", C.c_int), ("chunk", _P)]




class GlobalState(NestState):
    """The global grid on the GPU (include/tgx_global.h): the same state,
    outputs and workspace as a nest domain, run through the global forms of
    the routines (dt.F, duv.F, minor.F, oplus.F, comp_o2o.F + comp.F) with
    their ring filters.  grid: synth.make_grid(..., glb=True); fields: a
    state whose halos are filled (synth.make_global_state).  The global forms
    run the Thomas scheme."""

    _NAMES = {"tgx_ng_oplus": "tgx_glb_oplus", "tgx_ng_minors": "tgx_glb_minors", "tgx_ng_dt": "tgx_glb_dt",
              "tgx_ng_duv": "tgx_glb_duv", "tgx_ng_comp": "tgx_glb_comp", "tgx_ng_step": "tgx_glb_step",
              "tgx_ng_step_serial": "tgx_glb_step_serial"}

    def __init__(self, grid, fields, device="cuda", capture=False):
        assert grid.get("glb"), "GlobalState needs a grid made with glb=True"
        super().__init__(grid, fields, device=device, scheme="thomas", capture=capture)
        self.chunk = (C.c_int * max(1, len(grid["chunk"])))(*[int(c) for c in grid["chunk"]])
        self.gg = GlbGrid()
        self.gg.g = self.g
        self.gg.nlat_filter = len(grid["chunk"])
        self.gg.chunk = C.cast(self.chunk, _P)
        nbytes = _lib().tgx_glb_workspace_bytes(C.byref(self.gg))  # the nest form's + the ring filters' table
        if nbytes > 8 * self.work.numel():
            self.work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=self.work.device)

    def _call(self, name, stream=None):
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        f, o = (self.f, self.o)
        phases = getattr(self, "_phases", None)
        stepping = name in ("tgx_ng_step", "tgx_ng_step_serial")
        if phases and stepping:
            f, o = phases[self.phase]
        self.gg.g = self.g
        rc = getattr(_lib(), self._NAMES[name])(C.byref(self.gg), C.byref(f), C.byref(o),
                                                C.c_void_p(self.work.data_ptr()), C.c_void_p(s))
        if rc != 0:
            raise _tgx.TgxError(_tgx.last_error())
        if phases and stepping:
            self.phase ^= 1

    def comp_mixed(self, nsweep, stream=None):
        raise NotImplementedError("the precision form of comp is built for nest domains only")

    def autotune(self, stream=None):
        pass  # the global forms run the Thomas scheme

    def schemes(self):
        return {r: "thomas" for r in ROUTINES}

    def fill_halos(self, names=PROGNOSTIC, stream=None):
        """tgx_glb_halos on the current value of each field: its periodic
        points and the rows across the poles (as the model does before a
        step, mk_polelat.F), sign-flipped for synth.POLE_ODD."""
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        for n in names:
            t = self.current(n)
            nk = t.shape[0] if t.dim() == 3 else 1
            sign = -1. if n in _synth.POLE_ODD else 1.
            rc = _lib().tgx_glb_halos(C.c_void_p(t.data_ptr()), nk, self.grid["nlon"], self.grid["nlat"],
                                      self.ld, C.c_double(sign), C.c_void_p(s))
            if rc != 0:
                raise _tgx.TgxError(_tgx.last_error())
