"""Nest levels coupled on the device: the recursion of advance_ng.F90:1-216
with its map_in, set_bndry and map_out (levels_inner_ng.F90:131-410) and the
time interpolation of interp_fields_ng.F90:17-63, around the fused column
step of every domain.

Level 0 is the outermost domain (its lateral data are given); level l >= 1
is a nest inside level l-1 that takes nstep_ng[l] sub-steps per step of
level l-1.  Per step of level l-1 (advance_ng):

- set_bndry: the outer domain's new state interpolated (interp3d, log-linear
  in z for the reference's zlog fields) to the nest's four lateral boundary
  lines, and map_in: the outer's lower-boundary fields interpolated (interp2d)
  to the nest, both stored as the nest's end-of-step anchors;
- before each nest sub-step s: the lateral data and lower-boundary fields at
  s, linear in time between the previous anchors (s = 0) and the new ones
  (s = nstep_ng), and the O+ lateral data of the sub-step's nstep_sub O+
  sub-cycles on the finer O+ time axis;
- after the nest's sub-steps (each followed by the next level's own
  advance): the anchors move to s = 0, and map_out interpolates the nest's
  state back onto the outer grid points inside the nest.

Every operation is a kernel or a device copy on the current stream.  The
fields coupled are those the column step reads or writes: set_bndry the
lateral fields of the step (TN, UN, VN, O2, O, He, O+, N2D, N4S, NO, Ar;
the reference's OMEGA is not read by the column step), map_in t_lbc, u_lbc,
v_lbc and flx_he (the electric fields and z_lbc are not read by it), map_out
the step's prognostic 3-D outputs (the reference's fmap less TI, TE, NE,
O2P, Z and OMEGA, which the column step does not write).
"""
import torch

from . import step as _step
from . import tgx as _tgx

# set_bndry's fields (fields_ng.F90:88-91) the column step reads, with the
# reference's zlog flag (fields_ng.F90:100-105)
BNDRY = (("tn", True), ("un", False), ("vn", False), ("o2", True), ("o1", True), ("he", True), ("op", True),
         ("n2d", True), ("n4s", True), ("no", True), ("ar", True))
# map_in's 2-D fields (fields_ng.F90 f2din) the column step reads
LBC2D = ("t_lbc", "u_lbc", "v_lbc", "flx_he")
# map_out's fields (fmap, fields_ng.F90:80-85) the column step writes
MAPOUT = tuple((n, z) for n, z in BNDRY) + tuple((n + "_nm", z) for n, z in BNDRY)


def mapout_box(gout, gin):
    """The outer grid points strictly inside the nest (levels_inner_ng.F90:
    265-276): (lonbeg, lonend, latbeg, latend), 0-based and inclusive, or None
    when the nest covers no outer point."""
    go, ga = gout["glon"], gout["glat"]
    gi, gj = gin["glon"], gin["glat"]
    lonbeg = next((i for i in range(len(go)) if go[i] > gi[0]), len(go))
    lonend = next((i for i in range(len(go) - 1, -1, -1) if go[i] < gi[-1]), -1)
    latbeg = next((j for j in range(len(ga)) if ga[j] > gj[0]), len(ga))
    latend = next((j for j in range(len(ga) - 1, -1, -1) if ga[j] < gj[-1]), -1)
    if lonbeg < lonend and latbeg < latend:
        return lonbeg, lonend, latbeg, latend
    return None


class CoupledNests:
    """domains: [(grid, fields), ...] outermost first; nstep_ng[l] sub-steps
    of level l per step of level l-1 (nstep_ng[0] steps of level 0 per run).
    comp_nsweep: None for the FP64 step, or 0..2 for the step with comp in
    its FP32 + FP64-refinement form (forked steps)."""

    def __init__(self, domains, nstep_ng, scheme="thomas", device="cuda", forked=True, init_from_outer=True,
                 comp_nsweep=None):
        self.states = [_step.NestState(g, F, device=device, scheme=scheme) for g, F in domains]
        for st in self.states:
            st.enable_feedback()
            st.comp_nsweep = comp_nsweep  # comp in its precision form (tgx_ng_step_mixed), forked steps
        self.grids = [g for g, _ in domains]
        self.nstep_ng = list(nstep_ng)
        self.forked = forked
        kw = dict(dtype=torch.float64, device=device)
        self.axes = [dict(z=torch.tensor(g["zpint"], **kw), x=torch.tensor(g["glon"], **kw),
                          y=torch.tensor(g["glat"], **kw)) for g in self.grids]
        for ax in self.axes:  # the two lateral boundary lines of each direction
            ax["x2"] = ax["x"][[0, -1]].contiguous()
            ax["y2"] = ax["y"][[0, -1]].contiguous()
        self.anchors = [None]
        self.boxes = [None]
        for l in range(1, len(self.states)):
            st = self.states[l]
            a = {}
            for n, _ in BNDRY:
                for side in ("lonb_", "latb_"):
                    one = st.inp[side + n] if n != "op" else st.inp[side + n][0]
                    a[side + n] = [torch.zeros_like(one), torch.zeros_like(one)]
            for n in LBC2D:
                a[n] = [torch.zeros_like(st.inp[n]), torch.zeros_like(st.inp[n])]
            self.anchors.append(a)
            self.boxes.append(mapout_box(self.grids[l - 1], self.grids[l]))
        # init_fields_ng.F90:128-175: every nest's prognostic fields and its
        # boundary data at s = 0 interpolated from the outermost domain's
        # state (or, without init_from_outer, the nests keep their own states
        # and the block-0 data come from their parents)
        for l in range(1, len(self.states)):
            src = 0 if init_from_outer else l - 1
            if init_from_outer:
                self.init_fields(l)
            self.set_bndry(l, 0, src)
            self.map_in(l, 0, src)

    # ---- the coupling operations -----------------------------------------------
    def _input(self, st, name):
        """The tensor the domain's next step reads for `name`."""
        if name in _step.PROGNOSTIC and getattr(st, "_phases", None) and st.phase == 1:
            return st.out[name]
        return st.inp[name]

    def init_fields(self, l):
        """init_fields_ng.F90:128-140: the nest's prognostic fields (the state
        its next step reads) interpolated from the outermost domain's state
        over the nest's whole grid, ghost points included (one launch)."""
        s0, a0, ai, st = self.states[0], self.axes[0], self.axes[l], self.states[l]
        _tgx.interp3d_batch_dev([s0.current(n) for n, _ in MAPOUT], [st.current(n) for n, _ in MAPOUT],
                                [z for _, z in MAPOUT], a0["z"], a0["x"], a0["y"], ai["z"], ai["x"], ai["y"],
                                st.nlatd * st.ld, 1, st.ld)

    def set_bndry(self, l, slot, src=None):
        """levels_inner_ng.F90:299-410: the outer state at the nest's lateral
        boundary lines into the nest's anchors[slot]; src is the source
        level (the parent, or the outermost domain at initialisation).  One
        launch for the longitude sides of every field (the two lines are the
        target x points, side-major in each anchor), one for the latitude
        sides."""
        src = l - 1 if src is None else src
        so, ao, ai = self.states[src], self.axes[src], self.axes[l]
        nlatd, ld = self.states[l].nlatd, self.states[l].ld
        a = self.anchors[l]
        srcs = [so.current(n) for n, _ in BNDRY]
        zl = [z for _, z in BNDRY]
        _tgx.interp3d_batch_dev(srcs, [a["lonb_" + n][slot] for n, _ in BNDRY], zl, ao["z"], ao["x"], ao["y"],
                                ai["z"], ai["x2"], ai["y"], 2 * nlatd, nlatd, 1)
        _tgx.interp3d_batch_dev(srcs, [a["latb_" + n][slot] for n, _ in BNDRY], zl, ao["z"], ao["x"], ao["y"],
                                ai["z"], ai["x"], ai["y2"], 2 * ld, 1, ld)

    def map_in(self, l, slot, src=None):
        """levels_inner_ng.F90:131-214 (the 2-D lower-boundary fields): the
        outer domain's fields of its current step into anchors[slot]."""
        src = l - 1 if src is None else src
        so, ao, ai = self.states[src], self.axes[src], self.axes[l]
        ld = self.states[l].ld
        srcs = [self._input(so, n) if n == "flx_he" else so.inp[n] for n in LBC2D]
        _tgx.interp2d_batch_dev(srcs, [self.anchors[l][n][slot] for n in LBC2D], ao["x"], ao["y"], ai["x"], ai["y"],
                                1, ld)

    def substep_inputs(self, l, s):
        """interp_fields_ng.F90:17-63: the nest's lateral data and lower-boundary
        fields of sub-step s (1..nstep_ng), and its O+ lateral data of the
        sub-step's O+ sub-cycles (one launch)."""
        st, a, t1 = self.states[l], self.anchors[l], self.nstep_ng[l]
        nsub = int(self.grids[l]["nstep_sub"])
        jobs = []
        for key, pair in a.items():
            if key.endswith("_op"):
                # op_lon_b / op_lat_b at (s-1)*nsub+1 .. s*nsub of 0..t1*nsub
                jobs.append((pair[0], pair[1], st.inp[key], t1 * nsub, (s - 1) * nsub + 1, s * nsub))
                continue
            dst = self._input(st, key) if key == "flx_he" else st.inp[key]
            if s == t1:
                jobs.append((None, pair[1], dst, 0, 0, 0))
            else:
                jobs.append((pair[0], pair[1], dst, t1, s, s))
        _tgx.time_interp_batch_dev(jobs)

    def map_out(self, l):
        """levels_inner_ng.F90:216-297: the nest's state onto the outer grid
        points inside it (one launch)."""
        box = self.boxes[l]
        if box is None:
            return
        lonbeg, lonend, latbeg, latend = box
        so, si, ao, ai = self.states[l - 1], self.states[l], self.axes[l - 1], self.axes[l]
        x = ao["x"][lonbeg:lonend + 1].contiguous()
        y = ao["y"][latbeg:latend + 1].contiguous()
        off = latbeg * so.ld + lonbeg
        _tgx.interp3d_batch_dev([si.current(n) for n, _ in MAPOUT], [(so.current(n), off) for n, _ in MAPOUT],
                                [z for _, z in MAPOUT], ai["z"], ai["x"], ai["y"], ao["z"], x, y, so.nlatd * so.ld,
                                1, so.ld)

    # ---- the recursion --------------------------------------------------------
    def _step(self, l):
        st = self.states[l]
        st.step() if self.forked else st.step_serial()

    def _advance(self, l):
        """advance_ng.F90:1-216 for nest level l >= 1 (after a step of l-1)."""
        self.set_bndry(l, 1)
        self.map_in(l, 1)
        for s in range(1, self.nstep_ng[l] + 1):
            self.substep_inputs(l, s)
            self._step(l)
            if l + 1 < len(self.states):
                self._advance(l + 1)
        _tgx.time_interp_batch_dev([(None, pair[1], pair[0], 0, 0, 0) for pair in self.anchors[l].values()])
        self.map_out(l)

    def run(self):
        """nstep_ng[0] steps of level 0, each followed by the nests' advance."""
        for _ in range(self.nstep_ng[0]):
            self._step(0)
            if len(self.states) > 1:
                self._advance(1)

    # ---- graph capture ------------------------------------------------------
    def _phases(self):
        return [st.phase for st in self.states]

    def _set_phases(self, ph):
        for st, p in zip(self.states, ph):
            st.phase = p

    def capture(self):
        """Records run() into HIP graphs (after one eager run that tunes AUTO
        schemes).  Every domain's state alternates between two buffer sets, so
        when a domain takes an odd number of steps per run the second graph
        starts from the other set; replay() alternates the graphs."""
        torch.cuda.synchronize()
        self.run()
        torch.cuda.synchronize()
        start = self._phases()
        self.graphs = []
        while True:
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self.run()
            self.graphs.append((graph, self._phases()))
            if self._phases() == start or len(self.graphs) == 2:
                break
        self._set_phases(start)  # nothing ran while capturing
        self._next = 0
        torch.cuda.synchronize()

    def replay(self):
        graph, end = self.graphs[self._next]
        graph.replay()
        self._set_phases(end)
        self._next = (self._next + 1) % len(self.graphs)
