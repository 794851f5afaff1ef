"""The nest coupling of advance_ng.F90 / levels_inner_ng.F90 /
interp_fields_ng.F90 composed from the oracle (test infrastructure): the same
recursion as tiegcm3.0_ng_amd/nest.py, with the oracle's interpolation
(interp.F90, pinned by tests/test_oracle_interp.py) and the oracle's column
step on states in the reference's Fortran layout (numpy, C order: 3-D
(nlatd, nlond, nk), lonb (nlatd, 2, nk), latb (2, nlond, nk), lonb_op
(nsub, nlatd, 2, nk), latb_op (nsub, 2, nlond, nk))."""
import numpy as np


class OracleNests:
    def __init__(self, oracle, pkg, domains, nstep_ng, init_from_outer=True):
        self.o, self.pkg = oracle, pkg
        self.nest = __import__(pkg.__name__ + ".nest", fromlist=["nest"])
        self.grids = [g for g, _ in domains]
        self.states = [{k: np.array(v, copy=True) for k, v in Fo.items()} for _, Fo in domains]
        self.nstep_ng = list(nstep_ng)
        self.anchors = [None]
        for l in range(1, len(domains)):
            a = {}
            for n, _ in self.nest.BNDRY:
                for side in ("lonb_", "latb_"):
                    one = self.states[l][side + n] if n != "op" else self.states[l][side + n][0]
                    a[side + n] = [np.zeros_like(one), np.zeros_like(one)]
            for n in self.nest.LBC2D:
                a[n] = [np.zeros_like(self.states[l][n]), np.zeros_like(self.states[l][n])]
            self.anchors.append(a)
        for l in range(1, len(domains)):
            src = 0 if init_from_outer else l - 1
            if init_from_outer:
                self.init_fields(l)
            self.set_bndry(l, 0, src)
            self.map_in(l, 0, src)

    def _ax(self, l):
        g = self.grids[l]
        return np.array(g["zpint"]), np.array(g["glon"]), np.array(g["glat"])

    def init_fields(self, l):
        z0, x0, y0 = self._ax(0)
        zi, xi, yi = self._ax(l)
        for n, zlog in self.nest.MAPOUT:
            self.states[l][n] = self.o.interp3d(zi, xi, yi, z0, x0, y0, self.states[0][n], zlog)

    def set_bndry(self, l, slot, src_level=None):
        src_level = l - 1 if src_level is None else src_level
        zo, xo, yo = self._ax(src_level)
        zi, xi, yi = self._ax(l)
        a = self.anchors[l]
        for n, zlog in self.nest.BNDRY:
            src = self.states[src_level][n]
            for side, xs in ((0, xi[:1]), (1, xi[-1:])):
                a["lonb_" + n][slot][:, side, :] = self.o.interp3d(zi, xs, yi, zo, xo, yo, src, zlog)[:, 0, :]
            for side, ys in ((0, yi[:1]), (1, yi[-1:])):
                a["latb_" + n][slot][side] = self.o.interp3d(zi, xi, ys, zo, xo, yo, src, zlog)[0]

    def map_in(self, l, slot, src_level=None):
        src_level = l - 1 if src_level is None else src_level
        _, xo, yo = self._ax(src_level)
        _, xi, yi = self._ax(l)
        for n in self.nest.LBC2D:
            self.anchors[l][n][slot][...] = self.o.interp2d(xi, yi, xo, yo, self.states[src_level][n])

    def substep_inputs(self, l, s):
        st, a, t1 = self.states[l], self.anchors[l], self.nstep_ng[l]
        nsub = int(self.grids[l]["nstep_sub"])
        for key, (a0, a1) in a.items():
            if key.endswith("_op"):
                st[key] = self.o.time_interp(a0, a1, t1 * nsub, (s - 1) * nsub + 1, s * nsub)
            elif s == t1:
                st[key] = a1.copy()
            else:
                st[key] = self.o.time_interp(a0, a1, t1, s, s)[0]

    def map_out(self, l):
        box = self.nest.mapout_box(self.grids[l - 1], self.grids[l])
        if box is None:
            return
        lonbeg, lonend, latbeg, latend = box
        zo, xo, yo = self._ax(l - 1)
        zi, xi, yi = self._ax(l)
        for n, zlog in self.nest.MAPOUT:
            out = self.o.interp3d(zo, xo[lonbeg:lonend + 1], yo[latbeg:latend + 1], zi, xi, yi, self.states[l][n],
                                  zlog)
            self.states[l - 1][n][latbeg:latend + 1, lonbeg:lonend + 1, :] = out

    def _step(self, l):
        res = self.pkg.step.run_oracle_step(self.o, self.grids[l], self.states[l])
        for n in self.pkg.step.PROGNOSTIC:
            self.states[l][n] = res[n]

    def _advance(self, l):
        self.set_bndry(l, 1)
        self.map_in(l, 1)
        for s in range(1, self.nstep_ng[l] + 1):
            self.substep_inputs(l, s)
            self._step(l)
            if l + 1 < len(self.states):
                self._advance(l + 1)
        for pair in self.anchors[l].values():
            pair[0][...] = pair[1]
        self.map_out(l)

    def run(self):
        for _ in range(self.nstep_ng[0]):
            self._step(0)
            if len(self.states) > 1:
                self._advance(1)
