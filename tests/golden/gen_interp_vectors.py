"""Writes tests/golden/interp_vectors.npz: inputs and outputs of the
reference's own interp1d / interp2d / interp3d (interp.F90:1-128), compiled
from the reference source by `make -C oracle ref` into oracle/_ref (run this in
a checkout where that library is built).  Data only: seeded coordinates and
fields shaped like a nest's (an outer grid with its ghost points, an inner
grid partly outside it, the same and different level sets), and the compiled
reference's results, so the oracle and the device interpolation can be pinned
where the reference tree is absent.

    python tests/golden/gen_interp_vectors.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_interp.so")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def cases(seed=20261203):
    """(name, z, x, y, zp, xp, yp, fp (nyp, nxp, nzp) C order, zlog)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(6):
        nzp, nxp, nyp = int(rng.integers(5, 12)), int(rng.integers(6, 14)), int(rng.integers(5, 12))
        dz = rng.uniform(0.2, 0.6)
        zp = -7. + dz * np.arange(nzp)
        xp = rng.uniform(10., 40.) + rng.uniform(1., 3.) * np.arange(nxp)
        yp = rng.uniform(-30., 10.) + rng.uniform(1., 3.) * np.arange(nyp)
        nz, nx, ny = int(rng.integers(3, 15)), int(rng.integers(2, 12)), int(rng.integers(1, 10))
        # inner points inside, on nodes and beyond both ends of the outer range
        z = np.sort(rng.uniform(zp[0] - 0.5, zp[-1] + 0.5, nz)) if c % 2 else zp[:nz % nzp + 1].copy()
        # the log form takes logarithms of the horizontally interpolated
        # values: keep its targets inside the horizontal range
        m = 0. if c % 3 != 1 else 1.
        x = np.sort(rng.uniform(xp[0] - m, xp[-1] + m, nx))
        y = np.sort(rng.uniform(yp[0] - m, yp[-1] + m, ny))
        if c == 2:
            x[0], y[-1] = xp[3], yp[2]
        fp = rng.uniform(0.5, 2.0, (nyp, nxp, nzp)) * np.exp(-0.3 * zp)[None, None, :]
        if c % 3 == 1:
            fp = fp - 1.1  # signed values for the linear-in-z form
        out.append(("case%d" % c, z, x, y, zp, xp, yp, fp, c % 3 != 1))
    return out


def main():
    ref = C.CDLL(REF)
    arrays = {}
    for name, z, x, y, zp, xp, yp, fp, zlog in cases():
        f3 = np.empty((len(y), len(x), len(z)))
        ref.tgx_ref_interp3d(_p(z), len(z), _p(x), len(x), _p(y), len(y), _p(zp), len(zp), _p(xp), len(xp),
                             _p(yp), len(yp), _p(np.ascontiguousarray(fp)), int(zlog), _p(f3))
        lev = np.ascontiguousarray(fp[:, :, 0])
        f2 = np.empty((len(y), len(x)))
        ref.tgx_ref_interp2d(_p(x), len(x), _p(y), len(y), _p(xp), len(xp), _p(yp), len(yp), _p(lev), _p(f2))
        col = np.ascontiguousarray(fp[0, 0, :])
        f1 = np.empty(len(z))
        ref.tgx_ref_interp1d(_p(z), len(z), _p(zp), len(zp), _p(col), _p(f1))
        for k, v in dict(z=z, x=x, y=y, zp=zp, xp=xp, yp=yp, fp=fp, zlog=np.int64(zlog), f3=f3, f2=f2,
                         f1=f1).items():
            arrays["%s_%s" % (name, k)] = v
    np.savez_compressed(os.path.join(HERE, "interp_vectors.npz"), **arrays)


if __name__ == "__main__":
    main()
