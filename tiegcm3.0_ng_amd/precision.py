"""Precision of the column solves: FP32 elimination with FP64 block pivots and
0, 1 or 2 FP64 residual refinement sweeps (tgx_blk_mixed_dev), against the
FP64 solution, on the very systems the fused step assembles for each field.

The rows come from the Thomas column kernels in capture mode (one column step
on seeded synthetic state).  A seeded fraction of the columns is made stiff:
in every row the diagonal of the diagonal block is rescaled to the row's
off-diagonal mass times (1 + m), m ~ U(1e-5, 1e-3), so the diagonal-dominance
margin is m.  The FP64 reference solution of every system is computed by the
caller (the oracle's trsolv_ng / blktri_ng / blktri3 in the tests and the
benchmark); this module builds the systems and runs the device solver.
"""
import ctypes as C

import numpy as np
import torch

from . import step as _step
from . import tgx as _tgx

_P = C.c_void_p
_SIGS = {
    "tgx_blk_mixed_dev": ([C.c_int, _P, _P, _P, _P, _P, C.c_long, C.c_int, C.c_int, _P, C.c_long, C.c_int, _P,
                           _P], C.c_int),
    "tgx_blk_mixed_work_bytes": ([C.c_int, C.c_int, C.c_long], C.c_ulonglong),
}

# field -> (block size, first capture plane, rows = nk - row_drop); the capture
# planes are documented with tgx_ng_grid.capture in include/tgx_step.h
FIELDS = {
    "TN": (1, 0, 1), "O+": (1, 4, 1),
    "NO": (1, 8, 0), "N4S": (1, 12, 0), "N2D": (1, 16, 0), "Ar": (1, 20, 0),
    "UN,VN": (2, 24, 1), "O2,O,He": (3, 38, 1),
}
COMPONENTS = {"UN,VN": ("UN", "VN"), "O2,O,He": ("O2", "O", "He")}


def _lib():
    L = _tgx.lib()
    for name, (args, res) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def capture_systems(grid, fields, device="cuda"):
    """One column step with row capture; returns (capture tensor (planes, ncol,
    nk) on the device, valid-column mask (ncol,), nk)."""
    ns = _step.NestState(grid, fields, device=device, scheme="thomas", capture=True)
    ns.step_serial()
    torch.cuda.synchronize()
    nlond, nlatd = ns.nlond, ns.nlatd
    valid = np.zeros((nlatd, nlond), dtype=np.uint8)
    valid[1:-1, 1:-1] = 1  # lateral-boundary columns carry no system
    return ns.capture, valid.reshape(-1), ns.nk


def split(cap, field):
    """The block planes (A, B, C, F) of `field`: views of shape (N*N, ncol, nk)
    (F: (N, ncol, nk))."""
    n, pl, _ = FIELDS[field]
    nn = n * n
    return (cap[pl:pl + nn], cap[pl + nn:pl + 2 * nn], cap[pl + 2 * nn:pl + 3 * nn],
            cap[pl + 3 * nn:pl + 3 * nn + n])


def make_stiff(A, B, Cc, valid, frac, seed, lo=1e-5, hi=1e-3):
    """Rescales the diagonal of B in a seeded `frac` of the valid columns to
    (off-diagonal row mass) * (1 + m), m ~ U(lo, hi).  In place; returns the
    stiff-column mask."""
    nn, ncol, nk = B.shape
    n = int(round(nn ** 0.5))
    rng = np.random.default_rng(seed)
    stiff = (rng.random(ncol) < frac) & (valid != 0)
    m = torch.tensor(rng.uniform(lo, hi, size=(ncol, nk)), dtype=B.dtype, device=B.device)
    sel = torch.tensor(stiff, device=B.device)
    for r in range(n):
        mass = torch.zeros((ncol, nk), dtype=B.dtype, device=B.device)
        for c in range(n):
            mass += A[r * n + c].abs() + Cc[r * n + c].abs()
            if c != r:
                mass += B[r * n + c].abs()
        d = B[r * n + r]
        sign = torch.where(d < 0, -torch.ones_like(d), torch.ones_like(d))
        newd = sign * mass * (1. + m)
        B[r * n + r] = torch.where(sel[:, None], newd, d)
    return stiff


def solve_mixed(field, A, B, Cc, F, valid_dev, nk, nsweep, stream=None):
    """The FP32 + FP64-refinement solution (N, ncol, nk) of one field's systems."""
    n, _, drop = FIELDS[field]
    ncol = B.shape[1]
    nrow = nk - drop
    X = torch.zeros((n, ncol, nk), dtype=torch.float64, device=B.device)
    L = _lib()
    wb = L.tgx_blk_mixed_work_bytes(n, nrow, ncol)
    work = torch.empty((wb + 7) // 8, dtype=torch.float64, device=B.device)
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    rc = L.tgx_blk_mixed_dev(n, A.data_ptr(), B.data_ptr(), Cc.data_ptr(), F.data_ptr(), X.data_ptr(),
                             ncol * nk, nk, nrow, valid_dev.data_ptr(), ncol, nsweep, work.data_ptr(), s)
    if rc != 0:
        raise _tgx.TgxError(_tgx.last_error())
    return X


def to_oracle_layout(A, F):
    """Capture planes -> the oracle's Fortran layouts: blocks (N,N,nk,ncol)
    column-major (element (r,c) at r + N c), F (N,nk,ncol); numpy, C order
    arrays whose flat order is the Fortran order."""
    nn, ncol, nk = A.shape
    n = int(round(nn ** 0.5))
    a = A.reshape(n, n, ncol, nk).permute(2, 3, 1, 0).contiguous().cpu().numpy()
    f = F.permute(1, 2, 0).contiguous().cpu().numpy()
    return a, f


def errors(x, xref, valid, stiff):
    """Max and mean |x - xref| over the field's valid columns, relative to the
    field's maximum |xref|; also over the stiff columns alone."""
    out = {}
    v = valid.astype(bool)
    for tag, mask in (("all", v), ("stiff", v & stiff)):
        if not mask.any():
            continue
        d = np.abs(x[:, mask, :] - xref[:, mask, :])
        scale = max(np.abs(xref[:, v, :]).max(), 1e-300)
        out[tag] = (float(d.max() / scale), float(d.mean() / scale))
    return out
