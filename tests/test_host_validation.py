"""Argument checks of the device entry points that run before any device work
(CPU only: each call must return TGX_ERR_ARG with a message and launch
nothing, so these run without a GPU): the ring filter and its fused global
form, the batched nest-coupling launches, the global step's grid checks and
the mixed-precision step."""
import ctypes as C
import os

import pytest

from helpers import ROOT

ERR_ARG = 1


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(os.path.join(ROOT, "tiegcm3.0_ng_amd", "libtgx.so"))
    L.tgx_last_error.restype = C.c_char_p
    return L


def _err(lib, rc, *words):
    assert rc == ERR_ARG, rc
    msg = lib.tgx_last_error().decode()
    for w in words:
        assert w in msg, msg
    lib.tgx_clear_error()


def test_ringfilter_dev_rejects_bad_shapes(lib):
    f = C.c_void_p(0x1000)  # never dereferenced: the checks come first
    chunk = (C.c_int * 2)(3, 9)
    fn = lib.tgx_ringfilter_dev
    fn.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    _err(lib, fn(f, 0, 72, 36, 80, 3, 0, 0.1, 2, chunk, None), "tgx_ringfilter_dev")      # nk = 0
    _err(lib, fn(f, 29, 72, 36, 70, 3, 0, 0.1, 2, chunk, None), "tgx_ringfilter_dev")     # ld < nlon + 4
    _err(lib, fn(f, 29, 72, 36, 80, 3, 0, 0.1, 65, chunk, None), "tgx_ringfilter_dev")    # too many rows
    _err(lib, fn(f, 29, 72, 36, 80, 9, 0, 0.1, 2, chunk, None), "order")                 # order > 8
    _err(lib, fn(f, 29, 72, 36, 80, 3, 0, 0.1, 2, None, None), "tgx_ringfilter_dev")      # no chunk table


def test_ringfilter_post_rejects_bad_arguments(lib):
    fn = lib.tgx_ringfilter_post_dev
    P4 = C.c_void_p * 4
    p = P4(0x1000, 0x2000, 0x3000, 0x4000)
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 8 + \
        [C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    chunk = (C.c_int * 2)(3, 9)
    tab = C.c_void_p(0x5000)
    args = lambda nf, kind, nlev, nk, gt: (nf, p, p, p, p, kind, nlev, nk, 72, 36, 80, 3, 0, 0.1, 2, chunk, gt, 3,
                                           None)
    _err(lib, fn(*args(5, 1, 28, 29, tab)), "ring filter + post")   # five fields
    _err(lib, fn(*args(1, 7, 28, 29, tab)), "ring filter + post")   # unknown post pass
    _err(lib, fn(*args(1, 1, 27, 29, tab)), "ring filter + post")   # a post pass below nk - 1 levels
    _err(lib, fn(*args(1, 1, 28, 29, None)), "ring filter + post")  # no table


def test_coupling_batches_reject_bad_jobs(lib):
    f3 = lib.tgx_ng_interp3d_batch_dev
    f3.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 4 + [C.c_void_p] * 3 + \
        [C.c_int] * 3 + [C.c_long] * 3 + [C.c_void_p] * 4
    P = (C.c_void_p * 33)(*([0x1000] * 33))
    Z = (C.c_int * 33)()
    ax = C.c_void_p(0x2000)
    _err(lib, f3(33, P, P, Z, 9, 13, 11, 16, ax, ax, ax, 7, 6, 5, 30, 1, 6, ax, ax, ax, None), "33 jobs")
    _err(lib, f3(2, P, P, Z, 1, 13, 11, 16, ax, ax, ax, 7, 6, 5, 30, 1, 6, ax, ax, ax, None), "zp")  # one level
    assert f3(0, None, None, None, 9, 13, 11, 16, ax, ax, ax, 7, 6, 5, 30, 1, 6, ax, ax, ax, None) == 0
    ft = lib.tgx_ng_time_interp_batch_dev
    ft.argtypes = [C.c_int] + [C.c_void_p] * 8
    n = (C.c_long * 2)(10, 10)
    t1 = (C.c_int * 2)(5, 5)
    ta = (C.c_int * 2)(1, 6)   # job 1: ta beyond t1
    tb = (C.c_int * 2)(1, 6)
    _err(lib, ft(2, P, P, P, n, t1, ta, tb, None), "job 1")


def test_global_grid_and_mixed_step_checks(lib):
    import sys
    sys.path.insert(0, os.path.join(ROOT))
    import helpers
    pkg = helpers.load_pkg()
    g = pkg.synth.make_grid("small", glb=True)
    st = pkg.step
    gg = st.GlbGrid()
    chunk = (C.c_int * 1)(5)   # does not divide nlon = 16
    gg.nlat_filter = 1
    gg.chunk = C.cast(chunk, C.c_void_p)
    gg.g.nlon, gg.g.nlat, gg.g.nk = g["nlon"], g["nlat"], g["nk"]
    fn = lib.tgx_glb_step
    fn.argtypes = [C.c_void_p] * 5
    fn.restype = C.c_int
    _err(lib, fn(C.byref(gg), None, None, C.c_void_p(0x1000), None), "chunk count 5")
    gg.g.nlon = 15               # odd
    _err(lib, fn(C.byref(gg), None, None, C.c_void_p(0x1000), None), "must be even")
