"""Host-side mirror of the reference's column-solver interface over libtgx.

Names, argument order and meaning follow the Fortran subroutines
(trsolv.F:2-3, trsolv_ng.F90:1-9, duv.F:801-820).  Host arrays are numpy
float64 arrays in the reference's memory layout (level fastest): a Fortran
array real(lev0:lev1, lon0:lon1) is a C-order numpy array of shape
(nlon, nlev).  Device entry points take torch tensors already on the GPU.

Everything here calls libtgx.so; there is no CPU fallback.  If the library is
missing or no GPU is present, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TGX_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("TGX_LIBRARY") or os.path.join(HERE, "libtgx.so")
_lib = None

SCHEME_AUTO, SCHEME_THOMAS, SCHEME_PCR = 0, 1, 2

_D = C.c_void_p
_I = C.c_int
_SIGS = {
    "tgx_version": ([], _I),
    "tgx_last_error": ([], C.c_char_p),
    "tgx_last_status": ([], _I),
    "tgx_clear_error": ([], None),
    "tgx_trsolv_f64": ([_D] * 5 + [_I] * 9, None),
    "tgx_trsolv_ng_f64": ([_D] * 5 + [_I] * 4, None),
    "tgx_blktri_f64": ([_D] * 8 + [_I] * 5, None),
    "tgx_blktri3_f64": ([_D] * 5 + [_I] * 3, None),
    "tgx_bndcmp": ([C.c_double, _D, _D], None),
    "tgx_trsolv_dev_f64": ([_D] * 5 + [_I] * 6 + [_D], _I),
    "tgx_autotune_choice": ([_I] * 4, _I),
    "tgx_trsolv_dev_mixed": ([_D] * 5 + [_I] * 6 + [_D], _I),
    "tgx_blktri2_dev_f64": ([_D] * 5 + [_I] * 4 + [_D], _I),
    "tgx_blktri3_dev_f64": ([_D] * 5 + [_I] * 4 + [_D], _I),
    "tgx_ringfilter_f64": ([_D, _I, _I, _I, C.c_char_p, _I, _I, _I, C.c_double, _I, _D], None),
    "tgx_ringfilter_dev": ([_D] + [_I] * 6 + [C.c_double, _I, _D, _D], _I),
    "tgx_ng_pack_levels": ([_D, _D] + [_I] * 4 + [_D], _I),
    "tgx_ng_interp3d_dev": ([_D] + [_I] * 4 + [_D] * 4 + [_I] * 3 + [C.c_long] * 3 + [_D] * 3 + [_I, _D], _I),
    "tgx_ng_interp2d_dev": ([_D] + [_I] * 3 + [_D] * 3 + [_I] * 2 + [C.c_long] * 2 + [_D] * 3, _I),
    "tgx_ng_time_interp_dev": ([_D] * 3 + [C.c_long] + [_I] * 3 + [_D], _I),
    "tgx_ng_unpack_levels": ([_D, _D] + [_I] * 4 + [_D], _I),
    "tgx_ng_interp3d_batch_dev": ([_I, _D, _D, _D] + [_I] * 4 + [_D] * 3 + [_I] * 3 + [C.c_long] * 3 + [_D] * 4,
                                  _I),
    "tgx_ng_interp2d_batch_dev": ([_I, _D, _D] + [_I] * 3 + [_D] * 2 + [_I] * 2 + [C.c_long] * 2 + [_D] * 3, _I),
    "tgx_ng_time_interp_batch_dev": ([_I] + [_D] * 7 + [_D], _I),
}


class TgxError(RuntimeError):
    pass


def lib():
    """Loads libtgx.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TgxError("libtgx.so is not built; run tiegcm3.0_ng_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
        try:
            from . import step as _step  # noqa: F401  (registers the step signatures)
        except ImportError:
            pass
    return _lib


def last_error():
    return lib().tgx_last_error().decode()


def _check_void():
    L = lib()
    if L.tgx_last_status() != 0:
        raise TgxError(last_error())


def _check(rc):
    if rc != 0:
        raise TgxError(last_error() or ("libtgx error %d" % rc))


def _hp(a):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data


def _dp(t):
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


# ---- drop-in host entry points ------------------------------------------------
def trsolv(a, b, c, f, x, lev0, lev1, k1, k2, lon0, lon1, lonmax, lat=0, idebug=0):
    """trsolv.F:2-3.  Arrays of shape (lon1-lon0+1, lev1-lev0+1); x updated in place."""
    lib().tgx_trsolv_f64(_hp(a), _hp(b), _hp(c), _hp(f), _hp(x), lev0, lev1, k1, k2, lon0, lon1,
                         lonmax, lat, idebug)
    _check_void()
    return x


def trsolv_ng(a, b, c, f, x, k2):
    """trsolv_ng.F90:1-9.  Arrays of shape (nlatd, nlond, nk); x updated in place."""
    nlatd, nlond, nk = a.shape
    lib().tgx_trsolv_ng_f64(_hp(a), _hp(b), _hp(c), _hp(f), _hp(x), nk, nlond, nlatd, k2)
    _check_void()
    return x


def blktri(a, b, c, f, x, lev0, lev1, lon0, lon1, lonmax, beta=None, gamma=None, y=None):
    """duv.F:801-820.  a,b,c of shape (ncol, nlev, 2(col), 2(row)); f,x (ncol, nlev, 2)."""
    z = C.c_void_p(0)
    lib().tgx_blktri_f64(_hp(a), _hp(b), _hp(c), _hp(f), z if beta is None else _hp(beta),
                         z if gamma is None else _hp(gamma), z if y is None else _hp(y), _hp(x),
                         lev0, lev1, lon0, lon1, lonmax)
    _check_void()
    return x


def blktri3(p, q, r, f, x, k2):
    """comp's 3x3 elimination.  p,q,r (ncol, nk, 3, 3) column-major blocks; f,x (ncol, nk, 3)."""
    ncol, nk = p.shape[0], p.shape[1]
    lib().tgx_blktri3_f64(_hp(p), _hp(q), _hp(r), _hp(f), _hp(x), nk, ncol, k2)
    _check_void()
    return x


def _chunks(chunk):
    arr = (C.c_int * max(1, len(chunk)))(*[int(c) for c in chunk])
    return arr, len(chunk)


def ringfilter(f, order, name, lat, nlon, nlat, dlamda, chunk):
    """ringfilter.F:1-111 on one latitude, in place.  f is the reference's
    f(nlonp4, lev0:lev1): a numpy array (nlev, nlon+4); chunk = the chunk
    table cons_module chunk_array(1:nlat_filter)."""
    arr, nf = _chunks(chunk)
    lib().tgx_ringfilter_f64(_hp(f), 1, f.shape[0], int(order), name.encode(), int(lat), int(nlon), int(nlat),
                             float(dlamda), nf, C.cast(arr, C.c_void_p))
    _check_void()
    return f


def bndcmp(dz):
    b = np.zeros(9)
    fb = np.zeros(3)
    lib().tgx_bndcmp(float(dz), _hp(b), _hp(fb))
    return b.reshape(3, 3), fb


# ---- device entry points -------------------------------------------------------
def trsolv_dev(a, b, c, f, x, k1, k2, scheme=SCHEME_AUTO, col0=0, ncol=None, stream=None):
    """Batched scalar solve on device tensors of shape (ncol, ldk) (level fastest)."""
    ldk = a.shape[-1]
    ncol = a.numel() // ldk - col0 if ncol is None else ncol
    _check(lib().tgx_trsolv_dev_f64(_dp(a), _dp(b), _dp(c), _dp(f), _dp(x), ldk, k1, k2, col0,
                                    ncol, scheme, _stream(stream)))
    return x


def trsolv_dev_mixed(a, b, c, f, x, k1, k2, nsweep=2, col0=0, ncol=None, stream=None):
    ldk = a.shape[-1]
    ncol = a.numel() // ldk - col0 if ncol is None else ncol
    _check(lib().tgx_trsolv_dev_mixed(_dp(a), _dp(b), _dp(c), _dp(f), _dp(x), ldk, k1, k2, col0,
                                      ncol, nsweep, _stream(stream)))
    return x


def autotune_choice(ldk, k1, k2, ncol):
    return lib().tgx_autotune_choice(ldk, k1, k2, ncol)


def blktri2_dev(a, b, c, f, x, nrow, stream=None):
    ncol, ldk = a.shape[0], a.shape[1]
    _check(lib().tgx_blktri2_dev_f64(_dp(a), _dp(b), _dp(c), _dp(f), _dp(x), ldk, nrow, 0, ncol,
                                     _stream(stream)))
    return x


# The ring filter's chunk counts per latitude from each pole, by model
# resolution (nlon, nlat): cons_module's chunk_array (cons.F:331-355), which
# the model passes in; the filter takes any such table.
RING_CHUNKS = {
    (72, 36): [3, 9, 18, 36, 36],
    (144, 72): [3, 3, 9, 9, 18, 18, 36, 36, 72, 72],
    (288, 144): [3, 3, 3, 9, 9, 9, 18, 18, 18, 36, 36, 36, 72, 72, 72, 72, 144, 144, 144, 144, 144],
    (576, 288): [3, 3, 3, 3, 3, 9, 9, 9, 9, 9, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36,
                 72, 72, 72, 72, 72, 72, 144, 144, 144, 144, 144, 144, 288, 288, 288, 288, 288, 288],
}


def pack_levels(src, dst, nk, nlond, nlatd, stream=None):
    """Model-layout device array f(nk, nlond, nlatd) (level fastest; a torch
    tensor (nlatd, nlond, nk)) into level planes dst (nk, nlatd, ld)."""
    _check(lib().tgx_ng_pack_levels(_dp(src), _dp(dst), nk, nlond, nlatd, dst.shape[-1], _stream(stream)))
    return dst


def unpack_levels(src, dst, nk, nlond, nlatd, stream=None):
    """Level planes src (nk, nlatd, ld) into the model layout dst (nlatd, nlond, nk)."""
    _check(lib().tgx_ng_unpack_levels(_dp(src), _dp(dst), nk, nlond, nlatd, src.shape[-1], _stream(stream)))
    return dst


def interp3d_dev(src, zp, xp, yp, dst, z, x, y, dk, di, dj, zlog, offset=0, stream=None):
    """interp3d (interp.F90) of the level-plane tensor src (nzp, nyp, ld) on
    the device axes zp, xp, yp to the target axes z, x, y, written to the
    flat view of dst at offset + k*dk + i*di + j*dj."""
    nzp, nyp, ld = src.shape
    dptr = C.c_void_p(dst.data_ptr() + 8 * int(offset))
    _check(lib().tgx_ng_interp3d_dev(_dp(src), nzp, len(xp), nyp, ld, _dp(zp), _dp(xp), _dp(yp), dptr, len(z),
                                     len(x), len(y), int(dk), int(di), int(dj), _dp(z), _dp(x), _dp(y),
                                     int(bool(zlog)), _stream(stream)))


def interp2d_dev(src, xp, yp, dst, x, y, di, dj, offset=0, stream=None):
    """interp2d of the plane src (nyp, ld) to dst[offset + i*di + j*dj]."""
    nyp, ld = src.shape
    dptr = C.c_void_p(dst.data_ptr() + 8 * int(offset))
    _check(lib().tgx_ng_interp2d_dev(_dp(src), len(xp), nyp, ld, _dp(xp), _dp(yp), dptr, len(x), len(y), int(di),
                                     int(dj), _dp(x), _dp(y), _stream(stream)))


def time_interp_dev(f0, f1, dst, t1, ta, tb, stream=None):
    """dst[(t - ta)*n + p] for t = ta..tb between f0 (t = 0) and f1 (t = t1)."""
    _check(lib().tgx_ng_time_interp_dev(_dp(f0), _dp(f1), _dp(dst), f0.numel(), int(t1), int(ta), int(tb),
                                        _stream(stream)))


def _ptrs(items):
    """Host array of device pointers: tensors, or (tensor, element offset) pairs."""
    vals = []
    for it in items:
        t, off = it if isinstance(it, tuple) else (it, 0)
        vals.append(t.data_ptr() + 8 * int(off))
    return (C.c_void_p * max(1, len(vals)))(*vals)


def interp3d_batch_dev(srcs, dsts, zlogs, zp, xp, yp, z, x, y, dk, di, dj, stream=None):
    """interp3d_dev of several level-plane tensors of one shape (nzp, nyp, ld)
    on the same axes to the same target points, one launch; dsts: tensors or
    (tensor, element offset) pairs."""
    nzp, nyp, ld = srcs[0].shape
    zl = (C.c_int * max(1, len(zlogs)))(*[int(bool(v)) for v in zlogs])
    _check(lib().tgx_ng_interp3d_batch_dev(len(srcs), C.cast(_ptrs(srcs), _D), C.cast(_ptrs(dsts), _D),
                                           C.cast(zl, _D), nzp, len(xp), nyp, ld, _dp(zp), _dp(xp), _dp(yp), len(z),
                                           len(x), len(y), int(dk), int(di), int(dj), _dp(z), _dp(x), _dp(y),
                                           _stream(stream)))


def interp2d_batch_dev(srcs, dsts, xp, yp, x, y, di, dj, stream=None):
    """interp2d_dev of several planes (nyp, ld) on the same axes, one launch."""
    nyp, ld = srcs[0].shape
    _check(lib().tgx_ng_interp2d_batch_dev(len(srcs), C.cast(_ptrs(srcs), _D), C.cast(_ptrs(dsts), _D), len(xp),
                                           nyp, ld, _dp(xp), _dp(yp), len(x), len(y), int(di), int(dj), _dp(x),
                                           _dp(y), _stream(stream)))


def time_interp_batch_dev(jobs, stream=None):
    """time_interp_dev of several fields in one launch: jobs of (f0, f1, dst,
    t1, ta, tb), or (None, src, dst, 0, 0, 0) for a plain copy dst = src."""
    n = len(jobs)
    f0 = _ptrs([j[0] if j[0] is not None else j[1] for j in jobs])
    f1 = _ptrs([j[1] for j in jobs])
    dst = _ptrs([j[2] for j in jobs])
    cnt = (C.c_long * max(1, n))(*[j[1].numel() for j in jobs])
    t1 = (C.c_int * max(1, n))(*[int(j[3]) for j in jobs])
    ta = (C.c_int * max(1, n))(*[int(j[4]) for j in jobs])
    tb = (C.c_int * max(1, n))(*[int(j[5]) for j in jobs])
    _check(lib().tgx_ng_time_interp_batch_dev(n, C.cast(f0, _D), C.cast(f1, _D), C.cast(dst, _D), C.cast(cnt, _D),
                                              C.cast(t1, _D), C.cast(ta, _D), C.cast(tb, _D), _stream(stream)))


def ringfilter_dev(f, nlon, nlat, order, oplus, dlamda, chunk, stream=None):
    """The ring filter on a device tensor (nk, nlat+4, ld) in the level-plane
    layout: every latitude the chunk table filters, all levels, in place."""
    nk, ld = f.shape[0], f.shape[-1]
    arr, nf = _chunks(chunk)
    _check(lib().tgx_ringfilter_dev(_dp(f), nk, int(nlon), int(nlat), ld, int(order), int(bool(oplus)),
                                    float(dlamda), nf, C.cast(arr, C.c_void_p), _stream(stream)))
    return f


def blktri3_dev(p, q, r, f, x, nrow, stream=None):
    ncol, ldk = p.shape[0], p.shape[1]
    _check(lib().tgx_blktri3_dev_f64(_dp(p), _dp(q), _dp(r), _dp(f), _dp(x), ldk, nrow, 0, ncol,
                                     _stream(stream)))
    return x
