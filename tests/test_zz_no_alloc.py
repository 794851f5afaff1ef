"""tgx_trsolv_dev_f64_noalloc (the allocation-free entry point) against
tgx_trsolv_dev_f64, called through the C-ABI with ctypes on device buffers.

Both run the same Thomas (or PCR) kernel on the same data, so the outputs must
be bitwise equal.  Two calls in a row, the second on a smaller batch, check
that the shared fixed scratch array carries nothing from one call to the next;
a batch whose Thomas ratios would not fit the array is rejected by the host
before any launch."""
import ctypes as C

import numpy as np
import pytest

from helpers import ROOT, diag_dominant_system, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

_P = C.c_void_p
_I = C.c_int
SCRATCH_DOUBLES = 8 * 1024 * 1024


def _lib():
    import os
    L = C.CDLL(os.path.join(ROOT, "tiegcm3.0_ng_amd", "libtgx.so"))
    for n in ("tgx_trsolv_dev_f64", "tgx_trsolv_dev_f64_noalloc"):
        fn = getattr(L, n)
        fn.argtypes = [_P] * 5 + [_I] * 6 + [_P]
        fn.restype = _I
    L.tgx_last_error.restype = C.c_char_p
    return L


def _system(ncol, nk, seed):
    import torch
    rng = np.random.default_rng(seed)
    a, b, c = diag_dominant_system(rng, ncol, nk)
    f = rng.normal(size=(ncol, nk)) * 100.
    dev = [torch.tensor(v, dtype=torch.float64, device="cuda").contiguous() for v in (a, b, c, f)]
    return dev


def _call(L, name, sys_, x, k1, k2, col0, ncol, scheme, stream):
    a, b, c, f = sys_
    rc = getattr(L, name)(a.data_ptr(), b.data_ptr(), c.data_ptr(), f.data_ptr(), x.data_ptr(),
                          a.shape[1], k1, k2, col0, ncol, scheme, stream)
    assert rc == 0, (name, rc, L.tgx_last_error())


@pytest.mark.parametrize("scheme", [1, 2], ids=["thomas", "pcr"])
def test_noalloc_equals_allocating_entry_point_twice(scheme):
    import torch
    L = _lib()
    stream = torch.cuda.current_stream().cuda_stream
    # first a large batch (113 levels x 40,000 columns = 4.5 M doubles of
    # ratios), then a smaller one at an offset into the batch
    for ncol_all, nk, col0, ncol, seed in ((40000, 113, 0, 40000, 11), (3000, 57, 700, 1900, 12)):
        print("case: scheme=%d ncol=%d nk=%d col0=%d" % (scheme, ncol, nk, col0), flush=True)
        sys_ = _system(ncol_all, nk, seed)
        x_ref = torch.full((ncol_all, nk), -7.0, dtype=torch.float64, device="cuda")
        x_new = x_ref.clone()
        _call(L, "tgx_trsolv_dev_f64", sys_, x_ref, 0, nk - 2, col0, ncol, scheme, stream)
        _call(L, "tgx_trsolv_dev_f64_noalloc", sys_, x_new, 0, nk - 2, col0, ncol, scheme, stream)
        torch.cuda.synchronize()
        assert torch.equal(x_ref, x_new)
        # untouched outside the solved rows and columns
        assert bool((x_new[:col0] == -7.0).all()) and bool((x_new[col0 + ncol:] == -7.0).all())
        assert bool((x_new[:, nk - 1] == -7.0).all())


def test_noalloc_rejects_batches_beyond_its_scratch():
    import torch
    L = _lib()
    nk = 128
    ncol = SCRATCH_DOUBLES // nk + 1
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    # the size check happens on the host before any launch: pointers are never read
    rc = L.tgx_trsolv_dev_f64_noalloc(t.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                      nk, 0, nk - 1, 0, ncol, 1, torch.cuda.current_stream().cuda_stream)
    assert rc == 1
    assert b"exceed the fixed scratch" in L.tgx_last_error()
