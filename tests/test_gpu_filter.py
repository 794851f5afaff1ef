"""GPU parity of the ring filter (ringfilter.F:1-111) through the C-ABI
against the oracle (tests/test_oracle_ringfilter.py pins it): the device entry
point over every latitude the chunk table filters (cons.F:331-355), at the
5 deg and 0.625 deg resolutions, for the TN / wind / minor form (3 harmonics),
the O2-O form (1 harmonic) and the O+ form; the host drop-in on one latitude.
Bound: 1e-12 of each row's largest magnitude (the Fourier sums are reduced in
a tree on the GPU, in longitude order in the reference)."""
import math

import numpy as np
import pytest

from helpers import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

TOL = 1e-12


def _field(nk, nlat, nlon, ld, seed, signed):
    import torch
    g = torch.Generator().manual_seed(seed)
    lon = torch.arange(nlon + 4, dtype=torch.float64) * (2 * math.pi / nlon)
    base = (torch.cos(2 * lon) + .5 * torch.sin(5 * lon + .3)).view(1, 1, -1)
    f = torch.randn(nk, nlat + 4, nlon + 4, generator=g, dtype=torch.float64) * 5. + 40. * base
    if not signed:
        f = f + 1000.
    out = torch.zeros(nk, nlat + 4, ld, dtype=torch.float64)
    out[..., :nlon + 4] = f
    return out


@pytest.mark.parametrize("res", [(72, 36, 29), (576, 288, 113)], ids=["g5", "g0p625"])
@pytest.mark.parametrize("form", [("TN", 3, 0, False), ("UN", 3, 0, True), ("O2O", 1, 0, False), ("OPLUS", 3, 1, False)],
                         ids=lambda f: f[0])
def test_ringfilter_dev_matches_oracle(pkg, oracle, res, form):
    import torch
    nlon, nlat, nk = res
    name, order, oplus, signed = form
    ld = (nlon + 4 + 15) // 16 * 16
    chunk = oracle.RING_CHUNKS[(nlon, nlat)]
    dlam = 2 * math.pi / nlon
    host = _field(nk, nlat, nlon, ld, 11 + nlon + order, signed)
    dev = host.cuda()
    print("case %s %dx%dx%d" % (name, nlon, nlat, nk), flush=True)
    pkg.tgx.ringfilter_dev(dev, nlon, nlat, order, oplus, dlam, chunk)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    ref = host.numpy().copy()
    worst = 0.
    for lat in range(1, nlat + 1):
        nc = oracle.ringfilter_chunks(lat, nlat, chunk)
        j = lat + 1
        if nc == 0:
            assert np.array_equal(got[:, j], ref[:, j]), lat   # not filtered: untouched
            continue
        row = np.ascontiguousarray(ref[:, j, :nlon + 4])
        oracle.ringfilter_row(row, nlon, order, oplus, nc, dlam)
        scale = np.abs(row).max(axis=1, keepdims=True)
        worst = max(worst, float((np.abs(got[:, j, :nlon + 4] - row) / scale).max()))
        assert np.array_equal(got[:, j, nlon + 4:], ref[:, j, nlon + 4:])   # row padding untouched
    for j in (0, 1, nlat + 2, nlat + 3):                                  # ghost rows untouched
        assert np.array_equal(got[:, j], ref[:, j])
    print("worst row-relative error %.3e" % worst)
    assert worst <= TOL


def test_ringfilter_host_dropin_matches_oracle(pkg, oracle):
    nlon, nlat, nk = 144, 72, 57
    chunk = oracle.RING_CHUNKS[(nlon, nlat)]
    dlam = 2 * math.pi / nlon
    rng = np.random.default_rng(2)
    for lat, name in ((1, "TN"), (3, "OPLUS"), (70, "UN"), (30, "TN")):
        f = rng.normal(size=(nk, nlon + 4)) * 3. + 250.
        ref = f.copy()
        nc = oracle.ringfilter_chunks(lat, nlat, chunk)
        oracle.ringfilter_row(ref, nlon, 3, name == "OPLUS", nc, dlam)
        got = f.copy()
        pkg.tgx.ringfilter(got, 3, name, lat, nlon, nlat, dlam, chunk)
        if nc == 0:
            assert np.array_equal(got, f)
        assert float((np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max()) <= TOL, (lat, name)
