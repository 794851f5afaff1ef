"""Index maps of the tiled workspace regions (CPU only).

The column kernels keep per-level values in tiled regions: Geo::tile orders
them by groups of 64 consecutive columns (thread-per-column kernels),
Geo::wtile by the wavefronts of a 32x8 tile-kernel launch.  Both must be
one-to-one over (level, column or thread, value) and stay inside one region
of nv values, each value nk * max(ceil(ncol/64)*64, 64 * wavefronts) doubles
(the size tgx_ng_workspace_bytes reserves per value).  These are checked here
on the shapes the tests and the bench use, by enumerating every index.
"""
import numpy as np
import pytest

TILE_X, TILE_Y = 32, 8

SHAPES = {  # name: (nlon, nlat, nk), as in the synthetic grids
    "tiny": (8, 6, 12), "small": (16, 12, 29), "tall": (12, 8, 113), "g5": (72, 36, 29),
    "nest2": (100, 50, 57),
}


def value_elems(nlon, nlat, nk):
    nlond, nlatd = nlon + 4, nlat + 4
    ncol = -(-nlond * nlatd // 64) * 64
    nwt = 4 * -(-nlond // TILE_X) * -(-nlatd // TILE_Y)
    return nk * max(ncol, 64 * nwt)


def tile_index(k, i, j, n, nv, nlond, ngw):
    c = j * nlond + i
    return (((k * ngw + (c >> 6)) * nv + n) * 64) + (c & 63)


def wtile_index(k, block, tid, n, nv, nwt):
    w = block * 4 + (tid >> 6)
    return (((k * nwt + w) * nv + n) * 64) + (tid & 63)


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("nv", [2, 4, 7, 12])
def test_column_tile_map_is_one_to_one_and_in_region(name, nv):
    nlon, nlat, nk = SHAPES[name]
    nlond, nlatd = nlon + 4, nlat + 4
    ngw = -(-nlond * nlatd // 64)
    k, j, i, n = np.meshgrid(np.arange(nk), np.arange(nlatd), np.arange(nlond), np.arange(nv), indexing="ij")
    idx = tile_index(k, i, j, n, nv, nlond, ngw).ravel()
    assert idx.min() >= 0 and idx.max() < nv * value_elems(nlon, nlat, nk)
    assert np.unique(idx).size == idx.size
    # a wavefront of 64 consecutive columns writes one value of one level as
    # one contiguous run of 64 doubles
    c = np.arange(64 * (nlond * nlatd // 64))
    run = tile_index(3 % nk, c % nlond, c // nlond, 1 % nv, nv, nlond, ngw).reshape(-1, 64)
    assert (np.diff(run, axis=1) == 1).all()


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("nv", [2, 6])
def test_wavefront_tile_map_is_one_to_one_and_in_region(name, nv):
    nlon, nlat, nk = SHAPES[name]
    nlond, nlatd = nlon + 4, nlat + 4
    ntiles = -(-nlond // TILE_X) * -(-nlatd // TILE_Y)
    nwt = 4 * ntiles
    k, b, t, n = np.meshgrid(np.arange(nk), np.arange(ntiles), np.arange(256), np.arange(nv), indexing="ij")
    idx = wtile_index(k, b, t, n, nv, nwt).ravel()
    assert idx.min() >= 0 and idx.max() < nv * value_elems(nlon, nlat, nk)
    assert np.unique(idx).size == idx.size


# values per column and level the kernels keep in tiled regions: comp inputs
# (7) and elimination (12); minor right-hand sides (4), ratios and forward
# solutions (4), sources (4) and source coefficients (6); wind (6), TN (2) and
# O+ (2) Thomas values; the O+ factorisation kept across sub-cycles (3);
# comp's precision form: the solution between refinement sweeps (3)
TILED_VALUES = 7 + 12 + 4 + 4 + 4 + 6 + 6 + 2 + 2 + 3 + 3


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_workspace_holds_every_tiled_region(pkg, name):
    """tgx_ng_workspace_bytes reserves at least the tiled regions' extent on
    top of the level planes (read from the library, no GPU call)."""
    import ctypes as C
    nlon, nlat, nk = SHAPES[name]
    ld = pkg.step.pitch(nlon + 4)
    g = pkg.step.NgGrid()
    g.nk, g.nlon, g.nlat, g.ld = nk, nlon, nlat, ld
    nbytes = pkg.step._lib().tgx_ng_workspace_bytes(C.byref(g))
    plane = nk * ld * (nlat + 4)
    tiled = TILED_VALUES * value_elems(nlon, nlat, nk)
    assert nbytes >= 8 * tiled
    assert (nbytes // 8 - tiled) % plane == 0  # the rest is whole level planes
    assert nbytes // 8 - tiled >= 20 * plane   # prelude, chemistry and O+ planes
