"""Pins of the oracle's ring filter (ringfilter.F:1-111) by what the
mathematics fixes, and planted mistakes (oracle.set_mutation 701-704) that
fail them.  The filter keeps a row's Fourier harmonics up to `order`, averages
the rest over nchnk equal chunks of longitude, and rebuilds it from the chunk
averages with a four-chunk reconstruction of the chunk edges: linear inside a
chunk, or (O+) the parabola with the chunk's average."""
import math

import numpy as np
import pytest

NLON = 72
DLAM = 2 * math.pi / NLON


def _theta():
    return np.arange(1, NLON + 1) * DLAM   # longitude of interior point i = 1..nlon


def _row(w):
    """f(nlon+4, nlev) of the reference for levels w (nlev, nlon) or one level (nlon,)."""
    w = np.atleast_2d(w)
    f = np.zeros((w.shape[0], NLON + 4))
    f[:, 2:NLON + 2] = w
    f[:, :2] = w[:, -2:]
    f[:, NLON + 2:] = w[:, :2]
    return f


def _harmonics(order, rng):
    th = _theta()
    w = np.full(NLON, rng.normal() * 100. + 500.)
    for m in range(1, order + 1):
        w = w + rng.normal() * 20. * np.cos(m * th) + rng.normal() * 20. * np.sin(m * th)
    return w


def harmonic_rows_pass(oracle):
    rng = np.random.default_rng(3)
    err = 0.
    for order in (1, 3):
        for oplus in (0, 1):
            for nchnk in (3, 9, 36, 72):
                w = _harmonics(order, rng)
                f = _row(w)
                oracle.ringfilter_row(f, NLON, order, oplus, nchnk, DLAM)
                err = max(err, np.abs(f[0, 2:NLON + 2] - w).max() / np.abs(w).max())
    return err


def nyquist_removed(oracle):
    """The alternating mode (-1)^i is orthogonal to every kept harmonic and
    averages to zero over chunks of even length: it is removed entirely."""
    rng = np.random.default_rng(4)
    err = 0.
    for oplus in (0, 1):
        for nchnk in (3, 9, 18, 36):
            w = _harmonics(3, rng)
            f = _row(w + 7. * (-1.) ** np.arange(NLON))
            oracle.ringfilter_row(f, NLON, 3, oplus, nchnk, DLAM)
            err = max(err, np.abs(f[0, 2:NLON + 2] - w).max() / np.abs(w).max())
    return err


def oplus_conserves_chunk_means(oracle):
    """The O+ parabola of each chunk has that chunk's average, and the kept
    harmonics are unchanged: every chunk keeps its mean."""
    rng = np.random.default_rng(5)
    err = 0.
    for nchnk in (3, 9, 18, 36):
        w = rng.uniform(1e4, 1e6, NLON)
        f = _row(w)
        oracle.ringfilter_row(f, NLON, 3, 1, nchnk, DLAM)
        out = f[0, 2:NLON + 2]
        err = max(err, np.abs(out.reshape(nchnk, -1).mean(1) - w.reshape(nchnk, -1).mean(1)).max() / w.mean())
    return err


def edges_exact_for_cubics(oracle):
    """With no harmonic kept (order 0, only the mean removed) and a row that
    samples a cubic P over the interior chunks, the chunk averages are cell
    averages of Q = P - P''/24 (midpoint rule, exact for cubics), and the
    edge reconstruction (-A(c-2) + 7A(c-1) + 7A(c) - A(c+1))/12 is exact for
    cubic cell averages: inside chunk c the linear variant runs from Q at the
    chunk's left edge to Q at its right edge."""
    nchnk, n = 9, NLON // 9
    x = np.arange(NLON, dtype=np.float64)
    P = lambda t: 3. + .2 * t - .01 * t ** 2 + 1e-4 * t ** 3
    Q = lambda t: P(t) - (-.02 + 6e-4 * t) / 24.
    w = P(x)
    f = _row(w)
    oracle.ringfilter_row(f, NLON, 0, 0, nchnk, DLAM)
    out = f[0, 2:NLON + 2]
    err = 0.
    for c in range(2, nchnk - 2):          # chunks whose stencil stays inside the cubic
        L, R = Q(c * n - .5), Q((c + 1) * n - .5)   # the chunk's edges
        for m in range(1, n + 1):
            want = (L * (n - m) + R * m) / n
            err = max(err, abs(out[c * n + m - 1] - want) / abs(want))
    return err


def periodic_points_and_untouched_rows(oracle):
    rng = np.random.default_rng(6)
    w = rng.normal(size=(3, NLON)) * 10. + 300.
    f = _row(w)
    g0 = f.copy()
    oracle.ringfilter_row(g0, NLON, 3, 0, 0, DLAM)          # nchnk 0: not filtered
    untouched = float(np.abs(g0 - f).max())
    oracle.ringfilter_row(f, NLON, 3, 0, 9, DLAM)
    wr = f[:, 2:NLON + 2]
    per = float(max(np.abs(f[:, 0:2] - wr[:, -2:]).max(), np.abs(f[:, NLON + 2:] - wr[:, :2]).max()))
    return max(untouched, per)


CHECKS = {"harmonics": (harmonic_rows_pass, 1e-12), "nyquist": (nyquist_removed, 1e-12),
          "oplus_chunk_means": (oplus_conserves_chunk_means, 1e-13), "cubic_edges": (edges_exact_for_cubics, 1e-12),
          "periodic": (periodic_points_and_untouched_rows, 0.0)}
MUTATIONS = {701: "edge reconstruction weight 6 for 7", 702: "O+ parabola's linear term with the wrong sign",
             703: "periodic points one longitude off", 704: "O+ parabola's quadratic term over n, not n^2"}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_ringfilter_oracle_check(oracle, name):
    fn, tol = CHECKS[name]
    e = fn(oracle)
    print(name, "%.3e" % e)
    assert e <= tol, (name, e)


@pytest.mark.parametrize("mid", sorted(MUTATIONS))
def test_ringfilter_planted_mistake_fails_a_check(oracle, mid):
    oracle.set_mutation(mid)
    try:
        failed = [n for n, (fn, tol) in CHECKS.items() if not fn(oracle) <= tol]
    finally:
        oracle.set_mutation(0)
    assert failed, (mid, MUTATIONS[mid])


def test_chunk_table_lookup(oracle):
    """ringfilter.F:23-34: the first and last nlat_filter latitudes take the
    table's chunk counts, counted from each pole; the others are not filtered."""
    chunk = oracle.RING_CHUNKS[(72, 36)]
    got = [oracle.ringfilter_chunks(lat, 36, chunk) for lat in range(1, 37)]
    assert got == chunk + [0] * 26 + chunk[::-1]
