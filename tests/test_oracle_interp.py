"""Pins of the oracle's interpolation between nest levels (interp.F90:1-128,
interp_fields_ng.F90:17-63): bit for bit against the reference's own compiled
interp.F90 (live where oracle/_ref is built, and through the committed
vectors tests/golden/interp_vectors.npz written from it), and by what the
mathematics fixes: the bracket search against brute force, exactness for
fields linear in (x, y, z) (inside and beyond the source range) and for
fields exponential in z times linear in (x, y) in the log-linear form, and
the time interpolation's end points and linearity.  Planted mistakes
(oracle.set_mutation 801-803) each fail a check."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "interp_vectors.npz")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_interp.so")


def _gold_cases():
    d = np.load(GOLD)
    names = sorted({k.split("_")[0] for k in d.files})
    return [(n, {k[len(n) + 1:]: d[k] for k in d.files if k.startswith(n + "_")}) for n in names]


def test_oracle_matches_committed_reference_vectors(oracle):
    for name, c in _gold_cases():
        f3 = oracle.interp3d(c["z"], c["x"], c["y"], c["zp"], c["xp"], c["yp"], c["fp"], bool(c["zlog"]))
        assert np.array_equal(f3, c["f3"], equal_nan=True), name
        f2 = oracle.interp2d(c["x"], c["y"], c["xp"], c["yp"], np.ascontiguousarray(c["fp"][:, :, 0]))
        assert np.array_equal(f2, c["f2"]), name
        f1 = oracle.interp1d(c["z"], c["zp"], np.ascontiguousarray(c["fp"][0, 0, :]))
        assert np.array_equal(f1, c["f1"]), name


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
def test_oracle_bitwise_equal_to_compiled_reference(oracle):
    ref = C.CDLL(REF)
    rng = np.random.default_rng(7)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    for trial in range(20):
        zp = np.sort(rng.uniform(-7, 7, rng.integers(3, 9)))
        xp = np.sort(rng.uniform(0, 50, rng.integers(3, 9)))
        yp = np.sort(rng.uniform(-40, 40, rng.integers(3, 9)))
        z = rng.uniform(-8, 8, rng.integers(1, 7))
        x = rng.uniform(-5, 55, rng.integers(1, 7))
        y = rng.uniform(-45, 45, rng.integers(1, 7))
        fp = rng.uniform(0.1, 3., (len(yp), len(xp), len(zp)))
        for zlog in (0, 1):
            want = np.empty((len(y), len(x), len(z)))
            ref.tgx_ref_interp3d(p(z), len(z), p(x), len(x), p(y), len(y), p(zp), len(zp), p(xp), len(xp), p(yp),
                                 len(yp), p(fp), zlog, p(want))
            got = oracle.interp3d(z, x, y, zp, xp, yp, fp, zlog)
            assert np.array_equal(got, want, equal_nan=True), (trial, zlog)


def _worse(err, e):
    """The larger error; a NaN (a planted mistake's 0/0) counts as infinite."""
    return np.inf if np.isnan(e) else max(err, e)


def find_brute(oracle, rng):
    """0 on success, else the number of wrong brackets."""
    bad = 0
    for trial in range(50):
        xp = np.sort(rng.uniform(-10, 10, rng.integers(2, 12)))
        for x0 in list(rng.uniform(-12, 12, 20)) + list(xp):
            want = 0 if x0 < xp[0] else (len(xp) if x0 >= xp[-1] else int(np.nonzero(xp <= x0)[0][-1]) + 1)
            bad += oracle.interp_find(xp, x0) != want
    return float(bad)


def linear_exact(oracle, rng):
    """Trilinear (separable) interpolation, extrapolated from the end
    intervals, reproduces a field linear in (x, y, z) everywhere."""
    err = 0.
    for trial in range(10):
        zp = np.sort(rng.uniform(-7, 7, 6)); xp = np.sort(rng.uniform(0, 30, 7)); yp = np.sort(rng.uniform(-20, 20, 5))
        a, b, c, d = rng.normal(size=4)
        fp = a + b * xp[None, :, None] + c * yp[:, None, None] + d * zp[None, None, :]
        z = rng.uniform(-9, 9, 8); x = rng.uniform(-5, 35, 6); y = rng.uniform(-25, 25, 4)
        got = oracle.interp3d(z, x, y, zp, xp, yp, fp, False)
        want = a + b * x[None, :, None] + c * y[:, None, None] + d * z[None, None, :]
        err = _worse(err, float(np.abs(got - want).max() / np.abs(fp).max()))
    return err


def loglinear_exact(oracle, rng):
    """The log form reproduces exp(a + b z) * (c + d x + e y) (positive):
    bilinear in (x, y) at every level, then linear in log f."""
    err = 0.
    for trial in range(10):
        zp = np.sort(rng.uniform(-7, 7, 6)); xp = np.sort(rng.uniform(0, 30, 7)); yp = np.sort(rng.uniform(-20, 20, 5))
        a, b = rng.normal(), rng.uniform(-1, 1)
        c, d, e = 50., rng.uniform(-.5, .5), rng.uniform(-.5, .5)
        L = lambda xx, yy: c + d * xx + e * yy
        fp = np.exp(a + b * zp)[None, None, :] * L(xp[None, :, None], yp[:, None, None])
        z = rng.uniform(-9, 9, 8); x = rng.uniform(0, 30, 6); y = rng.uniform(-20, 20, 4)
        got = oracle.interp3d(z, x, y, zp, xp, yp, fp, True)
        want = np.exp(a + b * z)[None, None, :] * L(x[None, :, None], y[:, None, None])
        err = _worse(err, float((np.abs(got - want) / want).max()))
    return err


def time_ends_and_linear(oracle, rng):
    """t = 0 and t = t1 give the end values (to rounding), the values in
    between lie on the line."""
    f0, f1 = rng.normal(size=50), rng.normal(size=50)
    err = 0.
    for t1 in (2, 5, 10, 50):
        f = oracle.time_interp(f0, f1, t1, 0, t1)
        t = np.arange(t1 + 1)[:, None]
        want = f0 + (f1 - f0) * t / t1
        err = _worse(err, float(np.abs(f - want).max()))
    return err


CHECKS = {"find": (find_brute, 0.0), "linear": (linear_exact, 5e-12), "loglinear": (loglinear_exact, 1e-13),
          "time": (time_ends_and_linear, 1e-14)}
MUTATIONS = {801: "degenerate bracket", 802: "bracket weights exchanged", 803: "log field interpolated linearly in z"}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_interp_oracle_check(oracle, name):
    fn, tol = CHECKS[name]
    e = fn(oracle, np.random.default_rng(11))
    print(name, "%.3e" % e)
    assert e <= tol, (name, e)


@pytest.mark.parametrize("mid", sorted(MUTATIONS))
def test_interp_planted_mistake_fails_a_check(oracle, mid):
    oracle.set_mutation(mid)
    try:
        with np.errstate(all="ignore"):
            failed = [n for n, (fn, tol) in CHECKS.items() if not fn(oracle, np.random.default_rng(11)) <= tol]
    finally:
        oracle.set_mutation(0)
    assert failed, (mid, MUTATIONS[mid])
