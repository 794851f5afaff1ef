"""Pins the oracle's horizontal stencils with closed forms (CPU only).

df_3d/fm_3d (dffm_ng.F90), smooth_ng (Shapiro filter = f - s*delta^4 f in the
interior), advec_ng (4th-order advection) and div_ng are checked on
polynomial fields whose finite differences are known exactly.
"""
import numpy as np
import pytest

from helpers import load_pkg


def grid_state(oracle, name="tiny"):
    pkg = load_pkg()
    g = pkg.synth.make_grid(name)
    F = pkg.synth.to_fortran_layout(pkg.synth.make_state(g))
    return g, F, oracle.OracleState(g, F)


def lonfield(g, fn):
    nlatd, nlond, nk = g["nlat"] + 4, g["nlon"] + 4, g["nk"]
    i = np.arange(nlond, dtype=np.float64).reshape(1, nlond, 1)
    return np.ascontiguousarray(np.broadcast_to(fn(i), (nlatd, nlond, nk)))


def latfield(g, fn):
    nlatd, nlond, nk = g["nlat"] + 4, g["nlon"] + 4, g["nk"]
    j = np.arange(nlatd, dtype=np.float64).reshape(nlatd, 1, 1)
    return np.ascontiguousarray(np.broadcast_to(fn(j), (nlatd, nlond, nk)))


@pytest.mark.parametrize("order", [1, 2, -1, -2])
def test_df_fm_on_linear_fields(oracle, order):
    g, _, _ = grid_state(oracle)
    f = lonfield(g, lambda i: 3. + 2. * i) if order > 0 else latfield(g, lambda j: 3. + 2. * j)
    df = oracle.df3d(f, order)
    assert np.array_equal(df, np.full_like(f, 2.))       # exact slope, edges one-sided
    fm = oracle.fm3d(f, order)
    assert np.array_equal(fm, f)                          # centred mean of a line


def test_smooth_constant_and_polynomials(oracle):
    g, _, st = grid_state(oracle)
    s = 0.01
    c = np.full((g["nlat"] + 4, g["nlon"] + 4, g["nk"]), 5.5)
    assert np.array_equal(st.smooth(c, s), c)
    # cubic in lon: delta^4 of a cubic is 0 -> interior unchanged
    f3 = lonfield(g, lambda i: (i - 3.) ** 3)
    out = st.smooth(f3, s)
    assert np.abs(out[:, 2:-2] - f3[:, 2:-2]).max() < 1e-11
    # quartic in lon: delta^4 i^4 = 24 -> f - 24 s in the interior
    f4 = lonfield(g, lambda i: i ** 4)
    out = st.smooth(f4, s)
    assert np.abs(out[:, 2:-2] - (f4[:, 2:-2] - 24. * s)).max() < 1e-9
    # next to the edge (smooth_ng.F90:300-306): f + 4 s (f(i+1) - 2 f + f(i-1)) = f + 8 s for i^2
    f2 = lonfield(g, lambda i: i ** 2)
    out = st.smooth(f2, s)
    assert np.abs(out[:, 1] - (f2[:, 1] + 8. * s)).max() < 1e-12
    assert np.array_equal(out[:, 0], f2[:, 0])
    # the same in latitude (first pass)
    f4y = latfield(g, lambda j: j ** 4)
    out = st.smooth(f4y, s)
    assert np.abs(out[2:-2] - (f4y[2:-2] - 24. * s)).max() < 1e-9


def test_advec_linear_fields(oracle):
    # f = alpha*lon_index, un = U, vn = 0:  advec = alpha*U*racs(lat)/dlamda
    g, F, _ = grid_state(oracle)
    F = dict(F)
    U, V, alpha, beta = 1234.5, -321., 0.75, -1.5
    F["un"] = np.full_like(F["un"], U)
    F["vn"] = np.zeros_like(F["vn"])
    st = oracle.OracleState(g, F)
    f = lonfield(g, lambda i: alpha * i)
    got = st.advec(f)
    racs = np.asarray(g["racs"]).reshape(-1, 1, 1)
    assert np.abs(got - alpha * U * racs / g["dlamda"]).max() <= 1e-15 * np.abs(got).max() * 10
    # f = beta*lat_index, un = 0, vn = V: advec = beta*V/(re*dphi)
    F["un"] = np.zeros_like(F["un"])
    F["vn"] = np.full_like(F["vn"], V)
    st = oracle.OracleState(g, F)
    got = st.advec(latfield(g, lambda j: beta * j))
    assert np.abs(got - beta * V / 6.37122e8 / g["dphi"]).max() <= 1e-15 * np.abs(got).max() * 10


def test_div_fourth_order_on_cubic(oracle):
    # 4/3 D1 - 1/3 D2 is exact for cubics: derivative of i^3 is 3 i^2 (interior)
    g, _, st = grid_state(oracle, "small")
    f = lonfield(g, lambda i: i ** 3)
    d = st.div(f, 1)
    i = np.arange(g["nlon"] + 4)
    assert np.abs(d[:, 2:-2, :] - (3. * i[2:-2] ** 2).reshape(1, -1, 1)).max() < 1e-10
    # edge columns use the 2-point centred/one-sided difference (div_ng.F90:337-338)
    assert np.array_equal(d[:, 0], f[:, 1] - f[:, 0])
