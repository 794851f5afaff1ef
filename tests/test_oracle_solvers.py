"""Pins the oracle's column solvers and small-matrix routines (CPU only).

No reference-held vectors exist for this path (the reference ships no tests
or fixtures), so every check here is fixed by mathematics: dense LU solves,
closed-form inverses, exact rational results, and special cases that reduce
one solver to another.
"""
import numpy as np
import pytest

from helpers import diag_dominant_system, relerr


def dense_tridiag(a, b, c, n):
    return np.diag(b[:n]) + np.diag(a[1:n], -1) + np.diag(c[:n - 1], 1)


@pytest.mark.parametrize("nk,k2", [(29, 28), (57, 57), (113, 112), (5, 1), (1, 1)])
def test_trsolv_ng_matches_dense_solve(oracle, nk, k2):
    rng = np.random.default_rng(nk * 7 + k2)
    ncol = 9
    a, b, c = diag_dominant_system(rng, ncol, nk)
    f = rng.normal(size=(ncol, nk)) * 100.
    x = np.full((ncol, nk), 7.25)
    oracle.trsolv_ng(a, b, c, f, x, nk, ncol, k2)
    for j in range(ncol):
        xs = np.linalg.solve(dense_tridiag(a[j], b[j], c[j], k2), f[j, :k2])
        assert relerr(x[j, :k2], xs) < 1e-13
    assert np.all(x[:, k2:] == 7.25), "rows beyond k2 must be left untouched"


def test_trsolv_diagonal_system_is_exact(oracle):
    # a = c = 0 reduces Thomas to x = f / b, bit for bit.
    rng = np.random.default_rng(1)
    b = rng.uniform(1, 3, size=(4, 10))
    f = rng.normal(size=(4, 10))
    z = np.zeros_like(b)
    x = np.zeros_like(b)
    oracle.trsolv_ng(z, b, z, f, x, 10, 4, 10)
    assert np.array_equal(x, f / b)


def test_trsolv_green_function_closed_form(oracle):
    # tridiag(-1,2,-1) of order n: (T^-1)_{ij} = min(i,j)(n+1-max(i,j))/(n+1).
    n = 31
    a = np.full((1, n), -1.)
    c = np.full((1, n), -1.)
    b = np.full((1, n), 2.)
    for jj in (1, 7, 16, 31):
        f = np.zeros((1, n))
        f[0, jj - 1] = 1.
        x = np.zeros((1, n))
        oracle.trsolv_ng(a, b, c, f, x, n, 1, n)
        i = np.arange(1, n + 1)
        exact = np.minimum(i, jj) * (n + 1 - np.maximum(i, jj)) / (n + 1)
        assert np.abs(x[0] - exact).max() < 1e-13


def test_trsolv_periodic_skip_and_agreement_with_trsolv_ng(oracle):
    # trsolv.F:32-35: lon0 == 1 skips columns 1,2; lon1 == lonmax skips the last two.
    rng = np.random.default_rng(2)
    nlev, lon0, lon1, lonmax = 29, 1, 76, 76
    ncol = lon1 - lon0 + 1
    a, b, c = diag_dominant_system(rng, ncol, nlev)
    f = rng.normal(size=(ncol, nlev))
    x = np.full((ncol, nlev), -3.)
    oracle.trsolv(a, b, c, f, x, 1, nlev, 1, nlev - 1, lon0, lon1, lonmax)
    assert np.all(x[:2] == -3.) and np.all(x[-2:] == -3.)
    assert np.all(x[:, nlev - 1] == -3.)
    y = np.full((ncol, nlev), -3.)
    oracle.trsolv_ng(a, b, c, f, y, nlev, ncol, nlev - 1)
    # same recurrence, no contraction: identical bits on the solved columns
    assert np.array_equal(x[2:-2], y[2:-2])
    # interior task (lon0 != 1, lon1 != lonmax): every column solved
    x2 = np.zeros((ncol, nlev))
    oracle.trsolv(a, b, c, f, x2, 1, nlev, 1, nlev - 1, 3, 3 + ncol - 1, 200)
    assert np.array_equal(x2[:, :nlev - 1], y[:, :nlev - 1])


def block_dense(P, Q, R, F, n):
    m = P.shape[-1]
    A = np.zeros((m * n, m * n))
    rhs = np.zeros(m * n)
    for k in range(n):
        A[m * k:m * k + m, m * k:m * k + m] = Q[k]
        if k > 0:
            A[m * k:m * k + m, m * k - m:m * k] = P[k]
        if k < n - 1:
            A[m * k:m * k + m, m * k + m:m * k + 2 * m] = R[k]
        rhs[m * k:m * k + m] = F[k]
    return np.linalg.solve(A, rhs).reshape(n, m)


def random_blocks(rng, ncol, nk, m):
    P = rng.normal(size=(ncol, nk, m, m))
    R = rng.normal(size=(ncol, nk, m, m))
    Q = rng.normal(size=(ncol, nk, m, m)) * .3 + np.eye(m) * (2 * m + 2)
    F = rng.normal(size=(ncol, nk, m))
    return P, Q, R, F


def to_colmajor(M):
    """(…, row, col) -> memory order of Fortran (row fastest): (…, col, row)."""
    return np.ascontiguousarray(np.swapaxes(M, -1, -2))


def test_blktri_matches_dense_block_solve(oracle):
    rng = np.random.default_rng(3)
    ncol, nlev = 6, 17
    P, Q, R, F = random_blocks(rng, ncol, nlev, 2)
    x = np.zeros((ncol, nlev, 2))
    oracle.blktri(to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F), x,
                  1, nlev, 5, 5 + ncol - 1, 144)
    for j in range(ncol):
        xs = block_dense(P[j], Q[j], R[j], F[j], nlev - 1)
        assert relerr(x[j, :nlev - 1], xs) < 1e-13


def test_blktri_diagonal_blocks_reduce_to_two_scalar_solves(oracle):
    # p, q, r diagonal: UN and VN decouple into two trsolv systems.
    rng = np.random.default_rng(4)
    ncol, nlev = 3, 20
    a, b, c = diag_dominant_system(rng, 2 * ncol, nlev)
    f = rng.normal(size=(2 * ncol, nlev))
    P = np.zeros((ncol, nlev, 2, 2)); Q = np.zeros_like(P); R = np.zeros_like(P)
    for comp in range(2):
        P[:, :, comp, comp] = a[comp::2]
        Q[:, :, comp, comp] = b[comp::2]
        R[:, :, comp, comp] = c[comp::2]
    F = np.stack([f[0::2], f[1::2]], axis=-1)
    x = np.zeros((ncol, nlev, 2))
    oracle.blktri(to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F), x,
                  1, nlev, 3, 3 + ncol - 1, 144)
    y = np.zeros((2 * ncol, nlev))
    oracle.trsolv_ng(a, b, c, f, y, nlev, 2 * ncol, nlev - 1)
    for comp in range(2):
        assert relerr(x[:, :nlev - 1, comp], y[comp::2, :nlev - 1]) < 1e-15


@pytest.mark.parametrize("nk,k2", [(29, 28), (57, 56), (8, 8)])
def test_blktri3_matches_dense_block_solve(oracle, nk, k2):
    rng = np.random.default_rng(nk)
    ncol = 5
    P, Q, R, F = random_blocks(rng, ncol, nk, 3)
    x = np.full((ncol, nk, 3), 9.)
    oracle.blktri3(to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F), x,
                   nk, ncol, k2)
    for j in range(ncol):
        xs = block_dense(P[j], Q[j], R[j], F[j], k2)
        assert relerr(x[j, :k2], xs) < 1e-13
    assert np.all(x[:, k2:] == 9.)


def test_matinv3_exact_integer_inverse(oracle):
    # det = 1, so adjugate/det is exact: catches a transposed adjugate or a sign.
    A = np.array([[1., 2., 3.], [0., 1., 4.], [5., 6., 0.]])
    inv = np.array([[-24., 18., 5.], [20., -15., -4.], [-5., 4., 1.]])
    assert np.array_equal(oracle.matinv3(A), inv)
    rng = np.random.default_rng(5)
    for _ in range(20):
        M = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        assert np.abs(oracle.matinv3(M) @ M - np.eye(3)).max() < 1e-14


def test_bndcmp_closed_form(oracle, pkg):
    # lbc.F:527-594: b = (e/dz - f/2)^-1 (e/dz + f/2), fb = (e/dz - f/2)^-1 g.
    for dz in (0.5, 0.25, 0.125):
        e = np.zeros((3, 3)); e[1, 1] = 1.
        f = np.array([[1., 1., 0.], [0., -1., 0.], [0., 0., 1.]])
        g = -np.array([0.22 + 0.014, 0., 0.1154E-5])
        w1 = e / dz - f / 2.
        b_ref = np.linalg.solve(w1, e / dz + f / 2.)
        fb_ref = np.linalg.solve(w1, g)
        b, fb = oracle.bndcmp(dz)
        assert np.abs(b - b_ref).max() < 1e-14 and np.abs(fb - fb_ref).max() < 1e-15
        bs, fbs = pkg.synth.bndcmp(dz)
        assert np.array_equal(np.array(bs), b) and np.array_equal(np.array(fbs), fb)
