"""GPU parity of the plain column solvers, through the C-ABI, against the oracle.

FP64 Thomas kernels follow the reference operation order without FMA
contraction, so they must agree with the oracle to within a few ulps
(tolerance 1e-14 relative, well inside the 1e-12 drop-in bar).  PCR and the
mixed-precision solver use a different elimination order; their tolerances
are stated per test.
"""
import numpy as np
import pytest

from helpers import diag_dominant_system, gpu_available, relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def torch_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def config1_system(rng, ncol, nk):
    """5 deg dt-shaped system: a,c = -U(0.5,2); b = -(a+c)+U(0.1,1); Dirichlet
    bottom through the ghost row (q(1) -= p(1), f(1) -= 2 p(1) T_lbc, p(1) = 0)
    and r(top) = 0 (dt_ng.F90:147-163); rhs = 1000-800 exp(-(z+7)/2) + N(0,5)."""
    a, b, c = diag_dominant_system(rng, ncol, nk)
    z = -7. + 0.5 * np.arange(nk)
    f = 1000. - 800. * np.exp(-(z + 7.) / 2.) + rng.normal(0., 5., size=(ncol, nk))
    tlbc = 181. + rng.normal(0., 1., size=ncol)
    b[:, 0] -= a[:, 0]
    f[:, 0] -= 2. * a[:, 0] * tlbc
    a[:, 0] = 0.
    c[:, nk - 2] = 0.
    return a, b, c, f


def test_trsolv_dropin_config1_per_latitude(pkg, oracle):
    # dt.F:330 calls trsolv per latitude: lev0..lev1 = 1..29, rows 1..28,
    # lon0..lon1 = 1..76 (nlonp4 = 76, periodic columns skipped); every one
    # of the 5 degree grid's 36 latitudes, as one step of the model calls it.
    rng = np.random.default_rng(0x7E6C0000)
    nlev, nlonp4 = 29, 76
    for lat in range(1, 37):
        a, b, c, f = config1_system(rng, nlonp4, nlev)
        x = np.full((nlonp4, nlev), 1e30)
        y = x.copy()
        pkg.tgx.trsolv(a, b, c, f, x, 1, nlev, 1, nlev - 1, 1, nlonp4, nlonp4, lat, 0)
        oracle.trsolv(a, b, c, f, y, 1, nlev, 1, nlev - 1, 1, nlonp4, nlonp4)
        assert np.array_equal(x[:, nlev - 1], y[:, nlev - 1])   # untouched row
        assert np.array_equal(x[:2], y[:2]) and np.array_equal(x[-2:], y[-2:])
        assert relerr(x, y) <= 1e-14, relerr(x, y)


@pytest.mark.parametrize("name,k2off", [("g5", 1), ("g2p5", 0), ("nest1", 1)])
def test_trsolv_ng_dropin(pkg, oracle, name, k2off):
    g = pkg.synth.make_grid(name)
    nk, nlond, nlatd = g["nk"], g["nlon"] + 4, g["nlat"] + 4
    rng = np.random.default_rng(g["seed"])
    a, b, c, f = config1_system(rng, nlond * nlatd, nk)
    shp = (nlatd, nlond, nk)
    a, b, c, f = (np.ascontiguousarray(v.reshape(shp)) for v in (a, b, c, f))
    k2 = nk - k2off
    x = np.zeros(shp)
    y = np.zeros(shp)
    pkg.tgx.trsolv_ng(a, b, c, f, x, k2)
    oracle.trsolv_ng(a, b, c, f, y, nk, nlond * nlatd, k2)
    assert relerr(x[..., :k2], y[..., :k2]) <= 1e-14


def test_blktri_dropin(pkg, oracle):
    from test_oracle_solvers import random_blocks, to_colmajor
    rng = np.random.default_rng(11)
    ncol, nlev, nlonp4 = 76, 57, 76
    P, Q, R, F = random_blocks(rng, ncol, nlev, 2)
    args = [to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F)]
    x = np.zeros((ncol, nlev, 2))
    y = np.zeros((ncol, nlev, 2))
    pkg.tgx.blktri(*args, x, 1, nlev, 1, nlonp4, nlonp4)
    oracle.blktri(*args, y, 1, nlev, 1, nlonp4, nlonp4)
    assert np.array_equal(x[:2], y[:2]) and np.array_equal(x[-2:], y[-2:])
    assert relerr(x, y) <= 1e-14


def test_blktri3_dropin(pkg, oracle):
    from test_oracle_solvers import random_blocks, to_colmajor
    rng = np.random.default_rng(12)
    ncol, nk = 300, 57
    P, Q, R, F = random_blocks(rng, ncol, nk, 3)
    args = [to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F)]
    x = np.zeros((ncol, nk, 3))
    y = np.zeros((ncol, nk, 3))
    pkg.tgx.blktri3(*args, x, nk - 1)
    oracle.blktri3(*args, y, nk, ncol, nk - 1)
    assert relerr(x, y) <= 1e-14


@pytest.mark.parametrize("nk", [29, 57, 113, 200])
@pytest.mark.parametrize("scheme", [1, 2, 0])
def test_trsolv_dev_schemes(pkg, oracle, nk, scheme):
    import torch
    rng = np.random.default_rng(nk + 100 * scheme)
    ncol = 4099
    a, b, c, f = config1_system(rng, ncol, nk)
    k2 = nk - 2
    y = np.zeros((ncol, nk))
    oracle.trsolv_ng(a, b, c, f, y, nk, ncol, k2)
    da, db, dc, df = (torch_dev(v) for v in (a, b, c, f))
    dx = torch.full((ncol, nk), -1.0, dtype=torch.float64, device="cuda")
    pkg.tgx.trsolv_dev(da, db, dc, df, dx, 0, k2 - 1, scheme=scheme)
    torch.cuda.synchronize()
    x = dx.cpu().numpy()
    tol = 1e-14 if scheme == 1 else 1e-12
    assert relerr(x[:, :k2], y[:, :k2]) <= tol
    assert np.all(x[:, k2:] == -1.0)
    if scheme == 0:
        assert pkg.tgx.autotune_choice(nk, 0, k2 - 1, ncol) in (1, 2)


@pytest.mark.parametrize("nsweep,tol", [(0, 1e-5), (1, 1e-9), (2, 1e-10)])
def test_trsolv_mixed_precision(pkg, oracle, nsweep, tol):
    import torch
    rng = np.random.default_rng(77 + nsweep)
    ncol, nk = 8192, 113
    a, b, c, f = config1_system(rng, ncol, nk)
    y = np.zeros((ncol, nk))
    oracle.trsolv_ng(a, b, c, f, y, nk, ncol, nk - 1)
    da, db, dc, df = (torch_dev(v) for v in (a, b, c, f))
    dx = torch.zeros((ncol, nk), dtype=torch.float64, device="cuda")
    pkg.tgx.trsolv_dev_mixed(da, db, dc, df, dx, 0, nk - 2, nsweep=nsweep)
    torch.cuda.synchronize()
    err = np.abs(dx.cpu().numpy()[:, :nk - 1] - y[:, :nk - 1]) / np.abs(y[:, :nk - 1])
    assert err.max() <= tol, err.max()


def test_dev_blktri2_blktri3_match_dropin(pkg, oracle):
    import torch
    from test_oracle_solvers import random_blocks, to_colmajor
    rng = np.random.default_rng(13)
    ncol, nk = 1000, 29
    for m, fn, ofn in ((2, pkg.tgx.blktri2_dev, None), (3, pkg.tgx.blktri3_dev, None)):
        P, Q, R, F = random_blocks(rng, ncol, nk, m)
        args = [to_colmajor(P), to_colmajor(Q), to_colmajor(R), np.ascontiguousarray(F)]
        y = np.zeros((ncol, nk, m))
        if m == 2:
            oracle.blktri(*args, y, 1, nk, 5, 5 + ncol - 1, 100000)
        else:
            oracle.blktri3(*args, y, nk, ncol, nk - 1)
        dx = torch.zeros((ncol, nk, m), dtype=torch.float64, device="cuda")
        fn(*[torch_dev(v) for v in args], dx, nk - 1)
        torch.cuda.synchronize()
        assert relerr(dx.cpu().numpy()[:, :nk - 1], y[:, :nk - 1]) <= 1e-14


def test_error_paths_do_not_abort(pkg):
    import torch
    t = torch.zeros((4, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(pkg.tgx.TgxError):
        pkg.tgx.trsolv_dev(t, t, t, t, t, 0, 12)          # k2 beyond the column
    with pytest.raises(pkg.tgx.TgxError):
        pkg.tgx.trsolv_dev_mixed(t, t, t, t, t, 0, 5, nsweep=99)
    # zero columns is a no-op
    pkg.tgx.trsolv_dev(t, t, t, t, t, 0, 5, ncol=0)
    # drop-in with bad bounds reports through tgx_last_error
    a = np.zeros((3, 5))
    with pytest.raises(pkg.tgx.TgxError):
        pkg.tgx.trsolv(a, a, a, a, a.copy(), 1, 5, 0, 9, 1, 3, 3)


def test_solvers_from_two_host_threads_and_two_streams(pkg):
    """Thomas and block solves (both use library scratch) queued from two host
    threads at once, and from one thread on two streams without a wait in
    between, give bitwise the results of the same solves run one by one."""
    import threading
    import torch
    from test_oracle_solvers import random_blocks, to_colmajor
    rng = np.random.default_rng(21)
    jobs = []
    for ncol, nk in ((6000, 113), (3000, 57)):
        a, b, c, f = config1_system(rng, ncol, nk)
        P, Q, R, F = random_blocks(rng, ncol // 4, 29, 3)
        jobs.append(([torch_dev(v) for v in (a, b, c, f)], nk,
                     [torch_dev(v) for v in (to_colmajor(P), to_colmajor(Q), to_colmajor(R),
                                             np.ascontiguousarray(F))], ncol // 4))

    def run(job, stream=None):
        (sys_, nk, blk, nb) = job
        x = torch.zeros(sys_[0].shape, dtype=torch.float64, device="cuda")
        y = torch.zeros((nb, 29, 3), dtype=torch.float64, device="cuda")
        for _ in range(3):
            pkg.tgx.trsolv_dev(*sys_, x, 0, nk - 2, scheme=1, stream=stream)
            pkg.tgx.blktri3_dev(*blk, y, 28, stream=stream)
        return x, y

    ref = [run(j) for j in jobs]
    torch.cuda.synchronize()

    out, errs = [None, None], []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                out[i] = run(jobs[i], stream=s.cuda_stream)
            s.synchronize()
        except Exception as e:  # reported in the main thread
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(s1):
        o1 = run(jobs[0], stream=s1.cuda_stream)
    with torch.cuda.stream(s2):
        o2 = run(jobs[1], stream=s2.cuda_stream)
    torch.cuda.synchronize()
    for got in (out, [o1, o2]):
        for (rx, ry), (gx, gy) in zip(ref, got):
            assert torch.equal(rx, gx) and torch.equal(ry, gy)
