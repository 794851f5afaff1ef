"""The coupling between nest levels on the GPU (tiegcm3.0_ng_amd/nest.py:
map_in, set_bndry, map_out and the time interpolation of the boundary data,
levels_inner_ng.F90:131-410, interp_fields_ng.F90:17-63) against the
oracle's interpolation (pinned bit for bit against the compiled reference
interp.F90 by tests/test_oracle_interp.py).

The interpolation kernels: the linear forms bit for bit, the log-linear form
(the device's log and exp) within 1e-14.  The coupled recursion: a global
grid, a nest in it and a nest in that one (different level sets), two steps
of the outer grid, every prognostic field of every domain against the same
recursion composed from the oracle, pointwise to 1e-11 (as the chained
nest test)."""
import numpy as np
import pytest

from helpers import field_errs, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _axes(rng, n, lo, step):
    return lo + step * np.arange(n) + rng.uniform(-0.2, 0.2, n) * step * (np.arange(n) % 2)


@pytest.mark.parametrize("zlog", [0, 1])
def test_interp3d_kernel_matches_oracle(pkg, oracle, zlog):
    import torch
    rng = np.random.default_rng(5 + zlog)
    nzp, nxp, nyp, ld = 9, 13, 11, 16
    zp, xp, yp = np.linspace(-7, 7, nzp), _axes(rng, nxp, 10., 2.), _axes(rng, nyp, -20., 3.)
    fp = rng.uniform(0.5, 2., (nyp, nxp, nzp)) * np.exp(-0.4 * zp)
    if not zlog:
        fp = fp - 1.0
    # targets inside, on nodes and outside the source range (log form: inside horizontally)
    m = 0. if zlog else 2.
    z = np.sort(np.concatenate([rng.uniform(-8, 8, 10), zp[:3]]))
    x = np.sort(np.concatenate([rng.uniform(xp[0] - m, xp[-1] + m, 6), xp[4:6]]))
    y = np.sort(np.concatenate([rng.uniform(yp[0] - m, yp[-1] + m, 5), yp[1:2]]))
    want = oracle.interp3d(z, x, y, zp, xp, yp, fp, zlog)                  # (ny, nx, nz)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    src = torch.zeros(nzp, nyp, ld, dtype=torch.float64, device="cuda")
    src[:, :, :nxp] = dev(np.moveaxis(fp, 2, 0))
    out = torch.full((len(z), len(y), len(x)), np.nan, dtype=torch.float64, device="cuda")
    pkg.tgx.interp3d_dev(src, dev(zp), dev(xp), dev(yp), out, dev(z), dev(x), dev(y),
                         len(y) * len(x), 1, len(x), zlog)
    torch.cuda.synchronize()
    got = np.moveaxis(out.cpu().numpy(), 0, 2)                             # (ny, nx, nz)
    if zlog:
        err = float((np.abs(got - want) / np.abs(want)).max())
        assert err <= 1e-14, err
    else:
        assert np.array_equal(got, want)


def test_interp2d_and_time_kernels_match_oracle(pkg, oracle):
    import torch
    rng = np.random.default_rng(9)
    nxp, nyp, ld = 12, 9, 16
    xp, yp = _axes(rng, nxp, 0., 1.5), _axes(rng, nyp, 30., 0.7)
    fp = rng.normal(size=(nyp, nxp))
    x, y = np.sort(rng.uniform(-2, 20, 7)), np.sort(rng.uniform(29, 37, 5))
    want = oracle.interp2d(x, y, xp, yp, fp)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    src = torch.zeros(nyp, ld, dtype=torch.float64, device="cuda")
    src[:, :nxp] = dev(fp)
    out = torch.zeros(len(y), len(x), dtype=torch.float64, device="cuda")
    pkg.tgx.interp2d_dev(src, dev(xp), dev(yp), out, dev(x), dev(y), 1, len(x))
    f0, f1 = rng.normal(size=(3, 40)), rng.normal(size=(3, 40))
    tw = oracle.time_interp(f0, f1, 20, 0, 20)
    tout = torch.zeros((21, 3, 40), dtype=torch.float64, device="cuda")
    pkg.tgx.time_interp_dev(dev(f0), dev(f1), tout, 20, 0, 20)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), want)
    assert np.array_equal(tout.cpu().numpy(), tw)


def test_batched_coupling_kernels_equal_single_launches(pkg):
    """The batched interp3d / interp2d / time-interpolation launches (one
    launch for up to 32 fields, as set_bndry, map_in, map_out and the
    sub-step inputs use them) give bit for bit what one launch per field
    gives, the log-linear and linear forms mixed in one batch, time jobs of
    different sizes and plain copies beside interpolations."""
    import torch
    rng = np.random.default_rng(11)
    nzp, nxp, nyp, ld = 9, 13, 11, 16
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    zp, xp, yp = dev(np.linspace(-7, 7, nzp)), dev(_axes(rng, nxp, 10., 2.)), dev(_axes(rng, nyp, -20., 3.))
    z, x, y = dev(np.sort(rng.uniform(-6, 6, 7))), dev(np.sort(rng.uniform(11, 33, 6))), dev(
        np.sort(rng.uniform(-18, 8, 5)))
    srcs = [dev(rng.uniform(0.5, 2., (nzp, nyp, ld))) for _ in range(5)]
    zlogs = [1, 0, 1, 1, 0]
    dk, di, dj = 6 * 5, 1, 6
    one = [torch.zeros(7 * 6 * 5 + 3, dtype=torch.float64, device="cuda") for _ in srcs]
    bat = [torch.zeros_like(t) for t in one]
    for sr, o, zl in zip(srcs, one, zlogs):
        pkg.tgx.interp3d_dev(sr, zp, xp, yp, o, z, x, y, dk, di, dj, zl, offset=3)
    pkg.tgx.interp3d_batch_dev(srcs, [(t, 3) for t in bat], zlogs, zp, xp, yp, z, x, y, dk, di, dj)
    planes = [sr[0] for sr in srcs]
    one2 = [torch.zeros(6 * 5, dtype=torch.float64, device="cuda") for _ in planes]
    bat2 = [torch.zeros_like(t) for t in one2]
    for sr, o in zip(planes, one2):
        pkg.tgx.interp2d_dev(sr, xp, yp, o, x, y, 1, 6)
    pkg.tgx.interp2d_batch_dev(planes, bat2, xp, yp, x, y, 1, 6)
    f0 = [dev(rng.normal(size=n)) for n in (40, 7, 123)]
    f1 = [dev(rng.normal(size=n)) for n in (40, 7, 123)]
    one3 = [torch.zeros(3 * 40, dtype=torch.float64, device="cuda"), torch.zeros(7, dtype=torch.float64,
                                                                                 device="cuda"),
            torch.zeros(123, dtype=torch.float64, device="cuda")]
    bat3 = [torch.zeros_like(t) for t in one3]
    pkg.tgx.time_interp_dev(f0[0], f1[0], one3[0], 10, 4, 6)
    pkg.tgx.time_interp_dev(f0[1], f1[1], one3[1], 5, 5, 5)
    one3[2].copy_(f1[2])
    pkg.tgx.time_interp_batch_dev([(f0[0], f1[0], bat3[0], 10, 4, 6), (f0[1], f1[1], bat3[1], 5, 5, 5),
                                   (None, f1[2], bat3[2], 0, 0, 0)])
    torch.cuda.synchronize()
    for a, b in zip(one + one2 + one3, bat + bat2 + bat3):
        assert torch.equal(a, b)


def test_coupled_nests_match_oracle_recursion(pkg, oracle):
    """small (16x12x29) with nest_s1 (20x12x29, 2 sub-steps) inside and
    nest_s2 (10x8x57, 3 sub-steps, another level set) inside that: two outer
    steps of the full recursion on the GPU against the oracle's."""
    import torch
    from nestref import OracleNests
    names, nsteps = ("small", "nest_s1", "nest_s2"), (2, 2, 3)
    grids = [pkg.synth.make_grid(n, nstep_sub=2 if i == 0 else 1, opdiffcap=0. if i == 0 else 1e8)
             for i, n in enumerate(names)]
    states = [pkg.synth.make_state(g) for g in grids]
    print("# coupled nests", names, nsteps, flush=True)
    dev = pkg.nest.CoupledNests(list(zip(grids, states)), nsteps)
    dev.run()
    torch.cuda.synchronize()
    ref = OracleNests(oracle, pkg, [(g, pkg.synth.to_fortran_layout(F)) for g, F in zip(grids, states)], nsteps)
    ref.run()
    worst = {}
    for l, name in enumerate(names):
        for n in pkg.step.PROGNOSTIC:
            got = dev.states[l].fortran(n)
            if got.ndim != 3:
                continue
            fe, pe = field_errs(n, got, ref.states[l][n])
            worst["%s.%s" % (name, n)] = max(fe, pe)
    bad = {k: v for k, v in worst.items() if not v <= 1e-11}
    print("worst %.3e" % max(worst.values()))
    assert not bad, bad
    # the nests changed the outer grids: map_out wrote the inner states back
    box = pkg.nest.mapout_box(grids[0], grids[1])
    assert box is not None


def test_coupled_nests_at_configured_sizes_graph_and_oracle(pkg, oracle):
    """The configured nest batch, coupled: the 2.5 deg grid (144x72x57), nest 1
    (120x60 at 0.5 deg, 2 sub-steps) and nest 2 (100x50 at 0.1 deg, 5 sub-steps
    per nest-1 step), the nests' synthetic states interpolated from the global
    one.  One outer step eager against the oracle's recursion (pointwise
    1e-11); two outer steps replayed from HIP graphs bitwise equal to two eager
    ones, every field finite."""
    import torch
    from nestref import OracleNests
    names, nsteps = ("g2p5", "nest1", "nest2"), (1, 2, 5)
    grids = [pkg.synth.make_grid(n, nstep_sub=1, opdiffcap=1e8 if n.startswith("nest") else 0.) for n in names]
    states = pkg.synth.nested_states(grids)
    print("# coupled nests at size", flush=True)
    eager = pkg.nest.CoupledNests(list(zip(grids, states)), nsteps)
    eager.run()
    torch.cuda.synchronize()
    ref = OracleNests(oracle, pkg, [(g, pkg.synth.to_fortran_layout(F)) for g, F in zip(grids, states)], nsteps)
    ref.run()
    worst = 0.
    for l in range(3):
        for n in pkg.step.PROGNOSTIC:
            got = eager.states[l].fortran(n)
            if got.ndim == 3:
                worst = max(worst, max(field_errs(n, got, ref.states[l][n])))
    print("oracle worst %.3e" % worst, flush=True)
    assert worst <= 1e-11, worst
    graph = pkg.nest.CoupledNests(list(zip(grids, states)), nsteps)
    graph.capture()            # one eager outer step, then the graphs
    graph.replay()             # the second outer step
    eager.run()
    torch.cuda.synchronize()
    for l in range(3):
        for n in pkg.step.PROGNOSTIC:
            got, want = graph.states[l].current(n), eager.states[l].current(n)
            assert torch.isfinite(want).all(), (names[l], n)
            assert torch.equal(got, want), (names[l], n)
