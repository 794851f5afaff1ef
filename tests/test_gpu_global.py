"""GPU parity of the global-grid forms (include/tgx_global.h), through the
C-ABI, against the oracle's global forms (tests/test_oracle_global.py pins
them): dt.F, duv.F, minor.F with the species wrappers, oplus.F with its
sub-cycles, comp_o2o.F + comp.F, each with its ring filter.  Every output is
compared on the interior latitudes (all longitudes, the periodic points
included), pointwise: |got - ref| / max(|ref|, floor) <= 1e-12 with the
floors of helpers.POINT_FLOORS, below the top level where that is spval (TN,
UN, VN) or zero (O+), which must then match exactly."""
import numpy as np
import pytest

from helpers import field_errs, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

TOL = 1e-12


def interior(g, a):
    return a[2:g["nlat"] + 2]


# Fields whose top level is spval (1e36) or zero on the global grid: their
# errors are measured below the top, which the tests check exactly (a 1e36
# value would otherwise set the scale of the field and of every wind column).
FIXED_TOP = ("tn", "tn_nm", "un", "un_nm", "vn", "vn_nm", "op", "op_nm")


def compare(g, ns, ref, names, label):
    worst = 0.
    for n in names:
        got, want = interior(g, ns.fortran(n)), interior(g, ref[n])
        if n in FIXED_TOP:
            assert np.array_equal(got[..., -1], want[..., -1]), (label, n)
            got, want = got[..., :-1], want[..., :-1]
        fe, pe = field_errs(n, got, want)
        print("%s %-7s field %.3e pointwise %.3e" % (label, n, fe, pe), flush=True)
        assert fe <= TOL and pe <= TOL, (label, n, fe, pe)
        worst = max(worst, pe)
    return worst


OUT = ("op", "op_nm", "xiop2p", "xiop2d", "Fe", "Fn", "n2d", "n2d_nm", "ar", "ar_nm", "n4s", "n4s_nm", "no",
       "no_nm", "tn", "tn_nm", "un", "un_nm", "vn", "vn_nm", "o2", "o2_nm", "o1", "o1_nm", "he", "he_nm")


# (configuration, O+ sub-cycles): the 5, 2.5 and 1.25 degree grids of the
# benchmark configurations, small grids with filtered latitudes next to both
# poles, and sub-cycled O+ (its rows across the poles held between sub-steps)
@pytest.fixture(scope="module", params=[("tiny", 2), ("small", 1), ("small", 3), ("tall", 1), ("g5", 1),
                                        ("g2p5", 1), ("g2p5", 10), ("g1p25", 1)],
                ids=lambda p: "%s-nsub%d" % p)
def case(request, pkg, oracle):
    import torch
    name, nsub = request.param
    g = pkg.synth.make_grid(name, nstep_sub=nsub, glb=True)
    F = pkg.synth.make_global_state(g)
    Fo = pkg.synth.to_fortran_layout(F)
    print("# global step %s nsub=%d" % (name, nsub), flush=True)
    ref = pkg.step.run_oracle_step(oracle, g, Fo)
    ns = pkg.step.GlobalState(g, F)
    ns.step()
    torch.cuda.synchronize()
    return name, g, ns, ref


def test_global_step_matches_oracle(case):
    name, g, ns, ref = case
    compare(g, ns, ref, OUT, name)


def test_global_step_2d_bookkeeping(case):
    name, g, ns, ref = case
    for n in ("tlbc", "tlbc_nm", "ulbc", "vlbc", "ulbc_nm", "vlbc_nm", "flx_he"):
        assert np.array_equal(interior(g, ns.fortran(n)), interior(g, ref[n])), n


def test_global_top_levels_and_periodic_points(case):
    name, g, ns, ref = case
    nlon = g["nlon"]
    for n, v in (("tn", 1e36), ("tn_nm", 1e36), ("un", 1e36), ("vn", 1e36), ("un_nm", 1e36), ("vn_nm", 1e36),
                 ("op", 0.), ("op_nm", 0.)):
        assert np.all(interior(g, ns.fortran(n))[..., -1] == v), n
    for n in OUT:
        a = interior(g, ns.fortran(n))
        for ip, isrc in ((0, nlon), (1, nlon + 1), (nlon + 2, 2), (nlon + 3, 3)):
            assert np.array_equal(a[:, ip], a[:, isrc]), (n, ip)


def test_global_forked_step_equals_serial(pkg):
    """The forked multi-stream global step and the single-stream reference
    order give bitwise the same outputs, also when called twice in a row."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2, glb=True)
    F = pkg.synth.make_global_state(g)
    a = pkg.step.GlobalState(g, F)
    b = pkg.step.GlobalState(g, F)
    a.step_serial()
    b.step()
    b.step()
    torch.cuda.synchronize()
    rows = slice(2, g["nlat"] + 2)
    for n in a.out:
        x, y = a.out[n], b.out[n]
        if x.dim() == 3:
            x, y = x[:, rows], y[:, rows]
        else:
            x, y = x[rows], y[rows]
        assert torch.equal(x, y), n


def test_global_routines_one_at_a_time_match_oracle(pkg, oracle):
    """tgx_glb_oplus, _minors, _dt, _duv and _comp called one by one in the
    model's order (each with its own pointwise chemistry) agree with the
    oracle."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2, glb=True)
    F = pkg.synth.make_global_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    ns = pkg.step.GlobalState(g, F)
    for r in ("oplus", "minors", "dt", "duv", "comp"):
        getattr(ns, r)()
    torch.cuda.synchronize()
    compare(g, ns, ref, OUT, "routines")


def test_global_steps_chained_with_halos(pkg, oracle):
    """Three global steps, each step's prognostic outputs the next one's
    inputs after tgx_glb_halos (the device) or synth.fill_global_halos (the
    oracle's inputs) refills their periodic points and pole rows."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2, glb=True)
    F = pkg.synth.make_global_state(g)
    Fo = pkg.synth.to_fortran_layout(F)
    ns = pkg.step.GlobalState(g, F)
    ns.enable_feedback()
    import pinglb
    res = None
    for step in range(3):
        ns.step()
        ns.fill_halos()
        res = pkg.step.run_oracle_step(oracle, g, Fo)
        Fo = dict(Fo)
        for n in pkg.step.PROGNOSTIC:
            Fo[n] = res[n].copy()
        pinglb.fill_halos(g, Fo)
    torch.cuda.synchronize()
    for n in ("tn", "un", "vn", "op", "n2d", "no", "o2", "o1", "he"):
        fe, pe = field_errs(n, interior(g, ns.fortran(n))[..., :-1], interior(g, Fo[n])[..., :-1])
        print("chained %-4s field %.3e pointwise %.3e" % (n, fe, pe))
        assert fe <= TOL and pe <= TOL, (n, fe, pe)


def test_glb_halos_kernel_matches_the_pole_rule(pkg):
    import torch
    g = pkg.synth.make_grid("small", glb=True)
    nk, nlat, nlon = 5, g["nlat"], g["nlon"]
    gen = torch.Generator().manual_seed(7)
    t = torch.randn(nk, nlat + 4, nlon + 4, generator=gen, dtype=torch.float64)
    want = pkg.synth.fill_global_halos({"un": t.clone()}, g)["un"]
    ld = pkg.step.pitch(nlon + 4)
    d = torch.zeros(nk, nlat + 4, ld, dtype=torch.float64, device="cuda")
    d[..., :nlon + 4] = t.cuda()
    rc = pkg.step._lib().tgx_glb_halos(d.data_ptr(), nk, nlon, nlat, ld, -1.0, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(d[..., :nlon + 4].cpu(), want)


def test_global_full_size_step_matches_oracle(pkg, oracle):
    """At the benchmark's size (0.625 deg global: 576x288 columns x 113
    levels, 40 filtered latitudes next to each pole), where the large-grid
    forms run and the column kernels apply the post pass themselves on the
    latitudes the ring filter does not reach: every output of the global step
    matches the oracle's global forms run over the whole grid (pointwise, the
    floors of helpers.POINT_FLOORS), the forked step equals the serial step
    bit for bit, the top levels are spval / zero, the periodic points repeat
    the interior, and the solution keeps the reference's floors and
    normalisation (comp.F:1000-1022, minor.F:590-597, dt.F:418-424,
    oplus.F:930-934)."""
    import torch
    g = pkg.synth.make_grid("g0p625", nstep_sub=1, glb=True)
    F = pkg.synth.make_global_state(g, device="cuda")
    print("oracle global step over 576x288x113", flush=True)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    a = pkg.step.GlobalState(g, F)
    a.step_serial()
    torch.cuda.synchronize()
    rows = slice(2, g["nlat"] + 2)
    serial = {n: (t[:, rows] if t.dim() == 3 else t[rows]).clone() for n, t in a.out.items()}
    del a
    ns = pkg.step.GlobalState(g, F)
    ns.step()
    torch.cuda.synchronize()
    for n, t in serial.items():
        assert torch.equal(t, ns.out[n][:, rows] if ns.out[n].dim() == 3 else ns.out[n][rows]), n
    del serial
    worst = compare(g, ns, ref, OUT, "g0p625")
    print("worst pointwise %.3e" % worst, flush=True)
    del ref
    nlon = g["nlon"]
    o = {n: interior(g, ns.fortran(n)) for n in OUT}
    for n in ("tn", "tn_nm", "un", "vn", "un_nm", "vn_nm"):
        assert np.all(o[n][..., -1] == 1e36), n
    for n in ("op", "op_nm"):
        assert np.all(o[n][..., -1] == 0.), n
    for n, a in o.items():
        for ip, isrc in ((0, nlon), (1, nlon + 1), (nlon + 2, 2), (nlon + 3, 3)):
            assert np.array_equal(a[:, ip], a[:, isrc]), (n, ip)
        assert np.all(np.isfinite(a)), n
    for sfx in ("", "_nm"):
        s3 = o["o2" + sfx] + o["o1" + sfx] + o["he" + sfx]
        assert (o["o2" + sfx] > 0).all() and (o["o1" + sfx] > 0).all() and (o["he" + sfx] > 0).all()
        assert (s3 < 1).all(), sfx
        for n in ("no", "n4s", "n2d", "ar"):
            assert (o[n + sfx] >= 1e-12).all(), n + sfx
        assert (o["op" + sfx][..., :-1] >= 0).all()
    assert (o["tn"][..., :-1] >= 100).all()


@pytest.mark.parametrize("nsub", [1, 2])
def test_global_step_without_filtered_latitudes(pkg, oracle, nsub):
    """A global grid whose chunk table filters no latitude (nlat_filter = 0):
    the column kernels then apply every routine's post pass on every row
    and no ring filter runs; the step matches the oracle's global forms with
    the same empty table, and the forked step equals the serial one."""
    import torch
    g = pkg.synth.make_grid("g5", nstep_sub=nsub, glb=True)
    g["chunk"] = []
    F = pkg.synth.make_global_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    ns = pkg.step.GlobalState(g, F)
    ns.step()
    ser = pkg.step.GlobalState(g, F)
    ser.step_serial()
    torch.cuda.synchronize()
    compare(g, ns, ref, OUT, "g5-nofilter")
    rows = slice(2, g["nlat"] + 2)
    for n in ns.out:
        a, b = ns.out[n], ser.out[n]
        if a.dim() == 3:
            a, b = a[:, rows], b[:, rows]
        else:
            a, b = a[rows], b[rows]
        assert torch.equal(a, b), n
