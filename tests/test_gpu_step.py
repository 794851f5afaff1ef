"""GPU parity of the fused column step, through the C-ABI, against the oracle.

Each reference routine (oplus_ng with its sub-cycles, the four minor species,
dt_ng, duv_ng, comp_o2o_ng + comp_ng) runs on device-resident synthetic state
and is compared field by field with the oracle, pointwise:
|got - ref| / max(|ref|, floor) <= 1e-12 at every point, with the floors of
helpers.POINT_FLOORS (the reference's own clamps).  The dt and duv Thomas
kernels follow the reference's operation order; the minor-species and comp
kernels take the temperature powers of their diffusion coefficients from one
logarithm and divide by dz and by the 3x3 determinant through reciprocals,
and the PCR scheme reorders the elimination, so a few ulp per term remain.
"""
import numpy as np
import pytest

from helpers import field_errs, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

TOL = 1e-12


def field_err(got, ref, name=None):
    """The larger of the field-normalised and the pointwise relative error
    (the pointwise one needs the field's name for its floor)."""
    if name is None:
        return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))
    return max(field_errs(name, got, ref))


# (grid, O+ sub-cycles, column-solve scheme): "thomas" follows the reference's
# operation order; "pcr" (one wavefront per column, cyclic reduction) reorders
# the elimination and is held to the same 1e-12 bound.
@pytest.fixture(scope="module", params=[("tiny", 2, "thomas"), ("small", 1, "thomas"), ("small", 3, "thomas"),
                                        ("tall", 1, "thomas"), ("tiny", 2, "pcr"), ("small", 1, "pcr"),
                                        ("tall", 1, "pcr"),
                                        # the 5, 2.5 and 1.25 degree global grids and both nests at
                                        # full size
                                        ("g5", 1, "thomas"), ("g2p5", 1, "thomas"), ("g2p5", 1, "pcr"),
                                        ("g1p25", 1, "thomas"), ("nest1", 2, "thomas"), ("nest2", 1, "pcr"),
                                        # the reference's default of ten O+ sub-cycles (the Thomas
                                        # scheme factors the O+ matrix once for all ten)
                                        ("small", 10, "thomas"), ("tall", 10, "thomas"), ("g2p5", 10, "pcr")],
                ids=lambda p: "%s-nsub%d-%s" % p)
def case(request, pkg, oracle):
    import torch
    name, nsub, scheme = request.param
    g = pkg.synth.make_grid(name, nstep_sub=nsub)
    F = pkg.synth.make_state(g)
    Fo = pkg.synth.to_fortran_layout(F)
    ref = pkg.step.run_oracle_step(oracle, g, Fo)
    ns = pkg.step.NestState(g, F, scheme=scheme)
    ns.step()
    torch.cuda.synchronize()
    return name, ns, ref


@pytest.mark.parametrize("fields", [
    ("op", "op_nm", "xiop2p", "xiop2d", "Fe", "Fn"),
    ("n2d", "n2d_nm", "ar", "ar_nm"),
    ("n4s", "n4s_nm", "no", "no_nm"),
    ("tn", "tn_nm"),
    ("un", "un_nm", "vn", "vn_nm"),
    ("o2", "o2_nm", "o1", "o1_nm", "he", "he_nm"),
], ids=["oplus", "minors_n2d_ar", "minors_n4s_no", "dt", "duv", "comp"])
def test_step_fields_match_oracle(case, fields):
    name, ns, ref = case
    for n in fields:
        fe, pe = field_errs(n, ns.fortran(n), ref[n])
        print("%s %-7s field %.3e pointwise %.3e" % (name, n, fe, pe))
        assert fe <= TOL and pe <= TOL, (name, n, fe, pe)


def test_step_2d_bookkeeping(case):
    name, ns, ref = case
    for n in ("tlbc", "tlbc_nm", "ulbc", "vlbc", "ulbc_nm", "vlbc_nm", "flx_he"):
        assert np.array_equal(ns.fortran(n), ref[n]), n


def test_concurrent_step_equals_serial(pkg):
    """The forked multi-stream step and the single-stream reference order give
    bitwise the same outputs, also when called twice in a row."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2)
    F = pkg.synth.make_state(g)
    a = pkg.step.NestState(g, F)
    b = pkg.step.NestState(g, F)
    a.step_serial()
    b.step()
    b.step()
    torch.cuda.synchronize()
    for n in a.out:
        assert torch.equal(a.out[n], b.out[n]), n


@pytest.mark.parametrize("name,nsub", [("small", 2), ("tall", 1)])
def test_routines_one_at_a_time_match_oracle(pkg, oracle, name, nsub):
    """The five routines called one by one (each with its own pointwise
    chemistry, as a model replacing them one at a time does) agree with the
    oracle, as the fused step does with its single chemistry pass."""
    import torch
    g = pkg.synth.make_grid(name, nstep_sub=nsub)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    ns = pkg.step.NestState(g, F)
    for r in ("oplus", "minors", "dt", "duv", "comp"):
        getattr(ns, r)()
    torch.cuda.synchronize()
    for n in ("tn", "un", "vn", "op", "xiop2p", "xiop2d", "no", "n4s", "n2d", "ar", "o2", "o1", "he"):
        err = field_err(ns.fortran(n), ref[n], n)
        assert err <= TOL, (n, err)


@pytest.mark.parametrize("scheme,calc_helium", [("thomas", 1), ("pcr", 1), ("thomas", 0)])
def test_comp_n2_extrapolation_matches_oracle(pkg, oracle, scheme, calc_helium):
    """A thin N2 top (N2 < 0.01 crossed inside every interior column) runs the
    post pass's N2 extrapolation above the crossing (comp_ng.F90:419-431)."""
    import torch
    g = pkg.synth.make_grid("tall", nstep_sub=1, n2_top=0.003, calc_helium=calc_helium)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    n2 = 1. - ref["o2"] - ref["o1"] - ref["he"]
    assert (n2[2:-2, 2:-2, -2:] < 0.01).all()
    for fn in ("step", "step_serial"):
        ns = pkg.step.NestState(g, F, scheme=scheme)
        getattr(ns, fn)()
        torch.cuda.synchronize()
        for n in ("o2", "o1", "he", "o2_nm", "o1_nm", "he_nm"):
            err = field_err(ns.fortran(n), ref[n], n)
            assert err <= TOL, (fn, n, err)


ALL_FIELDS = ("op", "op_nm", "xiop2p", "xiop2d", "Fe", "Fn", "n2d", "n2d_nm", "ar", "ar_nm", "n4s", "n4s_nm",
              "no", "no_nm", "tn", "tn_nm", "un", "un_nm", "vn", "vn_nm", "o2", "o2_nm", "o1", "o1_nm", "he",
              "he_nm")


@pytest.mark.parametrize("scheme", ["thomas", "pcr"])
def test_step_run_switches_match_oracle(pkg, oracle, scheme):
    """The run switches away from their defaults: no helium solve
    (calc_helium=0), the O+ diffusion cap binding on the upper levels
    (opdiffcap=1e8, oplus_ng.F90:145-148), no horizontal-diffusion latitude
    factor (difhor=0), another O-N2 collision factor, two O+ sub-cycles."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2, calc_helium=0, colfac=1.7, opdiffcap=1e8, difhor=0)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    for fn in ("step", "step_serial"):
        ns = pkg.step.NestState(g, F, scheme=scheme)
        getattr(ns, fn)()
        torch.cuda.synchronize()
        for n in ALL_FIELDS:
            err = field_err(ns.fortran(n), ref[n], n)
            assert err <= TOL, (fn, n, err)


def test_auto_scheme_tunes_and_matches_oracle(pkg, oracle):
    """TGX_SCHEME_AUTO times both schemes on its first step, records a choice
    per routine, and its results stay within the FP64 bound."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=1)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    ns = pkg.step.NestState(g, F, scheme="auto")
    ns.step()
    ns.step()
    torch.cuda.synchronize()
    assert set(ns.schemes().values()) <= {"thomas", "pcr"}
    assert set(ns.wave_columns().values()) <= {1, 2, 4}
    for n in ("tn", "un", "vn", "op", "no", "n4s", "n2d", "ar", "o2", "o1", "he"):
        assert field_err(ns.fortran(n), ref[n], n) <= TOL, n


@pytest.mark.parametrize("cols", [1, 2])
def test_wave_kernels_with_fewer_columns_per_block_equal_four(pkg, cols):
    """The wavefront-per-column kernels launched with 1 or 2 columns per block
    (the autotuner's other candidates) give bitwise the step of 4 per block."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2)
    F = pkg.synth.make_state(g)
    a = pkg.step.NestState(g, F, scheme="pcr")
    a.step()
    torch.cuda.synchronize()
    b = pkg.step.NestState(g, F, scheme="pcr")
    pkg.step._lib().tgx_ng_force_wave_columns(cols)
    try:
        b.step()  # every routine's wave kernels launched with `cols` per block
    finally:
        pkg.step._lib().tgx_ng_force_wave_columns(0)
    torch.cuda.synchronize()
    for n in a.out:
        assert torch.equal(a.out[n], b.out[n]), n


@pytest.mark.parametrize("name,glb", [("small", False), ("nest1", False), ("g2p5", True)])
def test_dt_tiles_loaded_straight_into_lds_equal_register_staging(pkg, name, glb):
    """The dt tile kernel with its tiles loaded from HBM straight into LDS
    (the default where TN, UN, VN, TN_NM are 16-byte aligned and the row pitch
    is even) gives bitwise the step of the form that stages them through
    registers."""
    import torch
    g = pkg.synth.make_grid(name, nstep_sub=2, glb=glb)
    if glb:
        st = [pkg.step.GlobalState(g, pkg.synth.make_global_state(g)) for _ in range(2)]
    else:
        F = pkg.synth.make_state(g)
        st = [pkg.step.NestState(g, F, scheme="thomas") for _ in range(2)]
    a, b = st
    assert a.ld % 2 == 0 and all(a.inp[n].data_ptr() % 16 == 0 for n in ("tn", "un", "vn", "tn_nm"))
    a.step()
    torch.cuda.synchronize()
    pkg.step._lib().tgx_ng_set_tile_dma(0)
    try:
        b.step()
    finally:
        pkg.step._lib().tgx_ng_set_tile_dma(1)
    torch.cuda.synchronize()
    for n in a.out:
        assert torch.equal(a.out[n], b.out[n]), n


def test_profile_hooks_count_calls(pkg):
    import torch
    g = pkg.synth.make_grid("tiny", nstep_sub=1)
    ns = pkg.step.NestState(g, pkg.synth.make_state(g))
    pkg.step.profile_begin(2)
    for _ in range(3):
        ns.step()
    torch.cuda.synchronize()
    prof, kern = pkg.step.profile_end()
    for r in pkg.step.ROUTINES:
        ms, n = prof[r]
        assert n == 2 and ms > 0.0, (r, prof[r])
        km, kn = kern[r]
        assert kn == 2 and 0.0 < km <= ms, (r, kern[r], prof[r])


def _latest(ns):
    """Every domain's latest prognostic fields (clones)."""
    import sys
    names = sys.modules[type(ns).__module__].PROGNOSTIC
    return [{n: st.current(n).clone() for n in names} for st in ns.states]


@pytest.mark.parametrize("forked", [False, True], ids=["serial", "forked"])
def test_nested_outer_step_graph_replay_matches_eager(pkg, forked):
    """The global grid and two nests stepped in the reference's recursion
    (level i runs prod(nstep_ng[:i+1]) sub-steps), every sub-step advancing
    its domain's state, captured into HIP graphs: each replay reproduces the
    next eager outer step bit for bit (including the domains that take an
    odd number of sub-steps, which alternate buffer sets), with unrelated
    eager work and synchronisation between the replays."""
    import torch
    doms = []
    for name, nsub in (("small", 1), ("tiny", 2), ("tall", 1)):
        g = pkg.synth.make_grid(name, nstep_sub=nsub)
        doms.append((g, pkg.synth.make_state(g)))
    eager = pkg.step.NestedStep(doms, [1, 2, 3], forked=forked)
    assert eager.substeps() == {0: 1, 1: 2, 2: 6}
    ref = []
    for _ in range(4):
        eager.run()
        torch.cuda.synchronize()
        ref.append(_latest(eager))
    # consecutive outer steps differ: the state advances
    assert not torch.equal(ref[0][0]["tn"], ref[1][0]["tn"])
    ns = pkg.step.NestedStep(doms, [1, 2, 3], forked=forked)
    ns.capture()           # its warm-up is outer step 1
    other = pkg.step.NestState(doms[1][0], doms[1][1])
    for rep in range(3):
        ns.replay()        # outer steps 2, 3, 4
        torch.cuda.synchronize()
        got = _latest(ns)
        for d in range(len(doms)):
            for n, t in got[d].items():
                assert torch.equal(t, ref[rep + 1][d][n]), (rep, d, n)
        other.step()       # eager work between replays
        torch.cuda.synchronize()


def test_nested_batch_at_its_sizes_matches_chained_oracle(pkg, oracle):
    """The nested configuration at its real sizes: the 2.5 deg global grid
    (144x72), nest 1 (120x60, NSTEP_NG=2) and nest 2 (100x50, NSTEP_NG=5 per
    nest-1 step), 57 levels, one outer step in the reference's recursion
    (1, 2 and 10 sub-steps), each sub-step advancing its domain's state.
    Every domain's state after its last sub-step matches the oracle stepped
    the same number of times (pointwise, the floors of helpers.POINT_FLOORS),
    and a graph replay of the next outer step equals the eager one.  The nests
    cap the O+ diffusion coefficient (opdiffcap, oplus_ng.F90:145-148): at
    0.1 deg the explicit field-aligned O+ diffusion of the synthetic state is
    unstable without it and the chained state runs away within ten steps."""
    import torch
    doms = []
    for name in ("g2p5", "nest1", "nest2"):
        g = pkg.synth.make_grid(name, nstep_sub=2, opdiffcap=1e8 if name.startswith("nest") else 0.)
        doms.append((g, pkg.synth.make_state(g)))
    ns = pkg.step.NestedStep(doms, [1, 2, 5], forked=True)
    sub = ns.substeps()
    assert sub == {0: 1, 1: 2, 2: 10}
    ns.run()
    torch.cuda.synchronize()
    fields = ("tn", "un", "vn", "op", "n2d", "n4s", "no", "ar", "o2", "o1", "he")
    for d, (g, F) in enumerate(doms):
        print("domain %s: %d sub-steps" % (g["name"], sub[d]), flush=True)
        ref = pkg.step.run_oracle_steps(oracle, g, pkg.synth.to_fortran_layout(F), sub[d])
        st = ns.states[d]
        for n in fields + tuple(f + "_nm" for f in fields):
            fe, pe = field_errs(n, st.fortran(n), ref[n])
            assert fe <= TOL and pe <= 1e-11, (g["name"], n, fe, pe)
    # outer step 2: eager on one copy, graph replay on another
    eager = _latest(ns)
    ns.run()
    torch.cuda.synchronize()
    step2 = _latest(ns)
    ns2 = pkg.step.NestedStep(doms, [1, 2, 5], forked=True)
    ns2.capture()
    ns2.replay()
    torch.cuda.synchronize()
    got = _latest(ns2)
    for d in range(3):
        for n in got[d]:
            assert torch.equal(got[d][n], step2[d][n]), (d, n)
    del eager


def test_step_from_two_host_threads(pkg):
    """Two host threads stepping two domains at once, each on its own torch
    stream, give the outputs of the same steps taken one after the other."""
    import threading
    import torch
    doms = []
    for name in ("small", "tiny"):
        g = pkg.synth.make_grid(name, nstep_sub=2)
        doms.append((g, pkg.synth.make_state(g)))
    ref = [pkg.step.NestState(g, F) for g, F in doms]
    for st in ref:
        st.step()
    torch.cuda.synchronize()
    par = [pkg.step.NestState(g, F) for g, F in doms]
    errs = []

    def work(st):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(3):
                    st.step(stream=s.cuda_stream)
            s.synchronize()
        except Exception as e:  # reported in the main thread
            errs.append(e)

    th = [threading.Thread(target=work, args=(st,)) for st in par]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    torch.cuda.synchronize()
    for a, b in zip(ref, par):
        for n in a.out:
            assert torch.equal(a.out[n], b.out[n]), n


@pytest.mark.parametrize("scheme", ["thomas", "pcr"])
def test_small_grid_forms_equal_large_grid_forms(pkg, oracle, monkeypatch, scheme):
    """On grids under 65,536 columns the chemistry pass and O+ stage 1 walk
    runs of two levels, O+ stage 2 single levels and duv's stage 1 one
    wavefront per column; the large-grid forms (TGX_LARGE_GRID_FORMS=1: runs
    of about eight levels, whole columns, one thread per column) give bitwise
    the same step, within the FP64 bound of the oracle."""
    import torch
    g = pkg.synth.make_grid("small", nstep_sub=2)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    a = pkg.step.NestState(g, F, scheme=scheme)
    a.step()
    torch.cuda.synchronize()
    monkeypatch.setenv("TGX_LARGE_GRID_FORMS", "1")
    b = pkg.step.NestState(g, F, scheme=scheme)
    b.step()
    torch.cuda.synchronize()
    for n in a.out:
        assert torch.equal(a.out[n], b.out[n]), n
    for n in ALL_FIELDS:
        assert field_err(b.fortran(n), ref[n], n) <= TOL, n


def test_full_size_step_properties(pkg, oracle):
    """At the benchmark's size (0.625 deg: 576x288 columns x 113 levels), where
    the large-grid forms run: every field of the step matches the oracle run
    over the whole grid (pointwise, the floors of helpers.POINT_FLOORS), the
    forked step equals the serial step bit for bit, and the solution keeps the
    reference's floors and normalisation (comp_ng.F90:415-457,
    minor_ng.F90:234-237, dt_ng.F90:208-213, oplus_ng.F90:337-340): O2, O, He
    positive with O2 + O + He < 1, the minor species >= 1e-12, TN >= 100,
    O+ >= 0, every value finite."""
    import torch
    g = pkg.synth.make_grid("g0p625", nstep_sub=1)
    F = pkg.synth.make_state(g, device="cuda")
    print("oracle step over 576x288x113", flush=True)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    a = pkg.step.NestState(g, F, scheme="thomas")
    a.step_serial()
    torch.cuda.synchronize()
    serial = {n: t.clone() for n, t in a.out.items()}
    del a
    b = pkg.step.NestState(g, F, scheme="thomas")
    b.step()
    torch.cuda.synchronize()
    for n in serial:
        assert torch.equal(serial[n], b.out[n]), n
    worst = {}
    for n in ALL_FIELDS:
        fe, pe = field_errs(n, b.fortran(n), ref[n])
        worst[n] = max(fe, pe)
    print({n: "%.1e" % e for n, e in worst.items()})
    assert max(worst.values()) <= TOL, worst
    del ref
    nl = g["nlon"] + 4
    o = {n: t[..., :nl] for n, t in b.out.items()}
    for n, t in o.items():
        assert torch.isfinite(t).all(), n
    for sfx in ("", "_nm"):
        s3 = o["o2" + sfx] + o["o1" + sfx] + o["he" + sfx]
        assert (o["o2" + sfx] > 0).all() and (o["o1" + sfx] > 0).all() and (o["he" + sfx] > 0).all()
        assert (s3 < 1).all(), sfx
        for n in ("no", "n4s", "n2d", "ar"):
            assert (o[n + sfx] >= 1e-12).all(), n + sfx
        assert (o["op" + sfx] >= 0).all()
    assert (o["tn"] >= 100).all()


@pytest.mark.parametrize("name,nsub", [("small", 2), ("tall", 1)])
def test_assembled_rows_match_the_pinned_oracle_rows(pkg, oracle, name, nsub):
    """The rows every Thomas column kernel hands its elimination (capture
    mode: dt, O+ of the last sub-step, the four minor species, the 2x2 wind
    blocks, the 3x3 O2/O/He blocks and their right-hand sides) against the
    rows of the oracle, whose assembly tests/test_oracle_pins.py holds to the
    model description, conservation and symmetry.  Each plane is compared per
    column, relative to the column's largest magnitude in that plane."""
    import torch
    g = pkg.synth.make_grid(name, nstep_sub=nsub)
    F = pkg.synth.make_state(g)
    Fo = pkg.synth.to_fortran_layout(F)
    nk, nlatd, nlond = g["nk"], g["nlat"] + 4, g["nlon"] + 4
    buf = np.zeros(oracle.CAPTURE_PLANES * nlatd * nlond * nk)
    oracle.set_capture(buf)
    try:
        pkg.step.run_oracle_step(oracle, g, Fo)
    finally:
        oracle.set_capture(None)
    ref = buf.reshape(oracle.CAPTURE_PLANES, nlatd, nlond, nk)[:, 2:-2, 2:-2]
    ns = pkg.step.NestState(g, F, scheme="thomas", capture=True)
    ns.step()
    torch.cuda.synchronize()
    got = ns.capture.cpu().numpy().reshape(oracle.CAPTURE_PLANES, nlatd, nlond, nk)[:, 2:-2, 2:-2]
    worst = {}
    for pl in range(oracle.CAPTURE_PLANES):
        rows = nk if 8 <= pl < 24 else nk - 1          # minor species solve every level
        a, b = got[pl, ..., :rows], ref[pl, ..., :rows]
        scale = np.abs(b).max(axis=-1, keepdims=True)
        scale[scale == 0] = 1.
        worst[pl] = float((np.abs(a - b) / scale).max())
    print({pl: "%.1e" % e for pl, e in worst.items() if e > 1e-15})
    bad = {pl: e for pl, e in worst.items() if e > TOL}
    assert not bad, bad


@pytest.mark.parametrize("name,nsub", [("small", 10), ("tall", 3), ("small", 2)])
def test_oplus_kept_factorisation_equals_full_assembly(pkg, name, nsub):
    """With several O+ sub-cycles the Thomas scheme factors the O+ matrix in
    the first sub-step and later sub-steps rebuild only the right-hand side;
    capture mode assembles every sub-step in full.  Both give bitwise the
    same step."""
    import torch
    g = pkg.synth.make_grid(name, nstep_sub=nsub)
    F = pkg.synth.make_state(g)
    a = pkg.step.NestState(g, F, scheme="thomas")
    b = pkg.step.NestState(g, F, scheme="thomas", capture=True)
    a.step()
    b.step()
    torch.cuda.synchronize()
    for n in a.out:
        assert torch.equal(a.out[n], b.out[n]), n
