"""The precision variant on the systems the fused step assembles: FP32
elimination with FP64 block pivots plus FP64 residual refinement
(tgx_blk_mixed_dev) against the oracle's FP64 solves of the same rows, per
field.  Two refinement sweeps must reach the 1e-10 bound on these
(well-conditioned) systems, and each sweep must not lose accuracy."""
import numpy as np
import pytest
import torch

from helpers import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def oracle_solution(oracle, field, A, B, Cc, F, nk, prec):
    n, _, drop = prec.FIELDS[field]
    a, f = prec.to_oracle_layout(A, F)
    b, _ = prec.to_oracle_layout(B, F)
    c, _ = prec.to_oracle_layout(Cc, F)
    ncol = a.shape[0]
    x = np.zeros_like(f)
    if n == 1:
        oracle.trsolv_ng(a, b, c, f, x, nk, ncol, nk - drop)
    elif n == 2:
        oracle.blktri_ng(a, b, c, f, x, nk, ncol)
    else:
        oracle.blktri3(a, b, c, f, x, nk, ncol, nk - drop)
    return np.ascontiguousarray(np.moveaxis(x.reshape(ncol, nk, n), 2, 0))  # (N, ncol, nk)


@pytest.fixture(scope="module")
def systems(pkg):
    import importlib
    prec = importlib.import_module(pkg.__name__ + ".precision")
    g = pkg.synth.make_grid("tall", nstep_sub=1)
    cap, valid, nk = prec.capture_systems(g, pkg.synth.make_state(g))
    return prec, cap, valid, nk


@pytest.mark.parametrize("field", ["TN", "O+", "NO", "N4S", "N2D", "Ar", "UN,VN", "O2,O,He"])
def test_refinement_reaches_fp64_bound(systems, oracle, field):
    prec, cap, valid, nk = systems
    A, B, Cc, F = prec.split(cap, field)
    xref = oracle_solution(oracle, field, A, B, Cc, F, nk, prec)
    vdev = torch.tensor(valid, device=cap.device)
    errs = []
    for nsweep in (0, 1, 2):
        x = prec.solve_mixed(field, A, B, Cc, F, vdev, nk, nsweep).cpu().numpy()
        e = prec.errors(x, xref, valid, np.zeros_like(valid, dtype=bool))["all"]
        print(field, nsweep, "max %.3e mean %.3e" % e)
        errs.append(e[0])
    assert errs[0] < 1e-3, errs            # FP32 elimination alone: single-precision accuracy
    assert errs[2] <= 1e-10, errs          # two FP64 refinement sweeps: the FP64 bound
    assert errs[1] <= errs[0] and errs[2] <= errs[1] * 1.0001 + 1e-15, errs


def test_precision_report_1p25deg(pkg, oracle):
    """The 1.25 deg grid (288x144 columns x 113 levels) with 5% stiff columns
    (diagonal-dominance margin U(1e-5, 1e-3)): max / mean error of FP32
    elimination alone, + 1 and + 2 FP64 refinement sweeps, per field, against
    the oracle's FP64 solves.  The table is printed and written as JSON to
    gpurun_out/precision_g1p25.json when that directory exists.  Bound checked:
    the well-conditioned (non-stiff) columns reach 1e-10 with two sweeps."""
    import importlib
    import json
    import os
    prec = importlib.import_module(pkg.__name__ + ".precision")
    g = pkg.synth.make_grid("g1p25", nstep_sub=1)
    cap, valid, nk = prec.capture_systems(g, pkg.synth.make_state(g, device="cuda"))
    vdev = torch.tensor(valid, device=cap.device)
    report = {"grid": "g1p25 288x144 columns x 113 levels", "stiff_fraction": 0.05,
              "stiff_margin": [1e-5, 1e-3], "fields": {}}
    for fi, field in enumerate(prec.FIELDS):
        A, B, Cc, F = prec.split(cap, field)
        stiff = prec.make_stiff(A, B, Cc, valid, 0.05, seed=0x7E6C0003 + fi)
        xref = oracle_solution(oracle, field, A, B, Cc, F, nk, prec)
        comps = prec.COMPONENTS.get(field, (field,))
        for nsweep in (0, 1, 2):
            x = prec.solve_mixed(field, A, B, Cc, F, vdev, nk, nsweep).cpu().numpy()
            for ci, cname in enumerate(comps):
                e = prec.errors(x[ci:ci + 1], xref[ci:ci + 1], valid, stiff)
                ens = prec.errors(x[ci:ci + 1], xref[ci:ci + 1], valid & ~stiff, np.zeros_like(stiff))["all"]
                rec = report["fields"].setdefault(cname, {})
                rec["sweeps_%d" % nsweep] = {"max": e["all"][0], "mean": e["all"][1],
                                             "stiff_max": e["stiff"][0], "stiff_mean": e["stiff"][1],
                                             "nonstiff_max": ens[0]}
                print("%-4s sweeps=%d max %.2e mean %.2e | stiff max %.2e | non-stiff max %.2e" % (
                    cname, nsweep, e["all"][0], e["all"][1], e["stiff"][0], ens[0]))
                if nsweep == 2:
                    assert ens[0] <= 1e-10, (cname, ens)
    out = os.path.join(os.environ.get("GRAFT_REPO_ROOT", "."), "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "precision_g1p25.json"), "w") as fh:
            json.dump(report, fh, indent=1)


COMP_FIELDS = ("o2", "o1", "he", "o2_nm", "o1_nm", "he_nm")


@pytest.mark.parametrize("name", ["small", "tall", "g5"])
def test_comp_precision_form_refines_to_the_fp64_step(pkg, oracle, name):
    """The fused precision form of comp (tgx_ng_comp_mixed: FP32 assembly and
    elimination, FP64 block pivots, FP64-residual refinement sweeps) against
    the FP64 comp kernel on the same state, pointwise with comp's floor
    (1e-9, the reference's clamp): FP32 alone is within 1e-4, each sweep
    gains accuracy and two sweeps reach the 1e-10 bound.  The FP64 kernel
    itself is held to the oracle at 1e-12 by the step tests; one case here
    checks it again."""
    from helpers import field_errs
    g = pkg.synth.make_grid(name, nstep_sub=1)
    F = pkg.synth.make_state(g)
    ns = pkg.step.NestState(g, F)
    ns.comp()
    torch.cuda.synchronize()
    ref = {n: ns.fortran(n).copy() for n in COMP_FIELDS}
    if name == "small":
        ora = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
        for n in COMP_FIELDS:
            assert max(field_errs(n, ref[n], ora[n])) <= 1e-12, n
    errs = []
    for nsweep in (0, 1, 2):
        ns.comp_mixed(nsweep)
        torch.cuda.synchronize()
        e = max(max(field_errs(n, ns.fortran(n), ref[n])) for n in COMP_FIELDS)
        print("%s nsweep %d max pointwise error %.3e" % (name, nsweep, e))
        errs.append(e)
    assert errs[0] <= 1e-4, errs
    assert errs[1] <= errs[0] and errs[2] <= errs[1], errs
    assert errs[2] <= 1e-10, errs
    ns.comp()  # the FP64 kernel is unaffected by the precision form's use of the workspace
    torch.cuda.synchronize()
    for n in COMP_FIELDS:
        assert np.array_equal(ns.fortran(n), ref[n]), n


@pytest.mark.parametrize("scheme", ["thomas", "pcr"])
def test_step_with_comp_precision_form(pkg, scheme):
    """tgx_ng_step_mixed: the forked step with comp in its precision form (two
    sweeps) gives the FP64 step's outputs bit for bit for every routine but
    comp, and O2, O, He within 1e-10 pointwise of it; the entry point
    rejects a sweep count outside 0..2."""
    from helpers import field_errs
    g = pkg.synth.make_grid("small", nstep_sub=2)
    F = pkg.synth.make_state(g)
    a = pkg.step.NestState(g, F, scheme=scheme)
    a.step()
    b = pkg.step.NestState(g, F, scheme=scheme)
    b.comp_nsweep = 2
    b.step()
    torch.cuda.synchronize()
    for n in a.out:
        if n in COMP_FIELDS:
            e = max(field_errs(n, b.fortran(n), a.fortran(n)))
            print("%s %s max pointwise error %.3e" % (scheme, n, e))
            assert e <= 1e-10, (n, e)
        else:
            assert torch.equal(a.out[n], b.out[n]), n
    b.comp_nsweep = 3
    with pytest.raises(pkg.tgx.TgxError):
        b.step()
