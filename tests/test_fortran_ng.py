"""The Fortran drop-ins of the nest routines (dt_ng, duv_ng, oplus_ng, comp_ng,
comp_{n2d,n4s,no,ar}_ng and trsolv_ng, with the reference's names and
arguments) driven by a Fortran program exactly as dynamics_ng.F90:99-233
calls them, on stand-ins of the model modules filled with the synthetic state.

GPU: every result field equals the device path's routines run one at a time
(bit for bit: the drop-ins only move data between the model's layout and the
level planes) and matches the oracle pointwise to 1e-12.  CPU: the error
paths (no device; a nest split over tasks) end in the model's shutdown with a
message, before any result is written."""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, field_errs, gpu_available

EXE = os.path.join(ROOT, "tests", "fortran", "ng_driver")
needs_driver = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran nest driver not built")

OUT3 = ("tn", "tn_nm", "un", "un_nm", "vn", "vn_nm", "o2", "o2_nm", "o1", "o1_nm", "he", "he_nm", "op", "op_nm",
        "n2d", "n2d_nm", "n4s", "n4s_nm", "no", "no_nm", "ar", "ar_nm", "xiop2p", "xiop2d", "Fe", "Fn")
OUT2 = ("tlbc", "tlbc_nm", "ulbc", "vlbc", "ulbc_nm", "vlbc_nm", "flx_he")


def write_case(d, g, Fo, mode):
    """run.txt, the level and latitude arrays and every field in the model's
    layout (synth.to_fortran_layout) as raw float64 files."""
    b = [x for row in g["b"] for x in row]
    rv = [g["dz"], g["dlamda"], g["dphi"], g["dtx2inv"], g["shapiro"], g["expzmid"], g["expzmid_inv"],
          g["colfac"], g["opdiffcap"], g["ar_glbm"][0], g["ar_glbm"][1]] + b + list(g["fb"]) + [0.]
    with open(os.path.join(d, "run.txt"), "w") as fh:
        fh.write("%d %d %d %d %d %d\n" % (mode, g["nk"], g["nlon"], g["nlat"], g["nstep_sub"], g["calc_helium"]))
        fh.write(" ".join(repr(float(x)) for x in rv) + "\n")
    for n in ("expz", "difk", "dift", "xmue", "zpint", "zpmid", "glat", "racs", "cs", "tanphi", "cor"):
        np.asarray(g[n], dtype=np.float64).tofile(os.path.join(d, n + ".bin"))
    for n, a in Fo.items():
        np.ascontiguousarray(a, dtype=np.float64).tofile(os.path.join(d, n + ".bin"))


def run_driver(d, timeout=120):
    return subprocess.run([EXE, str(d)], capture_output=True, text=True, timeout=timeout)


def read_out(d, g, name):
    a = np.fromfile(os.path.join(d, "out_" + name + ".bin"), dtype=np.float64)
    nlatd, nlond = g["nlat"] + 4, g["nlon"] + 4
    return a.reshape((nlatd, nlond, g["nk"]) if name in OUT3 else (nlatd, nlond))


@needs_driver
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
@pytest.mark.parametrize("name,nsub", [("small", 3), ("tall", 1), ("g5", 2)])
def test_fortran_nest_step_matches_device_path_and_oracle(tmp_path, pkg, oracle, name, nsub):
    import torch
    g = pkg.synth.make_grid(name, nstep_sub=nsub)
    F = pkg.synth.make_state(g)
    Fo = pkg.synth.to_fortran_layout(F)
    write_case(tmp_path, g, Fo, 1)
    print("# fortran nest step", name, "nsub", nsub, flush=True)
    r = run_driver(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = pkg.step.run_oracle_step(oracle, g, Fo)
    ns = pkg.step.NestState(g, F)
    for rt in ("oplus", "minors", "dt", "duv", "comp"):
        getattr(ns, rt)()
    torch.cuda.synchronize()
    for n in OUT3:
        got = read_out(tmp_path, g, n)
        assert np.array_equal(got, ns.fortran(n)), (name, n, "differs from the device path")
        fe, pe = field_errs(n, got, ref[n])
        assert fe <= 1e-12 and pe <= 1e-12, (name, n, fe, pe)
    for n in OUT2:
        got = read_out(tmp_path, g, n)
        assert np.array_equal(got, ref[n]), (name, n)


@needs_driver
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
def test_fortran_trsolv_ng_shim_matches_oracle(tmp_path, pkg, oracle):
    """trsolv_ng (trsolv_ng.F90:1-28) through its drop-in, dimensions from the
    stand-in modules: rows 1..nk-1 of every column, ghosts included."""
    g = pkg.synth.make_grid("small")
    nk, nlond, nlatd = g["nk"], g["nlon"] + 4, g["nlat"] + 4
    rng = np.random.default_rng(31)
    a = -rng.uniform(.5, 2., (nlatd, nlond, nk))
    c = -rng.uniform(.5, 2., (nlatd, nlond, nk))
    b = -(a + c) + rng.uniform(.1, 1., (nlatd, nlond, nk))
    f = rng.normal(size=(nlatd, nlond, nk))
    x0 = np.full((nlatd, nlond, nk), 7.25)
    write_case(tmp_path, g, dict(a=a, b=b, c=c, f=f, x=x0), 2)
    r = run_driver(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(os.path.join(tmp_path, "out_xout.bin"), dtype=np.float64).reshape(nlatd, nlond, nk)
    want = x0.copy()
    oracle.trsolv_ng(a, b, c, f, want, nk, nlond * nlatd, nk - 1)
    assert np.array_equal(got[..., nk - 1], x0[..., nk - 1])     # row nk untouched
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= 1e-14, err


@needs_driver
@pytest.mark.skipif(gpu_available(), reason="the no-device error path needs a host without a GPU")
def test_fortran_drop_in_without_a_device_shuts_down_with_a_message(tmp_path, pkg):
    g = pkg.synth.make_grid("tiny")
    write_case(tmp_path, g, pkg.synth.to_fortran_layout(pkg.synth.make_state(g)), 1)
    r = run_driver(tmp_path)
    assert r.returncode != 0
    assert "tgx ctx_create:" in r.stdout and "shutdown: tgx: ctx_create failed" in r.stdout, r.stdout
    assert not os.path.exists(os.path.join(tmp_path, "out_op.bin"))


@needs_driver
def test_fortran_drop_in_refuses_a_nest_split_over_tasks(tmp_path, pkg):
    """The fused step applies lateral data on all four sides: a nest whose
    west side is an interior task boundary is refused before any device work."""
    g = pkg.synth.make_grid("tiny")
    write_case(tmp_path, g, pkg.synth.to_fortran_layout(pkg.synth.make_state(g)), 3)
    r = run_driver(tmp_path)
    assert r.returncode != 0
    assert "needs the nest on one task" in r.stdout, r.stdout
    assert not os.path.exists(os.path.join(tmp_path, "out_op.bin"))


@needs_driver
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
@pytest.mark.parametrize("name,order,oplus", [("TN", 3, 0), ("OPLUS", 3, 1)])
def test_fortran_ringfilter_shim_matches_oracle(tmp_path, oracle, name, order, oplus):
    """ringfilter (ringfilter.F:1-111) through its Fortran drop-in, called on
    every latitude of the 5 deg grid as filter_tn does (dt.F:486), the module
    state (nlon, nlat, dlamda, the chunk table of cons.F:331-355) from the
    stand-ins: filtered rows within 1e-12 of the oracle's, the others
    untouched."""
    import math
    nlon, nlat, nk = 72, 36, 29
    chunk = oracle.RING_CHUNKS[(nlon, nlat)]
    dlam = 2 * math.pi / nlon
    rng = np.random.default_rng(41 + oplus)
    rows = rng.normal(size=(nlat, nk, nlon + 4)) * 3. + 250.
    rows[..., :2] = rows[..., nlon:nlon + 2]
    rows[..., nlon + 2:] = rows[..., 2:4]
    with open(tmp_path / "run.txt", "w") as fh:
        fh.write("4 %d %d %d %d %d\n" % (nk, nlon, nlat, order, oplus))
        fh.write(" ".join(repr(float(x)) for x in [0., dlam] + [0.] * 22) + "\n")
    with open(tmp_path / "chunk.txt", "w") as fh:
        fh.write("%d\n%s\n" % (len(chunk), " ".join(str(c) for c in chunk)))
    rows.tofile(str(tmp_path / "rows.bin"))
    r = run_driver(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(str(tmp_path / "out_rows.bin"), dtype=np.float64).reshape(nlat, nk, nlon + 4)
    for lat in range(1, nlat + 1):
        nc = oracle.ringfilter_chunks(lat, nlat, chunk)
        want = np.ascontiguousarray(rows[lat - 1])
        if nc == 0:
            assert np.array_equal(got[lat - 1], want), lat
            continue
        oracle.ringfilter_row(want, nlon, order, oplus, nc, dlam)
        err = float((np.abs(got[lat - 1] - want) / np.abs(want).max(axis=1, keepdims=True)).max())
        assert err <= 1e-12, (lat, err)


@needs_driver
def test_fortran_ringfilter_shim_refuses_a_too_high_order(tmp_path):
    """An order the filter cannot take (>= nlon, ringfilter.F:36) ends in the
    model's shutdown, as the reference's does; checked on a host without a
    GPU too, since the order is refused before any device work."""
    nlon, nlat, nk = 8, 6, 4
    with open(tmp_path / "run.txt", "w") as fh:
        fh.write("4 %d %d %d %d 0\n" % (nk, nlon, nlat, nlon))
        fh.write(" ".join(repr(float(x)) for x in [0., 0.785] + [0.] * 22) + "\n")
    with open(tmp_path / "chunk.txt", "w") as fh:
        fh.write("1\n2\n")
    np.full((nlat, nk, nlon + 4), 1.0).tofile(str(tmp_path / "rows.bin"))
    r = run_driver(tmp_path)
    assert r.returncode != 0
    assert "shutdown: ringfilter" in r.stdout, r.stdout
