"""The Fortran drop-ins (trsolv, blktri with the reference's names and
arguments) called from a Fortran program exactly as the model calls them,
against the oracle.  Needs the GPU and the shim driver built by build()."""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, gpu_available, relerr

EXE = os.path.join(ROOT, "tiegcm3.0_ng_amd", "fortran", "shim_driver")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU"),
              pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran shim driver not built")]


def run(tmp_path, mode, dims, arrays):
    with open(tmp_path / "dims.txt", "w") as fh:
        fh.write(" ".join(str(d) for d in (mode,) + dims) + "\n")
    for n, a in arrays.items():
        np.ascontiguousarray(a, dtype=np.float64).tofile(str(tmp_path / (n + ".bin")))
    subprocess.run([EXE, str(tmp_path)], check=True, timeout=120)
    return np.fromfile(str(tmp_path / "xout.bin"), dtype=np.float64)


def test_fortran_trsolv_matches_oracle(tmp_path, oracle):
    from test_gpu_solvers import config1_system
    rng = np.random.default_rng(21)
    nlev, nlonp4 = 29, 76
    a, b, c, f = config1_system(rng, nlonp4, nlev)
    x0 = np.full((nlonp4, nlev), 3.5)
    got = run(tmp_path, 1, (1, nlev, 1, nlonp4, nlonp4), dict(a=a, b=b, c=c, f=f, x=x0))
    y = x0.copy()
    oracle.trsolv(a, b, c, f, y, 1, nlev, 1, nlev - 1, 1, nlonp4, nlonp4)
    got = got.reshape(nlonp4, nlev)
    assert np.array_equal(got[:2], y[:2]) and np.array_equal(got[:, nlev - 1], y[:, nlev - 1])
    assert relerr(got, y) <= 1e-14


def test_fortran_blktri_matches_oracle(tmp_path, oracle):
    from test_oracle_solvers import random_blocks, to_colmajor
    rng = np.random.default_rng(22)
    nlev, nlonp4 = 57, 76
    P, Q, R, F = random_blocks(rng, nlonp4, nlev, 2)
    arrs = dict(a=to_colmajor(P), b=to_colmajor(Q), c=to_colmajor(R), f=F, x=np.zeros((nlonp4, nlev, 2)))
    got = run(tmp_path, 2, (1, nlev, 1, nlonp4, nlonp4), arrs).reshape(nlonp4, nlev, 2)
    y = np.zeros((nlonp4, nlev, 2))
    oracle.blktri(arrs["a"], arrs["b"], arrs["c"], np.ascontiguousarray(F), y, 1, nlev, 1, nlonp4, nlonp4)
    assert relerr(got, y) <= 1e-14
