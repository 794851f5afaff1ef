"""The synthetic state with a thin N2 top (synth.make_grid(..., n2_top=...))
drives the comp post pass into its N2 extrapolation (comp_ng.F90:419-431):
every interior column of the oracle's solution crosses N2 < 0.01 above its
bottom level and below its top two levels, so the extrapolated branch runs
there.  The default state never crosses.  The GPU parity test of that branch
(tests/test_gpu_step.py) relies on this."""
import numpy as np


def _n2_interior(pkg, oracle, n2_top):
    g = pkg.synth.make_grid("tall", nstep_sub=1, n2_top=n2_top)
    F = pkg.synth.make_state(g)
    ref = pkg.step.run_oracle_step(oracle, g, pkg.synth.to_fortran_layout(F))
    n2 = 1. - ref["o2"] - ref["o1"] - ref["he"]
    return n2[2:-2, 2:-2, :]  # (lat, lon, level), interior columns


def test_thin_n2_top_crosses_inside_every_column(pkg, oracle):
    n2 = _n2_interior(pkg, oracle, 0.003)
    below = n2 < 0.01
    assert below.any(axis=2).all()
    first = np.argmax(below, axis=2)
    nk = n2.shape[2]
    assert first.min() >= 1 and first.max() + 1 < nk - 1, (first.min(), first.max())
    assert (n2[:, :, -2:] < 0.008).all() and (n2[:, :, 0] > 0.02).all()


def test_default_state_never_crosses(pkg, oracle):
    assert not (_n2_interior(pkg, oracle, None) < 0.01).any()
