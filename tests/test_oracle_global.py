"""Pins of the oracle's global-grid forms (dt.F, duv.F, minor.F with the
species wrappers, oplus.F, comp_o2o.F, comp.F) by checks that do not restate
the oracle (tests/pinglb.py): periodicity, the pinned nest forms away from
their edges, the model description's coefficients where the global form
differs, the top levels and the ring filter's place.  Each check must hold on
the oracle, and each mistake planted in a global code path of the oracle
(MUT(801..809) in oracle/column.c) must fail at least one check."""
import pytest

import pinglb

ROUTINES = sorted(pinglb.CHECKS)


@pytest.mark.parametrize("routine", ROUTINES)
def test_global_form_meets_its_checks(routine):
    errs = pinglb.run_checks(routine)
    for n, e in sorted(errs.items()):
        print("%-28s %.3e (bound %.0e)" % (n, e, pinglb.TOL[n]))
    assert not pinglb.failures(errs), pinglb.failures(errs)


CASES = [(r, m) for r in sorted(pinglb.MUTATIONS) for m in sorted(pinglb.MUTATIONS[r])]


@pytest.mark.parametrize("routine,mid", CASES, ids=["%s-%d" % c for c in CASES])
def test_planted_global_mistake_fails_a_check(routine, mid):
    o = pinglb.oracle()
    o.set_mutation(mid)
    try:
        failed = pinglb.failures(pinglb.run_checks(routine))
    finally:
        o.set_mutation(0)
    print(mid, pinglb.MUTATIONS[routine][mid], "->", sorted(failed))
    assert failed, "planted mistake %d (%s) passes every check" % (mid, pinglb.MUTATIONS[routine][mid])


def test_global_halo_fill_is_the_pole_and_periodic_rule():
    """synth.fill_global_halos against mk_polelat.F:30-57 written out for a
    small grid: latitude 0 at Fortran longitude L takes latitude 1 at
    L + nlon/2 for L <= nlon/2 + 2 and at L - nlon/2 above, sign-flipped for
    the wind components."""
    import numpy as np
    p = pinglb.pkg()
    g = p.synth.make_grid("small", glb=True)
    F = p.synth.make_global_state(g)
    nlon, nlat = g["nlon"], g["nlat"]
    for name, sign in (("tn", 1.), ("un", -1.)):
        f = F[name].numpy()                  # (nk, nlat+4, nlon+4), i = L - 1
        imaxh, ilim = nlon // 2, nlon // 2 + 2
        for jnew, jold in ((1, 2), (0, 3), (nlat + 2, nlat + 1), (nlat + 3, nlat)):
            want = np.empty_like(f[:, jnew])
            ident = imaxh + 1
            for i in range(1, ilim + 1):     # mk_polelat.F:36-47, 1-based L
                want[:, ident + 2 - 1] = sign * f[:, jold, i + 2 - 1]
                want[:, i - 1] = sign * f[:, jold, ident - 1]
                ident += 1
            assert np.array_equal(f[:, jnew], want), (name, jnew)
        assert np.array_equal(f[:, :, 0], f[:, :, nlon]) and np.array_equal(f[:, :, nlon + 3], f[:, :, 3])
