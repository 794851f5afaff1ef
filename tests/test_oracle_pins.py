"""Pins of the oracle's six column-step routines (dt_ng, duv_ng, minor_ng with
the species wrappers, oplus_ng, comp_o2o_ng, comp_ng) by checks that do not
restate the oracle: the equations of the reference's model description,
conservation, mirror symmetry and closed forms of degenerate cases
(tests/pinlib.py).  Each check must hold on the oracle, and each mistake
planted in the oracle (oracle.set_mutation, the MUT sites of
oracle/column.c) must fail at least one check of its routine."""
import pytest

import pinlib

ROUTINES = sorted(pinlib.CHECKS)


@pytest.mark.parametrize("routine", ROUTINES)
def test_oracle_routine_meets_its_checks(routine):
    errs = pinlib.run_checks(routine)
    for n, e in sorted(errs.items()):
        print("%-28s %.3e (bound %.0e)" % (n, e, pinlib.TOL[n]))
    assert not pinlib.failures(errs), pinlib.failures(errs)


CASES = [(r, m) for r in ROUTINES for m in sorted(pinlib.MUTATIONS[r])]


@pytest.mark.parametrize("routine,mid", CASES, ids=["%s-%d" % c for c in CASES])
def test_planted_mistake_fails_a_check(routine, mid):
    with pinlib.mutation(mid):
        failed = pinlib.failures(pinlib.run_checks(routine))
    print(mid, pinlib.MUTATIONS[routine][mid], "->", sorted(failed))
    assert failed, "planted mistake %d (%s) passes every check" % (mid, pinlib.MUTATIONS[routine][mid])


def test_every_routine_has_checks_and_mutations():
    for r in ("dt_ng", "duv_ng", "minor_ng", "oplus_ng", "comp_o2o_ng", "comp_ng"):
        assert pinlib.CHECKS.get(r) and len(pinlib.MUTATIONS.get(r, {})) >= 4, r
