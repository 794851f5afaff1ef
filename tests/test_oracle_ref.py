"""Oracle vs the reference's own compiled matutil.F90: live against oracle/_ref
(built from the reference source where it lies by oracle/Makefile 'ref';
skipped when not built) and against the committed vectors written from it."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "libref_matutil.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
def test_matinv3_matinv2_bitwise_equal_to_reference(oracle):
    ref = C.CDLL(REF)
    rng = np.random.default_rng(0)
    for t in range(500):
        A = np.asfortranarray(rng.normal(size=(3, 3)) * 10 ** rng.uniform(-3, 3))
        out = np.zeros((3, 3), order="F")
        ref.tgx_ref_matinv3(A.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        assert np.array_equal(out, oracle.matinv3(np.ascontiguousarray(A)))
        B = np.asfortranarray(rng.normal(size=(2, 2)))
        out2 = np.zeros((2, 2), order="F")
        ref.tgx_ref_matinv2(B.ctypes.data_as(C.c_void_p), out2.ctypes.data_as(C.c_void_p))
        mine = np.zeros(4)
        oracle.lib().tgx_o_matinv2(B.reshape(-1, order="F").copy().ctypes.data_as(C.c_void_p),
                                   mine.ctypes.data_as(C.c_void_p))
        assert np.array_equal(out2.reshape(-1, order="F"), mine)


def test_oracle_matches_committed_reference_vectors(oracle):
    """The committed fixture (written from the compiled reference matutil by
    tests/golden/gen_matutil_vectors.py) pins the oracle's matinv3/matinv2
    bit for bit without the reference tree."""
    golden = os.path.join(ROOT, "tests", "golden", "matutil_vectors.npz")
    d = np.load(golden)
    for A, R in zip(d["a3"], d["inv3"]):
        assert np.array_equal(oracle.matinv3(np.ascontiguousarray(A)), R)
    for B, R in zip(d["a2"], d["inv2"]):
        mine = np.zeros(4)
        oracle.lib().tgx_o_matinv2(B.reshape(-1, order="F").copy().ctypes.data_as(C.c_void_p),
                                   mine.ctypes.data_as(C.c_void_p))
        assert np.array_equal(mine.reshape(2, 2, order="F"), R)
