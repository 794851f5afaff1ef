"""Oracle vs the reference's own compiled matutil.F90 (oracle/_ref, built from
the reference source where it lies by oracle/Makefile 'ref').  Skipped when
oracle/_ref has not been built."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "libref_matutil.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


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
