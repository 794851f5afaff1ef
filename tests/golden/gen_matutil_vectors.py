"""Writes tests/golden/matutil_vectors.npz: inputs and outputs of the
reference's own matinv3 / matinv2 (matutil.F90:100-118), compiled from the
reference source by `make -C oracle ref` into oracle/_ref (run this in a
checkout where that library is built).  The fixture is data only: seeded
inputs and the compiled reference's outputs, so the oracle and the device
block inversions can be pinned where the reference tree is absent.

    python tests/golden/gen_matutil_vectors.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_matutil.so")


def main(n=256, seed=20261202):
    ref = C.CDLL(REF)
    rng = np.random.default_rng(seed)
    # scales spanning the magnitudes of the comp / minor alpha blocks
    a3 = rng.normal(size=(n, 3, 3)) * 10.0 ** rng.uniform(-3, 3, size=(n, 1, 1))
    # a few diagonally dominant blocks like the comp pivots
    a3[: n // 4] += np.eye(3) * 5.0
    a2 = rng.normal(size=(n, 2, 2)) * 10.0 ** rng.uniform(-3, 3, size=(n, 1, 1))
    inv3 = np.zeros_like(a3)
    inv2 = np.zeros_like(a2)
    for t in range(n):
        A = np.asfortranarray(a3[t])
        out = np.zeros((3, 3), order="F")
        ref.tgx_ref_matinv3(A.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        inv3[t] = out
        B = np.asfortranarray(a2[t])
        out2 = np.zeros((2, 2), order="F")
        ref.tgx_ref_matinv2(B.ctypes.data_as(C.c_void_p), out2.ctypes.data_as(C.c_void_p))
        inv2[t] = out2
    np.savez_compressed(os.path.join(HERE, "matutil_vectors.npz"), a3=a3, inv3=inv3, a2=a2, inv2=inv2,
                        seed=np.int64(seed))
    print("wrote", n, "matinv3 and matinv2 vectors")


if __name__ == "__main__":
    main()
