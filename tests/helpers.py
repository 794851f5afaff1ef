"""Shared test helpers: package/oracle loaders, synthetic states, tolerances."""
import importlib
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "tiegcm3.0_ng_amd")
PKG_NAME = "tiegcm_ng_amd"


def load_pkg():
    """Imports the package directory tiegcm3.0_ng_amd/ as module tiegcm_ng_amd."""
    if PKG_NAME in sys.modules:
        return sys.modules[PKG_NAME]
    spec = importlib.util.spec_from_file_location(
        PKG_NAME, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[PKG_NAME] = mod
    spec.loader.exec_module(mod)
    for sub in ("synth", "tgx", "step", "nest"):
        setattr(mod, sub, importlib.import_module(PKG_NAME + "." + sub))
    return mod


def load_oracle():
    d = os.path.join(ROOT, "oracle")
    if d not in sys.path:
        sys.path.insert(0, d)
    import oracle
    return oracle


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def relerr(got, ref):
    """max |got-ref| / max |ref| (field-normalised relative error)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    return float(np.abs(got - ref).max() / (den if den > 0 else 1.0))


# Pointwise error floors: |got - ref| / max(|ref|, floor).  The floors are the
# reference's own clamps where a field has one: minor species >= 1e-12
# (minor_ng.F90:236-237), O2/O/He >= 1e-9 before normalisation (comp_ng.F90:
# 434-456), TN >= 100 K (dt_ng.F90:208-213).  O+ is clamped at 0
# (oplus_ng.F90:337-340): its floor is 1e-6 cm^-3, far below any value a
# thermosphere holds.  Fields that change sign (winds, the O+ flux terms Fe and
# Fn) are measured against the largest magnitude of their own column: the
# error a column solve leaves is a fraction of the column's solution vector,
# so where a wind passes through zero its error relative to the local value
# is meaningless.  xiop2p/xiop2d are positive pointwise chemistry with no
# floor needed.
POINT_FLOORS = {
    "no": 1e-12, "n4s": 1e-12, "n2d": 1e-12, "ar": 1e-12,
    "o2": 1e-9, "o1": 1e-9, "he": 1e-9,
    "tn": 100., "op": 1e-6, "xiop2p": 1e-30, "xiop2d": 1e-30,
}
SIGN_CHANGING = ("un", "vn", "Fe", "Fn")


def point_floor(name, ref):
    base = name[:-3] if name.endswith("_nm") else name
    if base in SIGN_CHANGING:  # per column (level is the last axis)
        return np.abs(np.asarray(ref, dtype=np.float64)).max(axis=-1, keepdims=True)
    return POINT_FLOORS[base]


def pointwise_err(got, ref, floor):
    """max over points of |got-ref| / max(|ref|, floor)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), floor)).max())


def field_errs(name, got, ref):
    """(field-normalised error, pointwise relative error with the field's floor)."""
    return relerr(got, ref), pointwise_err(got, ref, point_floor(name, ref))


def diag_dominant_system(rng, ncol, nk, margin=(0.1, 1.0)):
    """a,c = -U(0.5,2), b = -(a+c) + U(margin) (strictly diagonally dominant)."""
    a = -rng.uniform(0.5, 2.0, size=(ncol, nk))
    c = -rng.uniform(0.5, 2.0, size=(ncol, nk))
    b = -(a + c) + rng.uniform(*margin, size=(ncol, nk))
    return a, b, c
