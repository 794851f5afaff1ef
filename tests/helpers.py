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
    for sub in ("synth", "tgx", "step"):
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


def diag_dominant_system(rng, ncol, nk, margin=(0.1, 1.0)):
    """a,c = -U(0.5,2), b = -(a+c) + U(margin) (strictly diagonally dominant)."""
    a = -rng.uniform(0.5, 2.0, size=(ncol, nk))
    c = -rng.uniform(0.5, 2.0, size=(ncol, nk))
    b = -(a + c) + rng.uniform(*margin, size=(ncol, nk))
    return a, b, c
