"""The C-ABI library loads and exports every symbol include/*.h declares (CPU only;
no compute calls)."""
import ctypes
import os
import re

from helpers import ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(inc, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(tgx_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "tiegcm3.0_ng_amd", "libtgx.so"))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 10


def test_version_and_error_slot_without_gpu():
    lib = ctypes.CDLL(os.path.join(ROOT, "tiegcm3.0_ng_amd", "libtgx.so"))
    lib.tgx_version.restype = ctypes.c_int
    assert lib.tgx_version() == 100
    lib.tgx_last_error.restype = ctypes.c_char_p
    lib.tgx_clear_error()
    assert lib.tgx_last_error() == b""
