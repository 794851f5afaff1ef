"""Builds libtgx.so (HIP, gfx950) in-tree with hipcc.

Used by __graft_entry__.build() and runnable directly:
    python tiegcm3.0_ng_amd/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtgx.so")
SOURCES = ["solvers.hip", "step.hip", "api.cpp", "step_api.cpp", "noalloc.hip", "filter.hip", "pack.hip",
           "host_ctx.cpp", "nest.hip", "global.hip"]

# FP64 parity kernels follow the reference's arithmetic order: no contraction.
# Packed FP32 instructions are disabled for the whole library.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"),
]


def hipcc():
    for p in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def build(verbose=True, jobs=4):
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for s in srcs:
        obj = os.path.join(objdir, s + ".o")
        objs.append(obj)
        cmd = [hipcc()] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", obj]
        if s.endswith(".cpp"):
            cmd = [hipcc(), "-x", "hip"] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for p in procs if p[1].poll() is None]) >= jobs:
            procs[0][1].wait()
    for s, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        if p.returncode != 0:
            raise RuntimeError("hipcc failed on %s:\n%s" % (s, txt))
        if verbose and txt.strip():
            print(txt)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


def _fc(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_fortran(verbose=True):
    """The Fortran drop-ins and their test drivers, built with amdflang
    against libtgx.so.  Skipped if no Fortran compiler.

    - fortran/tgx_c.F90, tgx_ng_c.F90: iso_c_binding interfaces (product).
    - fortran/shim_driver: trsolv / blktri drop-ins (trsolv_shim.F90) called
      as the global model calls them.
    - tests/fortran/ng_driver: trsolv_ng_shim.F90 and the nest routine
      drop-ins (ng_shims.F90) and ringfilter_shim.F90 compiled against the
      test stand-ins of the model modules (tests/fortran/model_modules.F90),
      one nest step run through them in the model's order."""
    fc = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"
    if not os.path.exists(fc):
        if verbose:
            print("amdflang not found: Fortran shims not built")
        return None
    fdir = os.path.join(HERE, "fortran")
    tdir = os.path.join(ROOT, "tests", "fortran")
    bdir = os.path.join(HERE, "_build", "fortran")
    tbdir = os.path.join(bdir, "test")
    os.makedirs(tbdir, exist_ok=True)
    base = ["-fdefault-real-8", "-O2", "-cpp"]
    link = ["-L" + HERE, "-ltgx"]

    def compile_(src, moddir, extra=()):
        obj = os.path.join(moddir, os.path.basename(src) + ".o")
        _fc([fc] + base + list(extra) + ["-J", moddir, "-I", bdir, "-c", src, "-o", obj], verbose)
        return obj

    common = [compile_(os.path.join(fdir, s), bdir) for s in ("tgx_c.F90", "tgx_ng_c.F90")]
    objs = common + [compile_(os.path.join(fdir, s), bdir, ["-DTGX_NLONP4=76"])
                     for s in ("trsolv_shim.F90", "shim_driver.F90")]
    exe = os.path.join(fdir, "shim_driver")
    _fc([fc, "-o", exe] + objs + link + ["-Wl,-rpath,$ORIGIN/.."], verbose)
    tobjs = [compile_(os.path.join(tdir, "model_modules.F90"), tbdir)]
    tobjs += [compile_(os.path.join(fdir, s), tbdir)
              for s in ("trsolv_ng_shim.F90", "ng_shims.F90", "ringfilter_shim.F90")]
    tobjs += [compile_(os.path.join(tdir, "ng_driver.F90"), tbdir)]
    texe = os.path.join(tdir, "ng_driver")
    _fc([fc, "-o", texe] + tobjs + common + link + ["-Wl,-rpath,$ORIGIN/../../tiegcm3.0_ng_amd"], verbose)
    return exe


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    build_fortran(verbose="-q" not in sys.argv)
