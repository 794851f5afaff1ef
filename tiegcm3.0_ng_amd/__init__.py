"""tiegcm3.0_ng_amd: MI355X-native implicit vertical-column step of TIE-GCM NG.

The product is the C-ABI library libtgx.so (HIP kernels for gfx950) declared
in include/tgx.h.  This Python package is a thin host-side mirror of the
reference's Fortran interface (trsolv, trsolv_ng, blktri, dt_ng, duv_ng,
minor_ng, oplus_ng, comp_ng) over that C-ABI, plus the synthetic workload
generator used by the benchmark.
"""
