#!/usr/bin/env python3
"""Benchmark of the fused implicit column step (TIE-GCM NG) on one MI355X.

Workload (default): the 0.625 deg global grid, 576 x 288 columns x 113 levels,
all 11 unknowns per column per step (TN; UN+VN as the 2x2 block; O+; NO,
N4S, N2D, Ar; O2+O+He as the 3x3 block), one O+ sub-cycle, FP64, seeded
synthetic fields resident in HBM.  A step is one call of tgx_glb_step, the
whole column step of the global model over that state (dt.F, duv.F, minor.F,
oplus.F, comp.F with their ring filters, periodic longitudes and the rows
across the poles), the independent routines on forked streams, bitwise equal
to the reference's order (--serial times that order on one stream instead).
--domain nest runs the same grid as one nest domain (tgx_ng_step, lateral
Dirichlet data at its edges); the global run also reports that form's step
time beside it.

Prints one JSON line:
  value   = Mcolumn-level solves/s over the interior columns (whole job),
  roofline: the column kernel with the longest launch: its algorithmic bytes
            (the planes it must read and write once) / its hipEvent-timed
            duration (in a few serial steps after the timed region, where
            it runs alone), against the 8 TB/s HBM peak; traffic from the
            committed rocprofv3 --pmc summary of that kernel when present,
  cpu_baseline: the C oracle (a restatement of the reference) on the host
            cores, the workload split into latitude bands, one process each.
"""
import argparse
import importlib
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "tiegcm3.0_ng_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "tiegcm_ng_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    for sub in ("synth", "tgx", "step", "nest"):
        setattr(mod, sub, importlib.import_module(name + "." + sub))
    return mod


ROUTINES = ("oplus", "minors", "dt", "duv", "comp")  # the library's routine order


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cc_version():
    import subprocess
    try:
        return subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except Exception:
        return "gcc"


def _cpu_worker(conn, grid, Fo, reps):
    """Child process: times `reps` oracle steps of one latitude band."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        pkg = load_pkg()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pkg.step.run_oracle_step(oracle, grid, Fo)
            times.append(time.perf_counter() - t0)
        conn.send(times)
    except Exception as e:  # reported by the parent
        conn.send(repr(e))
    conn.close()


def cpu_baseline(pkg, config, nstep_sub, budget_s=30.0):
    """The oracle's column step (a C restatement of the reference routines,
    gcc -O3, reference arithmetic order) on the host cores: the workload's
    latitude rows split into P bands, one process per band, as the reference
    splits a grid over MPI tasks (the column step needs no communication
    inside a step).  P = the host cores this process may use, at most 16 (the
    CPU share that goes with one GPU).  Every band has the workload's
    longitudes, levels, timestep and physics; the value counts the interior
    columns of all P bands over the slowest band's median step time."""
    import multiprocessing as mp
    nlon, nlat, nk, dlev, step, dstep, region = pkg.synth.CONFIGS[config]
    P = max(1, min(16, len(os.sched_getaffinity(0))))
    rows = -(-nlat // P)
    sub = "cpu_band"
    pkg.synth.CONFIGS[sub] = (nlon, rows, nk, dlev, step, dstep, region)
    g = pkg.synth.make_grid(sub, nstep_sub=nstep_sub)
    Fo = pkg.synth.to_fortran_layout(pkg.synth.make_state(g))
    # one step of one band first, in this process, to size the repetitions
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.perf_counter()
    pkg.step.run_oracle_step(oracle, g, Fo)
    t_one = time.perf_counter() - t0
    reps = max(1, min(5, int(budget_s / max(t_one, 1e-3))))
    ctx = mp.get_context("fork")
    procs = []
    for _ in range(P):
        a, b = ctx.Pipe(duplex=False)
        pr = ctx.Process(target=_cpu_worker, args=(b, g, Fo, reps))
        pr.start()
        procs.append((pr, a))
    results = []
    for pr, a in procs:
        results.append(a.recv())
        pr.join()
    bad = [r for r in results if not isinstance(r, list)]
    if bad:
        raise RuntimeError("cpu baseline worker failed: %s" % bad[0])
    per_band = sorted(sorted(r)[len(r) // 2] for r in results)
    t_step = per_band[-1]
    cols = nlon * rows * P
    value = cols * nk * (10 + nstep_sub) / t_step / 1e6
    return {"value": round(value, 3), "unit": "Mcolumn-level solves/s", "cores": P, "kind": "port",
            "cpu": _cpu_model(), "compiler": _cc_version() + " -O3 -ffp-contract=off",
            "sample": "the whole workload: %d latitude bands of %dx%d columns x %d levels, one process per "
                      "band, oracle column step (C restatement of the reference routines), median of %d "
                      "steps per band, slowest band %.3f s/step" % (P, nlon, rows, nk, reps, t_step)}


def traffic_from_profiles(kernel, workload, scheme):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc summary
    (profiles/*.txt lines 'TRAFFIC <kernel>|<workload>|<scheme> <bytes>'), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    key = "%s|%s|%s" % (kernel.replace(" ", ""), workload, scheme)
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".txt"):
            continue
        with open(os.path.join(d, fn)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == key:
                    return int(float(parts[2]))
    return None


def bench_nests(pkg, args):
    """Nested-grid batching: the 2.5 deg global grid (144x72), nest 1 (120x60
    at 0.5 deg, NSTEP_NG=2) and nest 2 (100x50 at 0.1 deg, NSTEP_NG=5 per nest-1
    step), all 57 levels, coupled as advance_ng couples them (map_in,
    set_bndry, the time interpolation of the boundary data and map_out, all on
    the device), one outer step per HIP graph replay; the same domains without
    the coupling are timed beside it."""
    import torch
    names = ("g2p5", "nest1", "nest2")
    # the nests cap the O+ diffusion (opdiffcap, oplus_ng.F90:145-148): each
    # sub-step advances the state, and at 0.1 deg the explicit field-aligned
    # O+ diffusion is unstable without the cap; the nests' states are
    # interpolated from the global one, as a model initialises its nests
    grids = [pkg.synth.make_grid(nm, nstep_sub=args.nstep_sub, opdiffcap=1e8 if nm.startswith("nest") else 0.)
             for nm in names]
    doms = list(zip(grids, pkg.synth.nested_states(grids, device="cuda")))
    def timed(runner):
        runner.capture()
        for _ in range(args.warmup):
            runner.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runner.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    # the domains alone, lateral data given
    plain = pkg.step.NestedStep(doms, [1, 2, 5], scheme=args.scheme, forked=not args.serial)
    ms_plain = timed(plain)
    sub = plain.substeps()
    del plain
    torch.cuda.empty_cache()
    # the recursion with its map_in / set_bndry / time interpolation / map_out
    ns = pkg.nest.CoupledNests(doms, [1, 2, 5], scheme=args.scheme, forked=not args.serial)
    ms = timed(ns)
    # The synthetic nested state is no equilibrium: with its lateral data
    # held fixed, a strip of nest 1's O+ reaches zero in the third outer step
    # and the state turns non-finite from there (the oracle's recursion does
    # the same, at the same 154 points).  The kernels' work does not depend on
    # the values, so the timing stands; the differences below are taken after
    # one outer step, where every value is finite.
    finite = all(bool(torch.isfinite(st.current(n)).all()) for st in ns.states for n in pkg.step.PROGNOSTIC)
    one = pkg.nest.CoupledNests(doms, [1, 2, 5], scheme=args.scheme, forked=not args.serial)
    one.run()
    torch.cuda.synchronize()
    finite1 = all(bool(torch.isfinite(st.current(n)).all()) for st in one.states for n in pkg.step.PROGNOSTIC)
    # the same recursion with comp in its FP32 + FP64-refinement form
    # (tgx_ng_step_mixed): time over the same number of outer steps, and the
    # largest pointwise relative difference of every domain's prognostic
    # fields from the FP64 recursion's after one outer step from the same
    # state (floors of the reference's clamps; the winds against their column
    # maximum)
    floors = {"o2": 1e-9, "o1": 1e-9, "he": 1e-9, "no": 1e-12, "n4s": 1e-12, "n2d": 1e-12, "ar": 1e-12,
              "tn": 100., "op": 1e-6}
    precision = {}
    if not args.serial:
        for nsweep in (1, 2):
            nm = pkg.nest.CoupledNests(doms, [1, 2, 5], scheme=args.scheme, forked=True, comp_nsweep=nsweep)
            ms_m = timed(nm)
            del nm
            m1 = pkg.nest.CoupledNests(doms, [1, 2, 5], scheme=args.scheme, forked=True, comp_nsweep=nsweep)
            m1.run()
            torch.cuda.synchronize()
            err = {}
            for n in pkg.step.PROGNOSTIC:
                worst = 0.
                base = n[:-3] if n.endswith("_nm") else n
                for a, b in zip(one.states, m1.states):
                    x, y = a.current(n), b.current(n)
                    if base in floors:
                        den = torch.clamp(x.abs(), min=floors[base])
                    elif x.dim() == 3:
                        den = x.abs().amax(dim=0, keepdim=True).clamp(min=1e-30)
                    else:
                        den = x.abs().max().clamp(min=1e-30)
                    d = float(((x - y).abs() / den).max())
                    worst = d if d != d else max(worst, d)  # a NaN is reported, not dropped
                err[n] = float("%.3e" % worst)
            precision["comp_fp32_refine%d" % nsweep] = {"ms_per_step": round(ms_m, 4),
                                                        "vs_fp64": round(ms / ms_m, 4),
                                                        "max_rel_diff_vs_fp64_after_one_outer_step": err}
            del m1
            torch.cuda.empty_cache()
    del one
    unknowns = 10 + args.nstep_sub
    colev = sum(d[0]["nlon"] * d[0]["nlat"] * d[0]["nk"] * sub[i] for i, d in enumerate(doms)) * unknowns
    alg = sum(pkg.step.alg_bytes_per_column_level(r) * (d[0]["nlon"] + 4) * (d[0]["nlat"] + 4) * d[0]["nk"]
              * sub[i] * (args.nstep_sub if r == "oplus" else 1) for i, d in enumerate(doms) for r in ROUTINES)
    out = {
        "metric": "Mcolumn-level solves/s + %HBM roofline, FP32+refine vs FP64 at <=1e-10 rel",
        "value": round(colev / (ms * 1e-3) / 1e6, 2), "unit": "Mcolumn-level solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded profiles; the nests' fields interpolated from the global state)",
        "config": {"workload": "nested: g2p5 global 144x72 + nest1 120x60 (NSTEP_NG=2) + nest2 100x50 "
                               "(NSTEP_NG=5 per nest-1 step), 57 levels, %d unknowns, one outer step per "
                               "HIP graph replay" % unknowns,
                   "substeps": {names[i]: n for i, n in sub.items()}, "column_levels_per_step": colev,
                   "coupling": "map_in, set_bndry, time interpolation and map_out on the device",
                   "precision_form": precision,
                   "ms_without_coupling": round(ms_plain, 4),
                   "state_finite_after_one_outer_step": finite1,
                   "state_finite_after_run": finite,
                   "outer_steps_run": 1 + args.warmup + args.steps,
                   "scheme": args.scheme, "schemes": ns.states[0].schemes(),
                   "wave_columns_per_block": {names[i]: st.wave_columns() for i, st in enumerate(ns.states)},
                   "schedule": "serial" if args.serial else "concurrent"},
        "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole outer step (graph replay)", "alg_bytes_per_launch": alg,
                     "launch_ms": round(ms, 4)},
    }
    print(json.dumps(out), flush=True)


def bench_precision(pkg, args):
    """FP32 elimination + FP64 refinement against the plain FP64 solvers on the
    column systems the fused step assembles at 1.25 deg (288x144 x 113), 5%
    stiff columns.  Per field: solve throughput of each variant and its error
    relative to the FP64 device solution (the oracle comparison of the same
    variants is tests/test_gpu_precision.py)."""
    import importlib
    import numpy as np
    import torch
    prec = importlib.import_module(pkg.__name__ + ".precision")
    cfg = "g1p25" if args.config == "g0p625" else args.config
    g = pkg.synth.make_grid(cfg, nstep_sub=1)
    cap, valid, nk = prec.capture_systems(g, pkg.synth.make_state(g, device="cuda"))
    vdev = torch.tensor(valid, device=cap.device)
    ncol_valid = int(valid.sum())
    fields = {}
    tot = {"fp64": 0.0, 0: 0.0, 1: 0.0, 2: 0.0}
    colev = 0
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(args.steps):
            r = fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps, r

    for fi, field in enumerate(prec.FIELDS):
        n, _, drop = prec.FIELDS[field]
        A, B, Cc, F = prec.split(cap, field)
        prec.make_stiff(A, B, Cc, valid, 0.05, seed=0x7E6C0003 + fi)
        nrow = nk - drop
        ncol = B.shape[1]
        if n == 1:
            a, b, c, f = (t[0].contiguous() for t in (A, B, Cc, F))
            x64 = torch.zeros_like(f)
            ms64, _ = timed(lambda: pkg.tgx.trsolv_dev(a, b, c, f, x64, 0, nrow - 1, scheme=1))
            x64 = x64[None]
        else:
            # (ncol, nk, N, N) column-major blocks for the FP64 block solvers
            bl = [t.reshape(n, n, ncol, nk).permute(2, 3, 1, 0).contiguous() for t in (A, B, Cc)]
            fr = F.permute(1, 2, 0).contiguous()
            xb = torch.zeros_like(fr)
            solver = pkg.tgx.blktri2_dev if n == 2 else pkg.tgx.blktri3_dev
            ms64, _ = timed(lambda: solver(bl[0], bl[1], bl[2], fr, xb, nrow))
            x64 = xb.permute(2, 0, 1).contiguous()
        ref = x64.cpu().numpy()
        rec = {"fp64_ms": round(ms64, 4)}
        tot["fp64"] += ms64
        for nsweep in (0, 1, 2):
            ms, x = timed(lambda: prec.solve_mixed(field, A, B, Cc, F, vdev, nk, nsweep))
            e = prec.errors(x.cpu().numpy(), ref, valid, np.zeros_like(valid, dtype=bool))["all"]
            rec["sweeps_%d" % nsweep] = {"ms": round(ms, 4), "max_err": e[0], "mean_err": e[1]}
            tot[nsweep] += ms
        colev += ncol_valid * nk * n
        fields[field] = rec
    out = {
        "metric": "Mcolumn-level solves/s + %HBM roofline, FP32+refine vs FP64 at <=1e-10 rel",
        "value": round(colev / (tot[2] * 1e-3) / 1e6, 2), "unit": "Mcolumn-level solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(tot[2], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "column systems assembled by the fused step on synthetic state, 5% stiff columns",
        "config": {"workload": "precision: %s column systems of all 11 unknowns, FP32 elimination + 2 FP64 "
                               "refinement sweeps (plain solve, assembly not timed)" % cfg,
                   "fp64_Mcolev_per_s": round(colev / (tot["fp64"] * 1e-3) / 1e6, 2),
                   "fp32_only_Mcolev_per_s": round(colev / (tot[0] * 1e-3) / 1e6, 2),
                   "fp32_1sweep_Mcolev_per_s": round(colev / (tot[1] * 1e-3) / 1e6, 2),
                   "fields": fields},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="g0p625")
    ap.add_argument("--nstep-sub", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of oracle time per process for the CPU baseline")
    ap.add_argument("--only", default=None, help="time a single routine (profiling runs)")
    ap.add_argument("--no-precision-variant", action="store_true",
                    help="skip comp's FP32+refinement form")
    ap.add_argument("--oplus-sub", type=int, default=10,
                    help="also time the O+ routine at this many sub-cycles (1: skip)")
    ap.add_argument("--scheme", default="auto", choices=("thomas", "pcr", "auto"),
                    help="column solve: thread per column Thomas, wavefront per column PCR, or timed")
    ap.add_argument("--concurrent", action="store_true",
                    help="(default) tgx_ng_step: the independent routines on forked streams")
    ap.add_argument("--serial", action="store_true",
                    help="time tgx_ng_step_serial: the reference's order on one stream")
    ap.add_argument("--roofline-steps", type=int, default=5,
                    help="serial steps after the timed region that time each column kernel alone")
    ap.add_argument("--precision", action="store_true",
                    help="FP32 + FP64-refinement solves against the FP64 solvers on the assembled systems")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of known size (counter calibration)")
    ap.add_argument("--no-nest-form", action="store_true",
                    help="global runs: skip timing the same grid as a nest domain")
    ap.add_argument("--domain", default="global", choices=("global", "nest"),
                    help="the grid as the global model (tgx_glb_step) or as one nest domain (tgx_ng_step)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (the column step of one domain needs no "
              "communication between GPUs; a run over several would give each its own latitude "
              "bands or nests). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = load_pkg()
    cpu = None
    if (not args.no_cpu_baseline and not args.only and not args.precision and args.config != "nests"):
        # before anything touches the GPU: the baseline's worker processes are
        # forked from this one and must not inherit a HIP context
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(pkg, args.config, args.nstep_sub, args.cpu_budget)
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.config == "nests":
        return bench_nests(pkg, args)
    if args.precision:
        return bench_precision(pkg, args)
    glb = args.domain == "global" and pkg.synth.CONFIGS[args.config][6] is None
    g = pkg.synth.make_grid(args.config, nstep_sub=args.nstep_sub, glb=glb)
    print("# generating %s state on the GPU (%s)" % (args.config, "global" if glb else "nest"), file=sys.stderr,
          flush=True)
    if glb:
        F = pkg.synth.make_global_state(g, device="cuda")
        ns = pkg.step.GlobalState(g, F)
    else:
        F = pkg.synth.make_state(g, device="cuda")
        ns = pkg.step.NestState(g, F, scheme=args.scheme)
    del F
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    calls = {r: getattr(ns, r) for r in ROUTINES}
    routines = (args.only,) if args.only else ROUTINES

    if args.copy_calibration:
        n = 1 << 28  # 2 GiB of float64 read + 2 GiB written
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("# copy calibration: %d bytes read, %d bytes written per launch" % (8 * n, 8 * n),
              file=sys.stderr, flush=True)
        del a, b

    serial = args.serial or bool(args.only)
    step_fn = ns.step_serial if serial else ns.step
    if args.only:
        step_fn = calls[args.only]
    for _ in range(args.warmup):  # the first AUTO step times both schemes
        step_fn()
    torch.cuda.synchronize()
    print("# timing %d steps" % args.steps, file=sys.stderr, flush=True)
    # per-routine and per-column-kernel hipEvent pairs, recorded by the
    # library on the stream each runs on, inside the timed region
    pkg.step.profile_begin(args.steps * max(1, args.nstep_sub))
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    for s in range(args.steps):
        step_fn()
        evs[s + 1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof, kern = pkg.step.profile_end()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # per-step times between events on the caller's stream, which every
    # forked stream joins before the step returns
    step_ms = sorted(evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps))
    rt = {r: prof[r][0] / max(prof[r][1], 1) for r in routines}
    kt = {r: kern[r][0] / max(kern[r][1], 1) for r in routines}
    serial_ms = ms_per_step if serial else None
    if not serial:
        # On forked streams the kernels share the GPU, so their event times
        # are not their own: a few steps in the reference's order on one
        # stream time each routine and column kernel alone.
        nrs = max(1, args.roofline_steps)
        pkg.step.profile_begin(nrs * max(1, args.nstep_sub))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for s in range(nrs):
            ns.step_serial()
        torch.cuda.synchronize()
        serial_ms = (time.perf_counter() - t2) * 1e3 / nrs
        prof, kern = pkg.step.profile_end()
        rt = {r: prof[r][0] / max(prof[r][1], 1) for r in routines}
        kt = {r: kern[r][0] / max(kern[r][1], 1) for r in routines}

    nk, cols = g["nk"], g["nlon"] * g["nlat"]
    cols_all = (g["nlon"] + 4) * (g["nlat"] + 4)
    nsub = args.nstep_sub
    unknowns = sum(pkg.step.UNKNOWNS[r] * (nsub if r == "oplus" else 1) for r in routines)
    value = cols * nk * unknowns / (ms_per_step * 1e-3) / 1e6
    schemes = ns.schemes()
    # the dominant kernel: the column kernel with the longest average launch
    dom = max(routines, key=lambda r: kt[r])
    names = pkg.step.COLUMN_KERNELS[dom][0]
    kname = names[1] if schemes[dom] == "pcr" else names[0]
    alg = pkg.step.column_kernel_bytes_per_column_level(dom, schemes[dom]) * cols_all * nk
    achieved = alg / (kt[dom] * 1e-3) / 1e9
    traffic = traffic_from_profiles(kname, args.config + ("-global" if glb else ""), schemes[dom])
    step_bytes = sum(pkg.step.alg_bytes_per_column_level(r) * cols_all * nk * (nsub if r == "oplus" else 1)
                     for r in routines)
    out = {
        "metric": "Mcolumn-level solves/s + %HBM roofline, FP32+refine vs FP64 at <=1e-10 rel",
        "value": round(value, 2),
        "unit": "Mcolumn-level solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded profiles, device-generated)",
        "config": {"workload": "%s %s %dx%d columns x %d levels, %d unknowns/column/level, "
                               "nstep_sub=%d, fused assembly+solve FP64%s" % (
                                   args.config, "global grid (periodic, poles, ring filters)" if glb else
                                   "as one nest domain (lateral Dirichlet edges)", g["nlon"], g["nlat"], nk,
                                   unknowns, nsub, "" if not glb else ", tgx_glb_step"),
                   "domain": "global" if glb else "nest",
                   "columns": cols, "levels": nk, "nstep_sub": nsub, "precision": "fp64",
                   "schedule": "serial" if serial else "concurrent", "scheme": args.scheme,
                   "schemes": schemes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "alg_bytes_per_launch": alg, "launch_ms": round(kt[dom], 4),
                     "timing": "hipEvents around each launch, %s" % (
                         "inside the timed region" if serial else
                         "%d serial steps after the timed region" % max(1, args.roofline_steps))},
        "routine_ms": {r: round(rt[r], 4) for r in routines},
        "column_kernel_ms": {r: round(kt[r], 4) for r in routines},
        "step_ms": {"median": round(step_ms[len(step_ms) // 2], 4), "min": round(step_ms[0], 4),
                    "max": round(step_ms[-1], 4), "timing": "hipEvents between consecutive steps"},
        "step_alg_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "serial_ms_per_step": round(serial_ms, 4),
    }
    if not args.only:
        out["aux_kernels"] = bench_aux_kernels(pkg, g, ns)
    if not args.only and nsub == 1 and args.oplus_sub > 1:
        out["oplus_nstep_sub%d" % args.oplus_sub] = bench_oplus_subcycled(pkg, args, ns)
    nest = ns
    if glb and ((not args.only and not args.no_nest_form) or args.only == "comp") and \
            not (args.no_nest_form and args.no_precision_variant):
        # the same grid as one nest domain: the precision form of comp is built
        # there, and its step time shows what the global passes cost
        nest = None
        del ns
        torch.cuda.empty_cache()
        gn = pkg.synth.make_grid(args.config, nstep_sub=args.nstep_sub)
        F = pkg.synth.make_state(gn, device="cuda")
        nest = pkg.step.NestState(gn, F, scheme=args.scheme)
        del F
        torch.cuda.empty_cache()
        if not args.only and not args.no_nest_form:
            out["nest_form"] = bench_nest_form(pkg, args, nest, cols * nk * unknowns)
    if args.only in (None, "comp") and not args.no_precision_variant:
        out["precision_variant"] = bench_comp_precision(pkg, nest, cols * nk)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def bench_nest_form(pkg, args, nest, colev):
    """tgx_ng_step on the same grid run as one nest domain (lateral Dirichlet
    data, no ring filters): the global form's cost over it."""
    import torch
    for _ in range(args.warmup):
        nest.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nest.step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    return {"ms_per_step": round(ms, 4), "Mcolumn_level_solves_per_s": round(colev / (ms * 1e-3) / 1e6, 2),
            "schemes": nest.schemes(), "steps": args.steps}


def bench_aux_kernels(pkg, g, ns, reps=20):
    """The kernels beside the column step at this grid: the conversion of a
    model-layout field (nk, lond, latd) to level planes and back, and the
    polar ring filter (ringfilter.F, order 3 as for TN at dt.F:486) over
    every latitude the chunk table filters.  HBM bytes are one read and one
    write of what each launch touches."""
    import math
    import torch
    nk, nlon, nlat = g["nk"], g["nlon"], g["nlat"]
    nlond, nlatd = nlon + 4, nlat + 4
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    res = {}
    src = torch.randn(nlatd, nlond, nk, dtype=torch.float64, device="cuda")
    planes = torch.zeros(nk, nlatd, ns.ld, dtype=torch.float64, device="cuda")
    back = torch.empty_like(src)
    nbytes = 2 * 8 * nk * nlond * nlatd
    for name, fn in (("pack_levels", lambda: pkg.tgx.pack_levels(src, planes, nk, nlond, nlatd)),
                     ("unpack_levels", lambda: pkg.tgx.unpack_levels(planes, back, nk, nlond, nlatd))):
        ms = timed(fn)
        res[name] = {"ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    assert torch.equal(back, src)
    del src, planes, back
    chunk = pkg.tgx.RING_CHUNKS.get((nlon, nlat))
    if chunk is not None:
        f = ns.out["tn"].clone()
        rows = 2 * len(chunk)
        ms = timed(lambda: pkg.tgx.ringfilter_dev(f, nlon, nlat, 3, 0, 2 * math.pi / nlon, chunk))
        rbytes = 2 * 8 * rows * nk * (nlon + 4)
        res["ringfilter_tn"] = {"ms": round(ms, 4), "rows": rows, "levels": nk,
                                "GBps": round(rbytes / (ms * 1e-3) / 1e9, 1)}
        del f
    res["timing"] = "torch.cuda events around %d launches each" % reps
    return res


def bench_comp_precision(pkg, ns, colev, reps=10):
    """comp, the step's dominant routine, in its precision form
    (tgx_ng_comp_mixed: FP32 assembly and elimination, FP64 block pivots,
    0, 1 or 2 FP64-residual refinement sweeps) beside the FP64 routine, each
    timed alone on one stream; the error of each form is the largest
    pointwise relative difference from the FP64 outputs over the whole grid,
    with comp's floor 1e-9 (the reference's clamp)."""
    import torch
    fields = ("o2", "o1", "he", "o2_nm", "o1_nm", "he_nm")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        pkg.step.profile_begin(reps)
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        prof, kern = pkg.step.profile_end()
        return (prof["comp"][0] / max(prof["comp"][1], 1), kern["comp"][0] / max(kern["comp"][1], 1))

    ms64 = timed(ns.comp)
    ref = {n: ns.out[n].clone() for n in fields}
    res = {"routine": "comp (O2, O, He)", "unknowns": 3,
           "fp64": {"routine_ms": round(ms64[0], 4), "column_kernels_ms": round(ms64[1], 4),
                    "Mcolumn_level_solves_per_s": round(3 * colev / (ms64[0] * 1e-3) / 1e6, 1)}}
    for nsweep in (0, 1, 2):
        ms = timed(lambda: ns.comp_mixed(nsweep))
        err = {}
        for n in fields:
            d = (ns.out[n] - ref[n]).abs() / torch.clamp(ref[n].abs(), min=1e-9)
            err[n] = float(d.max())
        res["fp32_refine%d" % nsweep] = {
            "routine_ms": round(ms[0], 4), "column_kernels_ms": round(ms[1], 4),
            "Mcolumn_level_solves_per_s": round(3 * colev / (ms[0] * 1e-3) / 1e6, 1),
            "vs_fp64": round(ms64[0] / ms[0], 4), "max_rel_err": {n: float("%.3e" % e) for n, e in err.items()}}
    ns.comp()  # leave the FP64 outputs in place
    torch.cuda.synchronize()
    res["timing"] = "hipEvents around each routine call, %d calls each, one stream" % reps
    return res


def bench_oplus_subcycled(pkg, args, ns1, reps=10):
    """The O+ routine at the reference's default of ten sub-cycles
    (dynamics_ng.F90:99-121) beside ten times the one-sub-cycle routine, each
    called on its own (its chemistry included) on one stream: the sub-cycles
    share one factorisation, so the later ones rebuild only the right-hand
    side."""
    import torch
    nsub = args.oplus_sub

    def timed(ns):
        ns.autotune()  # the scheme choice a step of this shape would make
        ns.oplus()
        torch.cuda.synchronize()
        pkg.step.profile_begin(reps * nsub)
        for _ in range(reps):
            ns.oplus()
        torch.cuda.synchronize()
        prof, _ = pkg.step.profile_end()
        return prof["oplus"][0] / max(prof["oplus"][1], 1)

    ms1 = timed(ns1)
    glb = bool(ns1.grid.get("glb"))
    g = pkg.synth.make_grid(args.config, nstep_sub=nsub, glb=glb)
    if glb:
        F = pkg.synth.make_global_state(g, device="cuda")
        ns = pkg.step.GlobalState(g, F)
    else:
        F = pkg.synth.make_state(g, device="cuda")
        ns = pkg.step.NestState(g, F, scheme=args.scheme)
    del F
    ms = timed(ns)
    ns_scheme = ns.schemes()["oplus"]
    del ns
    torch.cuda.empty_cache()
    return {"routine_ms": round(ms, 4), "nstep_sub1_routine_ms": round(ms1, 4), "scheme": ns_scheme,
            "ratio_to_%dx_nstep_sub1" % nsub: round(ms / (nsub * ms1), 4), "reps": reps,
            "timing": "hipEvents around each oplus routine call, one stream"}


if __name__ == "__main__":
    main()
