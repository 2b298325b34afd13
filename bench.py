"""Benchmark: cells/s for a 1 h integration of saprc99 cells (RTOL 1e-3).

One step = one batched integration of the whole synthetic batch (default:
1,048,576 saprc99 cells, 50 % night / 50 % day with SUN ~ U(0.2, 1),
Rodas3, FP64), inputs resident in HBM; the concentrations are reset from a
pristine device copy at the start of each step.  ``value`` is the tuned path:
the learned scheduler (network inference on the batch: per-cell starting
step and a launch order by predicted step count) followed by the persistent
work-stealing kernel, all inside the timed region.  The plain
one-cell-per-lane kernel (Rosenbrock's default start, cells in index order)
and the persistent kernel without the scheduler are timed on the same cells
and reported beside it, and the CPU oracle (this project's C restatement of
KPP's int/rosenbrock.c, OpenMP over cells) is timed on a bounded sample of
the same cells on the host's cores.

Also timed on the same cells: the persistent kernel with each cell's
starting step estimated from f(t0, y0) (IPAR[4] = 1, the non-learned
alternative to the scheduler), and, for mechanisms built with one, the
cooperative kernel (one wavefront per cell, the LU in LDS).

Accuracy (untimed): on an evenly spaced sample of cells, the final states of
every run are compared with the CPU oracle's FP64 integration of the same
cells at RTOL 1e-10 (ATOL 1e-3); max / median / 99th percentile relative
errors are reported for species above ATOL and above ATOL/RTOL.

Roofline: the solver is bound by HBM bandwidth.  Algorithmic bytes per step
attempt (one unit) follow the streamed working set of each attempt:
  first attempt of a step: 8*[NNZ + 2NNZ + 2NNZ + S*NNZ + NF1*(2NVAR+NFIX+NREACT)
                             + (S^2/2+3S+6)*NVAR]
  each retry:              the same without the Jacobian write and NF2 Fun calls
with NF1/NF2 the Fun evaluations of a first attempt / retry (Rodas3: 4 / 2).
Attempt counts come from the solver's own per-cell counters (Nstp, Njac).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
HERE = os.path.dirname(os.path.abspath(__file__))
# HBM bytes per launch of the tuned kernel on the default workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by the same counters read on
# a copy of known size (``--copy-probe``); the file records the commands.
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_ros_persistent_saprc99_uneven_1M.json")
WEIGHTS = os.path.join(HERE, "kpp_amd", "weights", "saprc99_step_mlp.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_work(mech, method, istat, mixed=False):
    """(bytes, flops) of all step attempts of a batch from its counters.

    mixed: the FP32-LU integrator -- per attempt the FP32 copy of Ghimj
    (8+4 B per entry), the FP32 factorisation (2x4 B per entry), and per
    stage two FP32 solves (4 B per entry each), the FP64 residual over the
    Jacobian's nonzeros plus the diagonal and its three vectors, and the two
    FP32 vector conversions, instead of the FP64 LU and solves."""
    from kpp_amd import solver
    st = mech.stats
    nnz, nv, nf, nr = mech.lu_nonzero, mech.nvar, mech.nfix, mech.nreact
    S = solver.METHOD_STAGES[method]
    nf1, nf2 = {solver.ROS2: (3, 1), solver.ROS3: (3, 1), solver.ROS4: (4, 2),
                solver.RODAS3: (4, 2), solver.RODAS4: (7, 5)}[method]
    vec = (S * S / 2 + 3 * S + 6) * nv
    fun_io = 2 * nv + nf + nr
    if mixed:
        nz = mech.nonzero + nv
        lu = 12 * nnz + 8 * nnz
        per_stage = 8 * nnz + 8 * nz + 24 * nv + 16 * nv
        b_first = 8 * nnz + lu + S * per_stage + 8 * (nf1 * fun_io + vec)
        b_retry = 8 * nnz + lu + S * per_stage + 8 * (nf2 * fun_io + vec)
    else:
        b_first = 8 * (nnz + 2 * nnz + 2 * nnz + S * nnz + nf1 * fun_io + vec)
        b_retry = 8 * (2 * nnz + 2 * nnz + S * nnz + nf2 * fun_io + vec)
    fl = st["flops"]
    f_vec = (S * S + 3 * S + 6) * nv
    f_first = fl["jac"] + (nnz + nv) + fl["decomp"] + S * fl["solve"] + nf1 * fl["fun"] + f_vec
    f_retry = (nnz + nv) + fl["decomp"] + S * fl["solve"] + nf2 * fl["fun"] + f_vec
    nstp = istat[2].astype(np.int64)
    njac = istat[1].astype(np.int64)
    first = int(njac.sum())
    retry = int((nstp - njac).sum())
    return (first * b_first + retry * b_retry, first * f_first + retry * f_retry,
            first + retry, {"bytes_first_attempt": b_first, "bytes_retry": b_retry})


def host_cpu():
    """CPU model, physical cores of the host and of this process's affinity set."""
    model, cores = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            txt = f.read()
        for blk in txt.strip().split("\n\n"):
            kv = dict(l.split(":", 1) for l in blk.splitlines() if ":" in l)
            kv = {k.strip(): v.strip() for k, v in kv.items()}
            model = model or kv.get("model name")
            cores.add((kv.get("physical id"), kv.get("core id")))
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0))
    aff_cores = set()
    for c in aff:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            aff_cores.add((pkg, core))
        except OSError:
            aff_cores.add(("?", c))
    return {"model": model, "host_physical_cores": len(cores) or None,
            "affinity_logical_cpus": len(aff), "affinity_physical_cores": len(aff_cores)}


def cpu_baseline(wl_fn, ncells_full, seed, target_s=20.0):
    """Time the oracle on a bounded sample of the same workload: built here
    with gcc -O2 -march=native, OpenMP threads = the CPUs this process may
    use, bound close to cores."""
    import tempfile
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import ref
    cpu = host_cpu()
    threads = cpu["affinity_logical_cpus"]
    mech = wl_fn(ncells=1, seed=seed).mech
    try:
        path = ref.build_native(mech, tempfile.mkdtemp(prefix="kpp_oracle_"))
        flags = "gcc -O2 -march=native -fopenmp"
    except Exception as e:  # no compiler: the prebuilt portable timing build
        log(f"[bench] native oracle build failed ({e}); using the x86-64-v3 build")
        path, flags = None, "gcc -O2 -march=x86-64-v3 -fopenmp"
    run = lambda bb: ref.integrate(bb, nthreads=threads, fast=path is None, path=path)
    ncal = max(1024, 128 * threads)
    cal = wl_fn(ncells=ncal, seed=seed)
    t = time.time()
    run(cal)
    rate = ncal / max(time.time() - t, 1e-3)
    n = int(min(ncells_full, max(1024, rate * target_s)))
    for _ in range(2):
        # an evenly spaced subset of the batch's cells
        cells = np.linspace(0, ncells_full - 1, n).astype(np.int64)
        b = wl_fn(seed=seed, cells=cells)
        t = time.time()
        out = run(b)
        dt = time.time() - t
        # a calibration that overestimated the time per cell: size the sample again
        if dt >= 0.4 * target_s or n >= ncells_full:
            break
        n = int(min(ncells_full, n * target_s / max(dt, 1e-3)))
    return {"value": n / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{n} evenly spaced cells of the same {ncells_full}-cell batch, "
                      f"oracle ros_ref ({flags}, OpenMP dynamic schedule, "
                      f"OMP_PROC_BIND={os.environ['OMP_PROC_BIND']} "
                      f"OMP_PLACES={os.environ['OMP_PLACES']}), {dt:.1f} s",
            "cpu": cpu, "threads": threads, "compiler_flags": flags,
            "ierr_ok": bool((out["ierr"] == 1).all())}


def copy_probe(nbytes=1 << 30, reps=3):
    """A device-to-device copy of known size, for calibrating HBM counters."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    evs = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        evs.append((e0, e1))
    torch.cuda.synchronize()
    ms = min(e0.elapsed_time(e1) for e0, e1 in evs[1:] or evs)
    # the HBM rate a plain device copy reaches (read + write bytes per second)
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                      "reps": reps, "best_ms": ms, "hbm_gbs": 2 * nbytes / ms / 1e6}))


def measured_traffic(workload, ncells, kernel):
    """HBM bytes per launch of this kernel on this workload from the committed
    rocprofv3 counter passes (not measured in this run), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("ncells") != ncells or t.get("kernel") != kernel:
        return None
    return t.get("bytes_per_launch")


def rel_errors(X, ref, nvar, atol, rtol):
    """max / median / p99 relative error of X against ref for species above thresholds."""
    out = {}
    for name, thr in (("above_atol", atol), ("above_atol_over_rtol", atol / rtol)):
        m = np.abs(ref[:nvar]) > thr
        r = np.abs(X[:nvar] - ref[:nvar])[m] / np.abs(ref[:nvar])[m]
        out[name] = {"max": float(r.max()), "median": float(np.median(r)),
                     "p99": float(np.percentile(r, 99)), "n": int(m.sum())}
    return out


def model_call_bytes(mech):
    """Algorithmic HBM bytes of one call of each model routine on a cell's
    tile (read and written once): Fun reads V, F, RCONST and writes Vdot;
    Jac_SP reads V, F, RCONST and writes the LU_NONZERO entries; a
    factorisation reads and writes the LU entries; a solve reads the LU
    entries and reads and writes the right-hand side."""
    nv, nf, nr, nnz = mech.nvar, mech.nfix, mech.nreact, mech.lu_nonzero
    return {"fun": 8 * (2 * nv + nf + nr), "jac": 8 * (nv + nf + nr + nnz),
            "dec": 8 * 2 * nnz, "sol": 8 * (nnz + 2 * nv)}


def bench_integrator(args):
    """One of the other integrators on the workload's cells: HIP-event time of
    the integration (inputs resident, states reset between steps, untimed),
    cells/s, the roofline from the per-cell counters (Nfun, Njac, Ndec, Nsol
    times each routine's bytes; the complex factorisations and solves of the
    Runge-Kutta integrator count twice), and the CPU oracle of the same
    integrator on a bounded sample of the cells."""
    import torch

    from kpp_amd import solver, workloads
    from oracle import ref

    wl_fn = workloads.WORKLOADS[args.workload]
    n = args.ncells
    b = wl_fn(ncells=n, seed=args.seed)
    mech = solver.mechanism(b.mech)
    dev = torch.device("cuda:0")
    d = b.to_device(dev)
    C0 = d["C"].clone()
    nv = mech.nvar
    name = args.integrator
    ic = np.zeros(20, np.int32)
    rc = np.zeros(20)
    lam0 = None
    if name in ("f90", "h211b"):
        ic[1], ic[14] = 1, 1  # scalar tolerances, rates at every call
        if name == "h211b":
            rc[14], rc[15] = 2.0, 3.0
        label = {"f90": "F90 Rosenbrock (Rodas3)", "h211b": "F90 Rosenbrock with H211b"}[name]
    elif name == "sdirk":
        ic[1] = 1
        label = "SDIRK (Sdirk2A, the default of ICNTRL[2] = 0)"
    elif name == "rk":
        ic[1], ic[9] = 1, 1
        label = "Runge-Kutta (Radau-2A, SDIRK error estimate)"
    else:
        ic[1], ic[2] = 1, 4
        label = "discrete adjoint Rosenbrock (Rodas3, 1 adjoint vector)"
        lam0 = torch.ones((nv, n), dtype=torch.float64, device=dev)
    nchk = 128  # checkpoint slots per cell (a cell needing more ends with -12)
    model = None
    if args.learned_order:
        from kpp_amd import learned
        model = learned.StepModel.load(args.weights, dev)
        span = float(np.max(np.abs(b.tend - b.tstart)))

    def run():
        d["C"].copy_(C0)
        lam = None if lam0 is None else lam0.clone()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        order = None
        if model is not None:
            order = model.schedule(model.features(d["C"], d["temp"], d["sun"]), 1e-6, span)[2]
        if name in ("f90", "h211b"):
            out = solver.integrate_f90(b.mech, d["C"], d["temp"], d["cfactor"], tin=d["tstart"],
                                       tout=d["tend"], atol=b.atol, rtol=b.rtol, ICNTRL=ic,
                                       RCNTRL=rc, sun=d["sun"], h211b=name == "h211b",
                                       variant=solver.PERSISTENT, order=order)
        elif name == "sdirk":
            out = solver.sdirk(b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"],
                               tend=d["tend"], atol=b.atol, rtol=b.rtol, ICNTRL=ic, RCNTRL=rc,
                               sun=d["sun"], sun_mode=b.sun_mode, order=order)
        elif name == "rk":
            out = solver.runge_kutta(b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"],
                                     tend=d["tend"], atol=b.atol, rtol=b.rtol, ICNTRL=ic,
                                     RCNTRL=rc, sun=d["sun"], sun_mode=b.sun_mode, order=order)
        else:
            out = solver.rosenbrock_adj(b.mech, d["C"], d["temp"], d["cfactor"],
                                        tstart=d["tstart"], tend=d["tend"], atol=b.atol,
                                        rtol=b.rtol, Lambda=lam, ICNTRL=ic, RCNTRL=rc, nchk=nchk,
                                        sun=d["sun"], sun_mode=b.sun_mode, order=order)
        e1.record()
        torch.cuda.synchronize()
        code, ierr, istat, rstat = out
        if code != 0:
            raise SystemExit(f"{name} returned {code}")
        return e0.elapsed_time(e1), ierr.cpu().numpy(), istat.cpu().numpy()

    for _ in range(args.warmup):
        run()
    times = []
    for _ in range(args.steps):
        ms, ierr, istat = run()
        times.append(ms)
        log(f"[bench] {name}: {ms:.1f} ms")
    ms = float(np.mean(times))
    by = model_call_bytes(mech)
    cplx = 2 if name == "rk" else 1
    nbytes = float(istat[0].sum() * by["fun"] + istat[1].sum() * by["jac"] +
                   istat[5].sum() * by["dec"] * cplx + istat[6].sum() * by["sol"] * cplx)
    achieved = nbytes / (ms * 1e-3) / 1e9
    # the CPU oracle of the same integrator on a bounded sample of the cells
    cpu = None
    if not args.no_cpu:
        threads = host_cpu()["affinity_logical_cpus"]

        def oracle(ns):
            cells = np.linspace(0, n - 1, ns).astype(np.int64)
            bs = wl_fn(seed=args.seed, cells=cells)
            t = time.time()
            if name in ("f90", "h211b"):
                o = ref.integrate_f90(bs, ic, rc, nthreads=threads, fast=True,
                                      h211b=name == "h211b")
            elif name == "sdirk":
                o = ref.integrate_sdirk(bs, ic, rc, nthreads=threads, fast=True)
            elif name == "rk":
                o = ref.integrate_rk(bs, ic, rc, nthreads=threads, fast=True)
            else:
                o = ref.integrate_adj(bs, np.ones((nv, ns)), ic, rc, cap=nchk, nthreads=threads,
                                      fast=True)
            return o, time.time() - t

        # a calibration sample, then one sized to about 10 s of CPU work
        o, dt = oracle(min(n, 2048))
        ns = int(min(n, max(2048, 2048 * 10.0 / max(dt, 1e-3))))
        o, dt = oracle(ns)
        cpu = {"value": ns / dt, "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": f"{ns} evenly spaced cells of the same {n}-cell batch, the oracle's "
                         f"{name} integrator (gcc -O2 -march=x86-64-v3 -fopenmp), {dt:.1f} s",
               "ierr_ok": bool((o["ierr"] == 1).all())}
    print(json.dumps({
        "metric": f"cells/s integrated ({args.workload}): {label}",
        "value": n / (ms * 1e-3), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "mechanism": b.mech, "ncells": n,
                   "integrator": name, "learned_order": model is not None,
                   "all_ierr_1": bool((ierr == 1).all()),
                   "mean_counters": [float(x) for x in istat.mean(axis=1)]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": nbytes},
        "cpu_baseline": cpu}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="saprc99_uneven")
    ap.add_argument("--ncells", type=int, default=1048576)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--nblocks", type=int, default=0)
    ap.add_argument("--refill-min", type=int, default=0)
    ap.add_argument("--weights", default=WEIGHTS, help="trained step-scheduler network")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-coop", action="store_true",
                    help="skip the cooperative kernel (mechanisms built with one)")
    ap.add_argument("--no-heuristic", action="store_true",
                    help="skip the run with estimated starting steps (IPAR[4] = 1)")
    ap.add_argument("--accuracy-cells", type=int, default=4096)
    ap.add_argument("--eager-tuned", action="store_true",
                    help="run the tuned path as eager calls instead of one graph (same kernels; "
                         "for per-dispatch hardware counters)")
    ap.add_argument("--lds-entries", type=int, default=0,
                    help="with --eager-tuned: LU entries per cell mirrored in LDS (0, 20, 24)")
    ap.add_argument("--lu-fp32", action="store_true",
                    help="also time the mixed-precision integrator and report it as value")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--learned-order", action="store_true",
                    help="with --integrator: launch cells in the order the step "
                         "scheduler predicts (most steps first), inside the timed region")
    ap.add_argument("--integrator", default="",
                    choices=["", "f90", "h211b", "sdirk", "rk", "adj"],
                    help="time one of the other integrators instead (F90 Rosenbrock, H211b, "
                         "SDIRK, Runge-Kutta, discrete adjoint) on --workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): cells are "
              "independent and shard without communication; run one process per GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.integrator:
        bench_integrator(args)
        return

    import torch

    from kpp_amd import solver, workloads

    wl_fn = workloads.WORKLOADS[args.workload]
    t0 = time.time()
    b = wl_fn(ncells=args.ncells, seed=args.seed)
    mech = solver.mechanism(b.mech)
    log(f"[bench] workload {args.workload}: {b.ncells} cells generated in {time.time()-t0:.1f}s")
    dev = torch.device("cuda:0")
    d = b.to_device(dev)
    C0 = d["C"].clone()
    C = d["C"]
    n = b.ncells
    ierr = torch.empty(n, dtype=torch.int32, device=dev)
    istat = torch.empty((8, n), dtype=torch.int32, device=dev)
    rstat = torch.empty((2, n), dtype=torch.float64, device=dev)
    ws = solver.Workspace(dev)
    fallbacks = torch.zeros(n, dtype=torch.int32, device=dev)
    span = float(np.max(np.abs(b.tend - b.tstart)))

    model = None
    if not args.no_learned and b.mech == "saprc99":
        from kpp_amd import learned
        if os.path.exists(args.weights):
            model = learned.StepModel.load(args.weights, dev)
            log(f"[bench] step scheduler weights: {os.path.relpath(args.weights, HERE)}")
        else:
            log("[bench] no trained step scheduler found: training one on step logs of two "
                "fresh batches (seeds 101, 102; untimed)")
            model, _ = learned.train_model(b.mech, wl_fn, [101, 102], args.ncells, dev, log=log)

    # launch shape of the tuned path's eager calls: the autotuner's choice once
    # the graph path has made it, else --lds-entries (for counter runs)
    tuned_launch = {"lds_entries": args.lds_entries} if args.lds_entries else {}

    def step(variant, learned_path, lu_fp32=0, hstart_auto=False):
        C.copy_(C0)
        kw = {"nblocks": args.nblocks if variant == solver.PERSISTENT else 0,
              "refill_min": args.refill_min}
        if (learned_path or hstart_auto) and variant == solver.PERSISTENT:
            kw.update(tuned_launch)
        ipar = b.IPAR
        if hstart_auto:
            ipar = b.IPAR.copy()
            ipar[4] = 1
        ev0, evk, ev1 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        ev0.record()
        hs = order = None
        if learned_path:
            X = model.features(C, d["temp"], d["sun"])
            hs, _, order = model.schedule(X, 1e-6, span)
        evk.record()
        code, _, _, _ = solver.rosenbrock(
            b.mech, C, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
            atol=b.atol, rtol=b.rtol, IPAR=ipar, RPAR=b.RPAR, sun=d["sun"],
            sun_mode=b.sun_mode, hstart=hs, order=order, ierr=ierr, istat=istat, rstat=rstat,
            variant=variant, lu_fp32=lu_fp32, fallbacks=fallbacks, workspace=ws, **kw)
        ev1.record()
        assert code == 0, code
        return ev0, evk, ev1

    def timed(variant, learned_path, steps, warmup, name, lu_fp32=0, hstart_auto=False):
        for _ in range(warmup):
            step(variant, learned_path, lu_fp32, hstart_auto)
        torch.cuda.synchronize()
        evs = []
        t = time.perf_counter()
        for k in range(steps):
            evs.append(step(variant, learned_path, lu_fp32, hstart_auto))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        dev_t = [a.elapsed_time(e) / 1e3 for a, _, e in evs]
        kern = [k.elapsed_time(e) / 1e3 for _, k, e in evs]
        sched = [a.elapsed_time(k) / 1e3 for a, k, _ in evs]
        ok = bool((ierr == 1).all().item())
        ist = istat.cpu().numpy()
        log(f"[bench] {name}: {steps} steps, {wall / steps * 1e3:.1f} ms/step, "
            f"device {np.mean(dev_t) * 1e3:.1f} ms (scheduler {np.mean(sched) * 1e3:.2f} ms, "
            f"integrator {np.mean(kern) * 1e3:.1f} ms), attempts/cell {ist[2].mean():.2f}, "
            f"all IERR=1: {ok}")
        return {"wall": wall / steps, "device_s": float(np.mean(dev_t)),
                "kernel_s": float(np.mean(kern)), "sched_s": float(np.mean(sched)), "ok": ok,
                "istat": ist, "C": C.cpu().numpy(), "fallbacks": fallbacks.cpu().numpy()}

    def timed_graph(steps, warmup, name):
        """The tuned step recorded once as a graph (features, network, bucket
        sort, integrator), replayed on its own stream."""
        X = model.features(C, d["temp"], d["sun"])
        sched = model.schedule(X, 1e-6, span)
        torch.cuda.synchronize()
        t = time.time()
        best, best_ms, tried = solver.autotune(
            b.mech, C, C0, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
            atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"],
            sun_mode=b.sun_mode, hstart=sched[0], order=sched[2])
        launch = {"variant": best[0], "nblocks": best[1], "refill_min": best[2], "lanes": best[3],
                  "lds_entries": best[4]}
        log(f"[bench] autotuner: {tried} launch shapes in {time.time() - t:.1f}s, best "
            f"{launch} at {best_ms:.1f} ms")
        tuned_launch.update({k: v for k, v in launch.items() if k != "variant"})

        def body():
            model.features(C, d["temp"], d["sun"], out=X)
            model.schedule(X, 1e-6, span, out=sched)
            code, _, _, _ = solver.rosenbrock(
                b.mech, C, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
                atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"],
                sun_mode=b.sun_mode, hstart=sched[0], order=sched[2], ierr=ierr, istat=istat,
                rstat=rstat, workspace=ws, **launch)
            assert code == 0, code

        C.copy_(C0)
        body()  # sizes every buffer: nothing allocates while recording
        torch.cuda.synchronize()
        gs = torch.cuda.Stream()
        graph = solver.capture(body, gs)

        def one():
            with torch.cuda.stream(gs):
                C.copy_(C0)
                e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
                e0.record(gs)
                graph.launch(gs)
                e1.record(gs)
            return e0, e1

        for _ in range(warmup):
            one()
        torch.cuda.synchronize()
        t = time.perf_counter()
        evs = [one() for _ in range(steps)]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        dev_t = float(np.mean([a.elapsed_time(e) / 1e3 for a, e in evs]))
        graph.close()
        ok = bool((ierr == 1).all().item())
        ist = istat.cpu().numpy()
        log(f"[bench] {name}: {steps} graph replays, {wall * 1e3:.1f} ms/step, device "
            f"{dev_t * 1e3:.1f} ms, attempts/cell {ist[2].mean():.2f}, all IERR=1: {ok}")
        return {"wall": wall, "device_s": dev_t, "kernel_s": dev_t, "sched_s": None, "ok": ok,
                "istat": ist, "C": C.cpu().numpy(), "launch": launch,
                "autotune": {"tried": tried, "best_ms": best_ms}}

    runs = {}
    if model is not None:
        if args.eager_tuned:
            runs["tuned"] = timed(solver.PERSISTENT, True, args.steps, args.warmup,
                                  "tuned path, eager calls")
        else:
            runs["tuned"] = timed_graph(args.steps, args.warmup,
                                        "tuned (learned schedule + autotuned persistent, one graph)")
            eager = timed(solver.PERSISTENT, True, 1, 0, "tuned path, eager calls (for the split)")
            runs["tuned"]["sched_s"] = eager["sched_s"]
            # the integrator's share of a replay: replay time less the scheduler's
            # time on the eager step (the graph's kernels are not timed one by one)
            runs["tuned"]["kernel_s"] = runs["tuned"]["device_s"] - eager["sched_s"]
    runs["persistent"] = timed(solver.PERSISTENT, False, max(1, args.steps if model is None else 1),
                               args.warmup, "persistent, default start")
    if not args.no_heuristic:
        # the non-learned alternative: each cell's starting step estimated from
        # f(t0, y0) (Hairer-Norsett-Wanner, IPAR[4] = 1), same launch shape as
        # the tuned path, cells in index order
        runs["heuristic"] = timed(solver.PERSISTENT, False, max(1, args.steps if model is None else 1),
                                  args.warmup, "persistent, estimated start (IPAR[4]=1)",
                                  hstart_auto=True)
    if args.lu_fp32:
        runs["mixed"] = timed(solver.PERSISTENT, False, max(1, args.steps), args.warmup,
                              "persistent, FP32 LU + FP64 refinement", lu_fp32=1)
    # large mechanisms: one wavefront per cell, the LU in LDS
    coop_ok = solver.workspace_bytes(b.mech, b.method, n, variant=solver.COOP) > 0
    if coop_ok and model is None and not args.no_coop:
        runs["coop"] = timed(solver.COOP, False, max(1, args.steps), args.warmup,
                             "cooperative kernel (one wavefront per cell, LU in LDS)")
    # small mechanisms: the kernel that keeps each cell's working set in VGPRs
    reg_ok = solver.workspace_bytes(b.mech, b.method, n, variant=solver.REGISTER) > 0
    if reg_ok and model is None and not args.lu_fp32:
        runs["register"] = timed(solver.REGISTER, False, max(1, args.steps), args.warmup,
                                 "register kernel (working set in VGPRs)")
    if not args.no_plain:
        runs["plain"] = timed(solver.PLAIN, False, max(1, args.steps), max(0, args.warmup), "plain")
    top = runs["tuned"] if model is not None else (
        runs["mixed"] if args.lu_fp32 else
        runs.get("coop", runs.get("register", runs["persistent"])))
    top_register = model is None and not args.lu_fp32 and "register" in runs and "coop" not in runs
    top_coop = model is None and not args.lu_fp32 and "coop" in runs

    # roofline of the persistent kernel on the tuned run's attempts
    top_mixed = (model is None and args.lu_fp32)
    top_lds = tuned_launch.get("lds_entries", 0) if model is not None else 0
    top_kernel = (f"ros_persistent_kernel<kppm_{b.mech},{solver.METHOD_STAGES[b.method]}"
                  f"{',MIX' if top_mixed else ''}{',LDS' + str(top_lds) if top_lds else ''}>")
    if top_coop:
        top_kernel = f"ros_coop_kernel<kppm_{b.mech},{solver.METHOD_STAGES[b.method]}>"
    nbytes, nflops, attempts, bterms = algorithmic_work(mech, b.method, top["istat"], top_mixed)
    kern_t = top["kernel_s"]

    accuracy = None
    if args.accuracy_cells > 0:
        # reference: the CPU oracle (the restatement of KPP's C Rosenbrock) in
        # FP64 at RTOL 1e-10, ATOL 1e-3 on an evenly spaced sample of the cells
        import dataclasses

        from oracle import ref
        cells = np.linspace(0, n - 1, min(n, args.accuracy_cells)).astype(np.int64)
        tb = wl_fn(seed=args.seed, cells=cells)
        tb = dataclasses.replace(tb, atol=np.full_like(tb.atol, 1e-3),
                                 rtol=np.full_like(tb.rtol, 1e-10))
        t = time.time()
        o = ref.integrate(tb, nthreads=len(os.sched_getaffinity(0)), fast=True)
        ref_C = o["C"]
        log(f"[bench] accuracy reference: CPU oracle, RTOL 1e-10, ATOL 1e-3, {len(cells)} cells "
            f"in {time.time() - t:.1f}s, all IERR=1: {bool((o['ierr'] == 1).all())}")
        at, rt = float(b.atol[0]), float(b.rtol[0])
        accuracy = {"reference": f"CPU oracle (ros_ref, FP64) at RTOL 1e-10 ATOL 1e-3 on "
                                 f"{len(cells)} evenly spaced cells of the batch",
                    "reference_all_ierr_1": bool((o["ierr"] == 1).all()),
                    "sample_cells": int(len(cells))}
        for k, r in runs.items():
            accuracy[k] = rel_errors(r["C"][:, cells], ref_C, mech.nvar, at, rt)
        if "mixed" in runs:
            accuracy["mixed_vs_fp64"] = rel_errors(runs["mixed"]["C"][:, cells],
                                                   runs["persistent"]["C"][:, cells], mech.nvar,
                                                   at, rt)
        if "plain" in runs and "tuned" in runs:
            accuracy["tuned_vs_plain"] = rel_errors(runs["tuned"]["C"][:, cells],
                                                    runs["plain"]["C"][:, cells], mech.nvar, at, rt)

    # PCIe-inclusive rate: the batch's inputs (C, TEMP, CFACTOR, SUN, Tstart,
    # Tend) copied in from pinned host memory and C copied back, around the
    # top path's device time (not `value`, which starts with inputs in HBM)
    ins = [C0, d["temp"], d["cfactor"], d["sun"], d["tstart"], d["tend"]]
    hosts = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for t in ins]
    for h, t in zip(hosts, ins):
        h.copy_(t)
    scratch = [torch.empty_like(t) for t in ins]
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    torch.cuda.synchronize()
    e0.record()
    for t, h in zip(scratch, hosts):
        t.copy_(h, non_blocking=True)
    e1.record()
    hosts[0].copy_(scratch[0], non_blocking=True)
    e2.record()
    torch.cuda.synchronize()
    h2d, d2h = e0.elapsed_time(e1) / 1e3, e1.elapsed_time(e2) / 1e3
    del scratch
    pcie = {"value": n / (h2d + top["device_s"] + d2h), "unit": "cells/s",
            "h2d_ms": h2d * 1e3, "d2h_ms": d2h * 1e3,
            "bytes_in": int(sum(t.numel() * t.element_size() for t in ins)),
            "bytes_out": int(C0.numel() * C0.element_size())}
    log(f"[bench] PCIe-inclusive: H2D {h2d * 1e3:.1f} ms, D2H {d2h * 1e3:.1f} ms, "
        f"{pcie['value']:.0f} cells/s")

    cpu = None
    if not args.no_cpu:
        log("[bench] timing the CPU oracle on a sample ...")
        cpu = cpu_baseline(wl_fn, n, args.seed)
        log(f"[bench] cpu baseline {cpu['value']:.0f} cells/s on {cpu['cores']} threads")
    value = n / top["wall"]
    achieved = nbytes / kern_t / 1e9
    if top_register:
        # no tile traffic: the cell's working set stays in VGPRs, so the bound
        # is FP64 vector arithmetic (the emitted operation counts per attempt)
        roof = {"bound": "fp64", "achieved": nflops / kern_t / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": nflops / kern_t / 1e12 / FP64_PEAK_TFLOPS,
                "traffic": None, "kernel": f"ros_reg_kernel<kppm_{b.mech},"
                                           f"{solver.METHOD_STAGES[b.method]}>",
                "kernel_ms": kern_t * 1e3, "algorithmic_flops": nflops}
    else:
        roof = None
    plain = runs.get("plain")
    line = {
        "metric": "cells/s integrated 1 h (saprc99, RTOL 1e-3): tuned vs plain GPU vs CPU",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": top["wall"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "mechanism": b.mech, "ncells": n,
                   "method": "Rodas3" if b.method == solver.RODAS3 else int(b.method),
                   "rtol": float(b.rtol[0]), "atol": float(b.atol[0]), "integrate_s": 3600,
                   "path": ("learned step scheduler (BF16 MFMA network: per-cell first step, "
                            "launch order by predicted step count) + autotuned persistent "
                            "work-stealing kernel, recorded as one graph"
                            if model is not None else
                            ("persistent work-stealing kernel, FP32 LU + FP64 refinement"
                             if args.lu_fp32 else
                             ("cooperative kernel: one wavefront per cell, LU in LDS"
                              if top_coop else
                              ("register kernel: one lane per cell, working set in VGPRs"
                               if top_register else "persistent work-stealing kernel")))),
                   "all_ierr_1": top["ok"], "attempts_per_cell": attempts / n},
        "roofline": roof or {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, n, top_kernel),
                     "traffic_source": (os.path.relpath(TRAFFIC_FILE, HERE) +
                                        ": rocprofv3 FETCH_SIZE/WRITE_SIZE passes of this kernel "
                                        "on this workload, not measured in this run"
                                        if measured_traffic(args.workload, n, top_kernel)
                                        else None),
                     "kernel": top_kernel,
                     "kernel_ms": kern_t * 1e3, "algorithmic_bytes": nbytes,
                     "fp64_tflops": nflops / kern_t / 1e12,
                     "fp64_frac": nflops / kern_t / 1e12 / FP64_PEAK_TFLOPS, **bterms},
        "scheduler_ms": (top["sched_s"] or 0.0) * 1e3,
        "launch": top.get("launch"),
        "autotune": top.get("autotune"),
        "plain": ({"value": n / plain["wall"], "kernel_ms": plain["kernel_s"] * 1e3,
                   "attempts_per_cell": float(plain["istat"][2].mean()),
                   "all_ierr_1": plain["ok"]} if plain else None),
        "coop": ({"value": n / runs["coop"]["wall"], "kernel_ms": runs["coop"]["kernel_s"] * 1e3,
                  "attempts_per_cell": float(runs["coop"]["istat"][2].mean()),
                  "speedup_vs_persistent": runs["persistent"]["wall"] / runs["coop"]["wall"],
                  "all_ierr_1": runs["coop"]["ok"]} if "coop" in runs else None),
        "register": ({"value": n / runs["register"]["wall"],
                      "kernel_ms": runs["register"]["kernel_s"] * 1e3,
                      "attempts_per_cell": float(runs["register"]["istat"][2].mean()),
                      "all_ierr_1": runs["register"]["ok"]} if "register" in runs else None),
        "persistent_default_start": {"value": n / runs["persistent"]["wall"],
                                     "attempts_per_cell": float(runs["persistent"]["istat"][2].mean())},
        "persistent_estimated_start": ({"value": n / runs["heuristic"]["wall"],
                                        "attempts_per_cell":
                                            float(runs["heuristic"]["istat"][2].mean()),
                                        "launch": dict(tuned_launch) or None,
                                        "all_ierr_1": runs["heuristic"]["ok"]}
                                       if "heuristic" in runs else None),
        "mixed": ({"value": n / runs["mixed"]["wall"],
                   "kernel_ms": runs["mixed"]["kernel_s"] * 1e3,
                   "attempts_per_cell": float(runs["mixed"]["istat"][2].mean()),
                   "fallback_attempts": int(runs["mixed"]["fallbacks"].sum()),
                   "cells_with_fallback": int((runs["mixed"]["fallbacks"] > 0).sum()),
                   "speedup_vs_fp64_persistent": runs["persistent"]["wall"] / runs["mixed"]["wall"],
                   "all_ierr_1": runs["mixed"]["ok"]} if "mixed" in runs else None),
        "speedup_vs_plain": (value / (n / plain["wall"])) if plain else None,
        "pcie_inclusive": pcie,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
