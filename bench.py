"""Benchmark: converged bias-point solves/s of the 2D n-MOSFET Id-Vg sweep.

Workload: 200x120 MOSFET (0.2 nm first cell at the Si/oxide interface,
L = 100 nm, t_ox = 2 nm, N_A = 1e17, 1e20 Gaussian S/D, Lombardi surface
mobility), Vd = 0.05 V, Vg 0 -> 1.0 V in 21 warm-started points, as
mosfet.id_vg_sweep runs it (equilibrium, then 21 solve_bias, drain
current after each).  One step = one full sweep (equilibrium + 21
converged bias points + terminal currents).  Inputs are built once and
resident on the GPU before the timed region.

Prints ONE JSON line.  `value` is the faster of the two GPU paths
(FP64 BiCGStab, or FP32-stored matrix BiCGStab inside FP64 refinement)
whose drain currents are within 1e-6 relative of the FP64 CPU oracle on
the oracle's sample points and of the FP64 GPU path on every point.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "tcad-dev_amd"), os.path.join(REPO, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = ("converged bias-point solves/sec at <=1e-6 current err vs FP64; "
          "p50/p99 latency")


def run_sweep(dev, spec, opts, lat=None):
    """One step: equilibrium + every bias point; returns drain currents."""
    import torch
    dev.solve_equilibrium(opts)
    out = []
    for bias in spec.sweep:
        t0 = time.perf_counter()
        dev.solve_bias(bias, opts)
        I = dev.terminal_currents()
        torch.cuda.synchronize()
        if lat is not None:
            lat.append(time.perf_counter() - t0)
        out.append(I)
    return out


def measure(spec, linsolve, steps, warmup, rtol=1e-10, forcing=False):
    import torch
    from tcadhip.device import NewtonOptions, device_from_spec
    opts = NewtonOptions(linsolve=linsolve, linsolve_rtol=rtol, forcing=forcing)
    dev = device_from_spec(spec)
    for _ in range(warmup):
        run_sweep(dev, spec, opts)
    torch.cuda.synchronize()
    lat = []
    its, kry = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        cur = run_sweep(dev, spec, opts, lat)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # per-point Newton / Krylov counts of the last sweep (re-run, untimed)
    dev.solve_equilibrium(opts)
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        kry.append(dev.last_krylov_iterations)
    drain = np.array([c["drain"] for c in cur])
    return dict(dev=dev, seconds=dt, latencies=np.array(lat), drain=drain,
                newton=its, krylov=kry)


def kernel_timing(dev, which, reps=200):
    import ctypes
    from tcadhip import _lib
    out = ctypes.c_double()
    _lib.check(_lib.lib().tcad_time_kernel(dev._gpu().handle, which, reps, ctypes.byref(out)),
               "tcad_time_kernel")
    return out.value


def cpu_baseline(spec, npoints, budget_s=30.0):
    """The FP64 oracle (NumPy + SciPy SuperLU, single thread) on the
    first `npoints` bias points of the same sweep, equilibrium included."""
    from oracle_build import oracle_device
    dev = oracle_device(spec)
    t0 = time.perf_counter()
    dev.solve_equilibrium()
    cur = []
    done = 0
    for bias in spec.sweep[:npoints]:
        dev.solve_bias(bias)
        cur.append(dev.terminal_current("drain"))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return done, dt, np.array(cur)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-points", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--points", type=int, default=None,
                    help="first K bias points only (short profiling runs)")
    ap.add_argument("--modes", default="bicgstab,bicgstab_mixed")
    ap.add_argument("--rtol", type=float, default=1e-10,
                    help="relative residual of each Newton linear solve")
    ap.add_argument("--forcing", action="store_true",
                    help="Newton forcing terms (inexact early solves)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (the reference "
                         "kernel for scaling HBM traffic counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every MI355X box here has one, and this "
              f"workload (one device sweep) is not sharded; --gpus {args.gpus} is not "
              f"supported.", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        n = args.copy_calibration << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        a.fill_(1)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 5}))
        return
    import warnings
    warnings.simplefilter("ignore")
    from tcadhip import configs

    spec = configs.config2_mosfet()
    if args.points:
        spec.sweep = spec.sweep[:args.points]
    npts = len(spec.sweep)
    res = {}
    modes = args.modes.split(",")
    for mode in modes:
        print(f"[bench] {mode}: warmup {args.warmup}, steps {args.steps}", file=sys.stderr,
              flush=True)
        res[mode] = measure(spec, mode, args.steps, args.warmup, args.rtol,
                            forcing=args.forcing)
        r = res[mode]
        print(f"[bench] {mode}: {npts * args.steps / r['seconds']:.3f} points/s "
              f"newton={r['newton']} krylov={r['krylov']}", file=sys.stderr, flush=True)

    ref64 = res["bicgstab"]["drain"] if "bicgstab" in res else res[modes[0]]["drain"]
    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu oracle on {args.cpu_points} points", file=sys.stderr, flush=True)
        done, dt, ccur = cpu_baseline(spec, args.cpu_points)
        cpu = dict(points=done, seconds=dt, drain=ccur)

    def err_vs(a, b):
        return float(np.max(np.abs(a - b) / np.abs(b)))

    chosen, best = None, -1.0
    for mode, r in res.items():
        ok = err_vs(r["drain"], ref64) <= 1e-6
        if cpu is not None:
            r["err_vs_oracle"] = err_vs(r["drain"][:cpu["points"]], cpu["drain"])
            ok = ok and r["err_vs_oracle"] <= 1e-6
        r["ok"] = ok
        rate = npts * args.steps / r["seconds"]
        if ok and rate > best:
            chosen, best = mode, rate
    if chosen is None:
        chosen = "bicgstab"
        best = npts * args.steps / res[chosen]["seconds"]
    r = res[chosen]

    # roofline: the stencil SpMV inside BiCGStab dominates the kernel time;
    # algorithmic bytes per launch = N * (9 + 5 * n_neighbour_blocks) * w
    # matrix entries + 3 reads of the source vector + 3 writes of the
    # product + 3 reads of the dotted vector (FP64 vectors), per node
    dev = r["dev"]
    N = spec.N
    nx, ny = spec.coords[0].size, spec.coords[1].size
    edges = 2 * ((nx - 1) * ny + nx * (ny - 1))          # directed neighbour blocks
    w = 4 if chosen == "bicgstab_mixed" else 8
    spmv_bytes = (9 * N + 5 * edges) * w + 3 * N * 8 * 3
    which = 1 if chosen == "bicgstab_mixed" else 0
    us = kernel_timing(dev, which)
    achieved = spmv_bytes / (us * 1e-6) / 1e9
    lat_ms = r["latencies"] * 1e3
    line = {
        "metric": METRIC,
        "value": round(best, 4),
        "unit": "bias points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 matrix / f64 vectors + f64 residual" if chosen == "bicgstab_mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "mosfet_200x120_idvg_vd0.05_vg0-1_21pts", "nodes": N,
                   "unknowns": 3 * N, "bias_points_per_step": npts, "path": chosen},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "paths": {m: {"points_per_s": round(npts * args.steps / v["seconds"], 4),
                      "newton_per_point": v["newton"], "krylov_per_point": v["krylov"],
                      "max_rel_err_vs_fp64_gpu": err_vs(v["drain"], ref64),
                      "max_rel_err_vs_oracle": v.get("err_vs_oracle"),
                      "meets_1e-6": v["ok"]} for m, v in res.items()},
        "roofline": {"bound": "hbm", "kernel": "k_bicg_spmv<%s>" % ("float" if w == 4 else "double"),
                     "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 5), "traffic": None,
                     "bytes_per_launch": spmv_bytes, "avg_launch_us": round(us, 3)},
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["points"] / cpu["seconds"], 5), "unit": "bias points/s",
            "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (NumPy + SciPy SuperLU) equilibrium + first "
                      f"{cpu['points']} of {npts} Vg points of the same 200x120 sweep, "
                      f"{cpu['seconds']:.1f} s"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
