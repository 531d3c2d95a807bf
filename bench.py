"""Benchmark: converged bias-point solves/s of the 3D tri-gate Id-Vg sweeps.

Headline workload (the largest single-GPU configuration): the 64x48x48
tri-gate (147,456 nodes, 442,368 unknowns; fin 10 nm wide and 30 nm tall,
channel 1e16, source/drain 1e20 with 5 nm Gaussian lateral roll-off,
1 nm oxide, top and side gates), Vd in {0.05, 0.7} V x Vg 0 -> 0.8 V in 17
warm-started points, as finfet3d.id_vg_sweep_3d runs it (equilibrium, then
every Vg point, terminal currents after each).  One step = both sweeps
(2 equilibria + 34 converged bias points + their currents).  Inputs are
built once and resident on the GPU before the timed region.

Prints ONE JSON line.  `value` is the FP64 path's bias points/s over the
timed steps.  Parity is checked where the reference's own numbers exist:
the reduced 24x18x14 tri-gate at both drain biases against the reference
fixtures (tests/golden/make_golden_3d.py), strict 1e-6 with no floor.
`roofline` is the dominant kernel's in-solver average (HIP events around
every launch of it during one extra, untimed step).  `cpu_baseline` is
the FP64 CPU restatement at full size with the reference's iterative
method for large 3D systems (GMRES(100) + node block-Jacobi, rtol 1e-10),
on the first bias point from the GPU's equilibrium state, under a stated
cap.  Extras (--extras): the 60x40 diode sweep, the 200x120 MOSFET sweep
(FP64 and FP32-matrix paths), small-signal AC, 256 batched diodes and the
serving loop at sustainable and saturating arrival rates.

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
HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes of the BiCGStab kernels on one 3D device (each distinct
# array element once): 39 FP64 stencil entries per node, FP64 vectors of 3
# planes (24 B per node), a node's tile-inverse rows 36 x 4 FP16 + 3 FP32
# scales (300 B).
TILE_ROW_BYTES = 36 * 8 + 3 * 4
PROBE_KERNELS = {0: "k_bicg_spmv<double>", 1: "k_bicg_s_tmem<double>",
                 2: "k_bicg_spmv_t<double>", 3: "k_bicg_u_tmem<double>",
                 4: "k_bicg_black_tmem<double>"}


def krylov_bytes(shape, rb=True):
    """Algorithmic bytes per launch of BiCGStab kernel 0..4 on a 3D device
    of shape (nz, ny, nx) with 2x1x8-node tiles:
      0 k_bicg_spmv: matrix 312, p-hat 24, v 24 (written), r-hat 24;
      1 k_bicg_s: r, v 48, s 24 (written), s-hat 24 (written) and the tile
        rows of the red tiles' nodes with the red-black ordering (rb), of
        every node without;
      2 k_bicg_spmv_t: matrix 312, s-hat 24, t 24 (written), s, r-hat 48;
      3 k_bicg_u: x, p-hat, s-hat, s, t, p, v 168, x, r, p 72 (written),
        p-hat 24 (written) and tile rows as kernel 1;
      4 k_bicg_black (rb only): for each black-tile node its source 24,
        tile rows 300, result 24 (written) and 20 (5 FP32) per existing
        neighbour in another tile, and each such red neighbour's value
        (24) once."""
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = (k * ny + j) * nx + i
    N = nx * ny * nz
    black = ((i // 2 + j + k // 8) & 1) == 1
    nb = int(black.sum()) if rb else 0
    nr = N - nb
    out = {0: N * (312 + 72), 2: N * (312 + 96),
           1: N * 72 + nr * (24 + TILE_ROW_BYTES), 3: N * 240 + nr * (24 + TILE_ROW_BYTES)}
    if not rb:
        return out
    moves = [((i % 2 == 0) & (i > 0), -1), ((i % 2 == 1) & (i < nx - 1), 1),
             (j > 0, -nx), (j < ny - 1, nx),
             ((k % 8 == 0) & (k > 0), -nx * ny), ((k % 8 == 7) & (k < nz - 1), nx * ny)]
    adj = np.zeros(N, dtype=bool)
    links = 0
    for cond, step in moves:
        sel = black & cond
        links += int(sel.sum())
        adj[g[sel] + step] = True
    out[4] = nb * (24 + TILE_ROW_BYTES + 24) + links * 20 + int(adj.sum()) * 24
    return out


# Launches issued after a solve converged return before reading the vectors:
# with HIP events on the 64x48x48 tri-gate they take 2.5-5 us, the working
# launches of every BiCGStab kernel 10 us and more.
EARLY_EXIT_US = 5.2


def probe_min_us(kernel, shape):
    """Launches of the probed kernel shorter than this are left out as early
    exits: under EARLY_EXIT_US, or so short that their algorithmic bytes
    would have moved at twice the HBM peak."""
    return max(EARLY_EXIT_US, krylov_bytes(shape)[kernel] / (2 * HBM_PEAK_GBPS * 1e9) * 1e6)


def _say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _sync():
    import torch
    torch.cuda.synchronize()


def _lib():
    from tcadhip import _lib as L
    return L


# ---------------------------------------------------------------------------
# headline: the tri-gate sweeps
# ---------------------------------------------------------------------------
def trigate_step(devs, specs, opts, eq_opts, lat=None, record=None, predictor=True,
                 opts_start=None):
    """One step: for each drain bias, equilibrium + every Vg point, each
    point warm-started from the previous one (with `predictor`, from the
    secant prediction of the last two; tcadhip.sweep.device_sweep).  A
    point's latency runs from the end of the previous point to its drain
    current."""
    from tcadhip import sweep
    for dev, spec in zip(devs, specs):
        dev.solve_equilibrium(eq_opts)
        t_last = [time.perf_counter()]

        def on_point(k, its, kry, dev=dev, spec=spec, t_last=t_last):
            I = dev.terminal_currents()
            t = time.perf_counter()
            if lat is not None:
                lat.append(t - t_last[0])
            t_last[0] = t
            if record is not None:
                record.append((spec.name, its, kry, I["drain"]))

        sweep.device_sweep(dev, spec.sweep, opts, predictor=predictor, on_point=on_point,
                           opts_start=opts_start)


# Newton iterations a tuning candidate may need above the untuned solve when
# both start from the predicted state: that start usually saves one iteration
# against the previous converged state (the reference's start), so such a
# candidate still needs no more iterations than the untuned solve from
# there, and the sweep stays within +-1 of the exact counts
PREDICTED_SLACK = 1


def tuning_start(dev, sw, opts, eq_opts, predictor):
    """Leave `dev` at the start the sweep gives the point the autotuner
    times, and return that point's index: the second point warm-started
    from the first, or with the predictor the third, from its secant
    prediction off the first two."""
    from tcadhip import sweep
    dev.solve_equilibrium(eq_opts)
    dev.solve_bias(sw[0], opts)
    if not predictor or len(sw) < 3:
        return 1
    B = dev._gpu()
    names = list(dev._s.bcs)
    with B.on_stream():
        prev = (B.psi.clone(), B.n.clone(), B.p.clone())
    v0 = sweep.bias_vector(dev._s, names)
    dev.solve_bias(sw[1], opts)
    v1 = sweep.bias_vector(dev._s, names)
    dev._set_voltages(sw[2])
    f = sweep.secant_ratio(v0, v1, sweep.bias_vector(dev._s, names))
    with B.on_stream():
        B.predict(prev, np.array([f]))
    B.stream_sync()
    return 2


def tune_sweep(dev, sw, opts, eq_opts, predictor):
    """The autotuner for a sweep: the options of the points that start from
    the previous converged state, tuned on the second point warm-started
    from the first, and with the predictor those of the predicted points,
    tuned on the third point from its prediction (newton_slack
    PREDICTED_SLACK).  Returns (options of predicted points, options of the
    other points, report)."""
    from tcadhip import tune as tuner
    tuning_start(dev, sw, opts, eq_opts, False)
    warm, rep_w = tuner.tune(dev, sw[1], opts, say=_say, start="warm")
    if not predictor or len(sw) < 3:
        return warm, warm, rep_w
    k = tuning_start(dev, sw, opts, eq_opts, True)
    pred, rep_p = tuner.tune(dev, sw[k], opts, say=_say, newton_slack=PREDICTED_SLACK,
                             start="predicted")
    return pred, warm, {"predicted_start": rep_p, "warm_start": rep_w}


def parse_options(text):
    """NewtonOptions fields from "key=value,key=value" (bench.py --set and
    --start-set), converted to each field's type."""
    import dataclasses
    from tcadhip.device import NewtonOptions
    types = {f.name: f.type for f in dataclasses.fields(NewtonOptions)}
    out = {}
    for item in filter(None, (t.strip() for t in text.split(","))):
        k, v = item.split("=", 1)
        if k not in types:
            raise SystemExit(f"unknown NewtonOptions field {k!r}")
        t = types[k] if isinstance(types[k], type) else \
            {"bool": bool, "int": int, "float": float, "str": str}[types[k]]
        out[k] = (v.lower() in ("1", "true", "yes")) if t is bool else t(v)
    return out


def bench_trigate(steps, warmup, mode, points=None, probe=4, forcing=False, precond="tile",
                  tune=True, forcing_slope=0.01, forcing_max=1e-4, tile_rows="auto",
                  probe_cuts=(), forcing_pow=1, predictor=True, overrides=None,
                  start_overrides=None):
    from tcadhip import configs
    from tcadhip.device import NewtonOptions, device_from_spec
    specs = [configs.config3_trigate(0.05), configs.config3_trigate(0.7)]
    if points:
        for s in specs:
            s.sweep = s.sweep[:points]
    devs = [device_from_spec(s) for s in specs]
    opts = NewtonOptions(linsolve=mode, forcing=forcing, precond=precond,
                         forcing_slope=forcing_slope, forcing_max=forcing_max,
                         forcing_pow=forcing_pow, tile_rows=tile_rows)
    eq_opts = NewtonOptions(max_iter=400)
    tuning = None
    opts_start = None
    if overrides or start_overrides:
        # a fixed configuration (profiling runs): the options of the
        # predicted points, and of the points without a prediction
        from dataclasses import replace
        base = opts
        opts = replace(base, **(overrides or {}))
        opts_start = replace(base, **(start_overrides or {}))
    if tune:
        # the autotuner times its candidates on Vg points of the Vd = 0.05 V
        # sweep started as the sweep starts them (untimed, before the warm-up
        # steps; tune_sweep)
        opts, opts_start, tuning = tune_sweep(devs[0], specs[0].sweep, opts, eq_opts,
                                              predictor)
    for _ in range(warmup):
        trigate_step(devs, specs, opts, eq_opts, predictor=predictor, opts_start=opts_start)
    _sync()
    lat, rec = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        rec = []
        trigate_step(devs, specs, opts, eq_opts, lat, rec, predictor, opts_start)
    _sync()
    dt = time.perf_counter() - t0
    npts = sum(len(s.sweep) for s in specs)
    out = dict(seconds=dt, points_per_step=npts, latencies=np.array(lat), record=rec,
               devs=devs, specs=specs, eq_iterations=[d.last_eq_iterations for d in devs],
               opts=opts, tuning=tuning, opts_start=opts_start)
    if opts.forcing or opts.linsolve != "bicgstab" or predictor:
        # the exact-Newton counts of the plain sweep (FP64 solves to rtol
        # every step, each point started from the previous one as the
        # reference starts it): the headline's counts are held to them +-1
        # per point, and its currents to theirs
        exact = []
        trigate_step(devs, specs, NewtonOptions(precond=precond), eq_opts, None, exact, False)
        out["newton_exact"] = [x[1] for x in exact]
        out["drain_exact"] = [x[3] for x in exact]
    if probe is not None and probe >= 0:
        # in-solver duration of one kernel: one more (untimed) step with HIP
        # events around every launch of it
        n, us = probe_step(devs, specs, opts, eq_opts, probe, None, predictor, opts_start)
        out["probe"] = dict(kernel=probe, launches=n, avg_us=us)
        # the same, one step per cut: how the average depends on where the
        # early exits are cut off
        out["probe_cuts"] = [dict(cut_us=c, launches=m, avg_us=round(a, 3) if a else None)
                             for c in probe_cuts
                             for m, a in [probe_step(devs, specs, opts, eq_opts, probe, c,
                                                     predictor, opts_start)]]
    return out


def probe_step(devs, specs, opts, eq_opts, probe, cut_us, predictor=True, opts_start=None):
    """One untimed step with HIP start/stop events on every launch of
    BiCGStab kernel `probe`; returns (launches, average us) over the
    launches longer than cut_us (default: probe_min_us)."""
    import ctypes
    L = _lib()
    for d in devs:
        L.check(L.lib().tcad_probe_kernel(d._gpu().handle, probe), "tcad_probe_kernel")
    trigate_step(devs, specs, opts, eq_opts, predictor=predictor, opts_start=opts_start)
    tot_n, tot_us = 0, 0.0
    for d in devs:
        n = ctypes.c_int64()
        us = ctypes.c_double()
        cut = probe_min_us(probe, tuple(specs[0].shape)) if cut_us is None else cut_us
        L.check(L.lib().tcad_probe_read(d._gpu().handle, cut, ctypes.byref(n), ctypes.byref(us)),
                "tcad_probe_read")
        L.check(L.lib().tcad_probe_kernel(d._gpu().handle, -1), "tcad_probe_kernel")
        tot_n += n.value
        tot_us += n.value * us.value
    return tot_n, (tot_us / tot_n if tot_n else None)


def reduced_trigate_parity(opts, predictor=True, opts_start=None):
    """The reduced 24x18x14 tri-gate at both drain biases against the
    reference's own sweeps (committed fixtures), with the headline's solver
    options and sweep predictor: Newton counts and every contact current
    per point, strict 1e-6 of the largest current."""
    from golden.fixtures import load
    from tcadhip import sweep
    from tcadhip.device import NewtonOptions, device_from_spec
    out = {}
    worst, pts, strict, newton_ok = 0.0, 0, 0, True
    for name in ("trigate_24x18x14_vd005_sweep.npz", "trigate_24x18x14_vd07_sweep.npz"):
        spec, d = load(name)
        dev = device_from_spec(spec)
        dev.solve_equilibrium(NewtonOptions(max_iter=400))
        its, cur = [], []

        def on_point(k, n_it, kry, dev=dev, spec=spec):
            its.append(n_it)
            I = dev.terminal_currents()
            cur.append([I[c.name] for c in spec.contacts])

        sweep.device_sweep(dev, spec.sweep, opts, predictor=predictor, on_point=on_point,
                           opts_start=opts_start)
        cur = np.array(cur)
        ref = d["currents"]
        for p in range(len(cur)):
            e = float(np.abs(cur[p] - ref[p]).max() / np.abs(ref[p]).max())
            worst = max(worst, e)
            strict += int(e <= 1e-6)
            pts += 1
        newton_ok = newton_ok and bool(np.all(np.abs(np.array(its) - d["newton_iterations"]) <= 1))
        out[name.split("_sweep")[0]] = {"newton_gpu": its,
                                        "newton_reference": d["newton_iterations"].tolist()}
    out.update({"points": pts, "points_within_1e-6_no_floor": strict,
                "max_rel_current_err_vs_reference": worst, "newton_within_1": newton_ok})
    return out


def cpu_trigate_baseline(gpu_dev, spec, cap_s, gpu_newton):
    """The FP64 CPU restatement at full size (NumPy assembly, SciPy
    GMRES(100) + node block-Jacobi at rtol 1e-10, one core), first Vg point
    of the Vd = 0.05 V sweep, warm-started from the GPU's equilibrium state;
    stops after the first Newton iteration past `cap_s`."""
    import warnings
    warnings.simplefilter("ignore")
    from oracle import physics as oph
    from oracle_build import oracle_device
    from tcadhip.device import NewtonOptions
    gpu_dev.solve_equilibrium(NewtonOptions(max_iter=400))
    odev = oracle_device(spec)
    odev.psi, odev.n, odev.p = gpu_dev.psi, gpu_dev.n, gpu_dev.p
    opts = oph.NewtonOptions(linsolve="gmres", linsolve_rtol=1e-10, block_size=3,
                             direct_fallback=False)
    t0 = time.perf_counter()
    err = None
    try:
        odev.solve_bias(spec.sweep[0], opts, deadline=t0 + cap_s)
    except Exception as e:                    # a GMRES that does not converge
        err = f"{type(e).__name__}: {e}"
    dt = time.perf_counter() - t0
    done = bool(odev.last_converged) if err is None and not odev.last_timed_out else False
    its = odev.last_iterations or 0
    if done:
        value = 1.0 / dt
        how = f"converged in {its} Newton iterations, {dt:.1f} s"
    else:
        # not finished under the cap: the rate the iterations done imply,
        # with the GPU's Newton count for the same point
        value = (its / max(gpu_newton, 1)) / dt if its else 0.0
        how = (f"NOT finished within the {cap_s:.0f} s cap: {its} of the {gpu_newton} Newton "
               f"iterations the GPU needs in {dt:.1f} s" + (f" ({err})" if err else "")
               + "; value = the fraction done per second")
    return {"value": round(value, 6), "unit": "bias points/s", "cores": 1, "kind": "port",
            "completed": done, "krylov_iterations": odev.last_krylov_iterations,
            "sample": "FP64 oracle restatement (NumPy assembly + SciPy GMRES(100) with node "
                      "block-Jacobi, rtol 1e-10, the reference's method for large 3D systems) "
                      "on the full 64x48x48 tri-gate, first point (Vd 0.05 V, Vg 0 V) from the "
                      f"GPU's equilibrium state: {how}"}


# ---------------------------------------------------------------------------
# extras
# ---------------------------------------------------------------------------
def bench_diode(steps=3, warmup=1):
    """The 60x40 diode, 0 -> 0.6 V in 13 points, FP64, against the
    reference's own sweep (fixture) and the CPU oracle's time for it."""
    from golden.fixtures import load
    from oracle_build import oracle_sweep
    from tcadhip.device import NewtonOptions, device_from_spec
    spec, d = load("diode_60x40_sweep.npz")
    dev = device_from_spec(spec)
    opts = NewtonOptions()
    names = [c.name for c in spec.contacts]

    def sweep(lat=None):
        dev.solve_equilibrium(opts)
        cur, its = [], []
        for bias in spec.sweep:
            t0 = time.perf_counter()
            dev.solve_bias(bias, opts)
            I = dev.terminal_currents()
            if lat is not None:
                lat.append(time.perf_counter() - t0)
            cur.append([I[n] for n in names])
            its.append(dev.last_iterations)
        return np.array(cur), its
    for _ in range(warmup):
        sweep()
    _sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        cur, its = sweep(lat)
    _sync()
    dt = time.perf_counter() - t0
    ref = d["currents"]
    errs = [float(np.abs(cur[p] - ref[p]).max() / np.abs(ref[p]).max()) for p in range(len(ref))]
    cons = [float(abs(ref[p].sum()) / np.abs(ref[p]).max()) for p in range(len(ref))]
    t1 = time.perf_counter()
    oracle_sweep(spec)
    dt_cpu = time.perf_counter() - t1
    lat_ms = np.array(lat) * 1e3
    return {"points": len(spec.sweep), "points_per_s": round(len(spec.sweep) * steps / dt, 3),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "newton_gpu": its, "newton_reference": d["newton_iterations"].tolist(),
            "points_within_1e-6_no_floor": int(sum(e <= 1e-6 for e in errs)),
            "max_rel_current_err_vs_reference": max(errs),
            "points_under_floor": int(sum(e > 1e-6 and e <= c + 1e-6 for e, c in zip(errs, cons))),
            "reference_conservation_rel_max": max(cons),
            "cpu_oracle_points_per_s": round(len(spec.sweep) / dt_cpu, 3),
            "cpu_sample": f"FP64 oracle (NumPy + SciPy SuperLU, 1 core), equilibrium + 13 points, "
                          f"{dt_cpu:.2f} s"}


MOPTS = ("linsolve", "forcing", "forcing_pow", "forcing_slope", "forcing_max", "forcing_fp32",
         "tile_rows", "graph", "tile_lag")


def bench_mosfet(steps=2, warmup=1, modes=("bicgstab", "bicgstab_mixed", "tuned"), cpu_points=3):
    """The 200x120 MOSFET Id-Vg sweep (21 points): the FP64 path and the
    FP32-matrix path with the reference's warm start, and the autotuned
    FP64 path with the secant predictor (the autotuner timed on the third
    Vg point from its predicted start); the CPU oracle on the first points
    for time and agreement."""
    from oracle_build import oracle_device
    from tcadhip import configs, sweep as sweepmod
    from tcadhip.device import NewtonOptions, device_from_spec
    spec = configs.config2_mosfet()
    out = {}
    ref64 = None
    for mode in modes:
        dev = device_from_spec(spec)
        opts = NewtonOptions(linsolve="bicgstab" if mode == "tuned" else mode)
        predictor = mode == "tuned"
        opts_start = None
        if mode == "tuned":
            opts, opts_start, rep = tune_sweep(dev, spec.sweep, opts, opts, True)

        def sweep(lat=None):
            dev.solve_equilibrium(opts)
            cur, its = [], []
            t_last = [time.perf_counter()]

            def on_point(k, n_it, kry):
                cur.append(dev.terminal_currents()["drain"])
                t = time.perf_counter()
                if lat is not None:
                    lat.append(t - t_last[0])
                t_last[0] = t
                its.append(n_it)

            sweepmod.device_sweep(dev, spec.sweep, opts, predictor=predictor, on_point=on_point,
                                  opts_start=opts_start)
            return np.array(cur), its
        for _ in range(warmup):
            sweep()
        _sync()
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            cur, its = sweep(lat)
        _sync()
        dt = time.perf_counter() - t0
        if ref64 is None:
            ref64 = cur
        lat_ms = np.array(lat) * 1e3
        out[mode] = {"points_per_s": round(len(spec.sweep) * steps / dt, 3),
                     "sweep_start": "secant predictor" if predictor else "plain warm start",
                     "options": {k: getattr(opts, k) for k in MOPTS},
                     "options_unpredicted_points": (None if opts_start is None else
                                                    {k: getattr(opts_start, k) for k in MOPTS}),
                     "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                     "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                     "newton_per_point": its,
                     "max_rel_drain_err_vs_fp64_gpu": float(np.max(np.abs(cur - ref64)
                                                                   / np.abs(ref64)))}
        out[mode]["_drain"] = cur
    odev = oracle_device(spec)
    t0 = time.perf_counter()
    odev.solve_equilibrium()
    ocur = []
    for bias in spec.sweep[:cpu_points]:
        odev.solve_bias(bias)
        ocur.append(odev.terminal_current("drain"))
    dt_cpu = time.perf_counter() - t0
    ocur = np.array(ocur)
    for mode in modes:
        e = float(np.max(np.abs(out[mode]["_drain"][:cpu_points] - ocur) / np.abs(ocur)))
        out[mode]["max_rel_drain_err_vs_oracle"] = e
        out[mode]["meets_1e-6_no_floor"] = bool(
            e <= 1e-6 and out[mode]["max_rel_drain_err_vs_fp64_gpu"] <= 1e-6)
        del out[mode]["_drain"]
    out["cpu_oracle_points_per_s"] = round(cpu_points / dt_cpu, 4)
    out["cpu_sample"] = (f"FP64 oracle (NumPy + SciPy SuperLU, 1 core), equilibrium + first "
                         f"{cpu_points} of 21 points, {dt_cpu:.1f} s")
    return out


def bench_ac(freqs=tuple(np.logspace(6, 12, 7)), cpu_freqs=2):
    """Small-signal Y(f) of the 200x120 MOSFET at Vg = 1.0 V, Vd = 0.05 V:
    4 ports x 7 frequencies on the GPU, f_T; the FP64 oracle (SciPy SuperLU
    per frequency) on the first frequencies from the same operating point."""
    from oracle import ac as oac
    from oracle_build import oracle_device
    from tcadhip import ac, configs
    from tcadhip.device import NewtonOptions, device_from_spec
    spec = configs.config2_mosfet()
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions())
    for bias in spec.sweep:
        dev.solve_bias(bias, NewtonOptions())
    ac.y_parameters(dev, [freqs[0]])                   # warm-up
    _sync()
    t0 = time.perf_counter()
    res = ac.y_parameters(dev, list(freqs))
    _sync()
    dt = time.perf_counter() - t0
    odev = oracle_device(spec)
    odev.psi, odev.n, odev.p = dev.psi, dev.n, dev.p
    for nm, bc in dev.bcs.items():
        obc = odev.bcs[nm]
        if hasattr(bc, "Vg"):
            obc.Vg = bc.Vg
        else:
            obc.V = bc.V
    odev._update_surface_mobility(odev.psi)
    t1 = time.perf_counter()
    ores = oac.y_parameters(odev, list(freqs[:cpu_freqs]))
    dt_cpu = time.perf_counter() - t1
    err = max(float(np.abs(res.Y[f] - ores.Y[f]).max() / np.abs(ores.Y[f]).max())
              for f in range(cpu_freqs))
    return {"operating_point": dict(spec.sweep[-1]), "ports": list(res.port_names),
            "frequencies": len(freqs), "seconds": round(dt, 3),
            "frequencies_per_s": round(len(freqs) / dt, 3),
            "krylov_iterations": res.iterations,
            "f_T_Hz": ac.cutoff_frequency(res, "gate", "drain"),
            "cpu_oracle_frequencies_per_s": round(cpu_freqs / dt_cpu, 3),
            "cpu_sample": f"FP64 oracle, SciPy SuperLU, first {cpu_freqs} frequencies, "
                          f"{dt_cpu:.1f} s",
            "max_err_vs_oracle_rel_to_max_Y": err}


def _cpu_diode(idx):
    """One batch diode's continuation sweep on the FP64 CPU oracle (NumPy +
    SciPy SuperLU), in a worker process of bench_batch's CPU baseline;
    returns (final currents in contact order, seconds)."""
    import warnings
    warnings.simplefilter("ignore")
    from current_bar import oracle_sweep
    from tcadhip import configs
    spec = configs.batch_diodes()[idx]
    t0 = time.perf_counter()
    _, cur = oracle_sweep(spec)
    return np.asarray(cur[-1]).tolist(), time.perf_counter() - t0


def bench_transient(t_end=1e-9, dt0=1e-12, cpu_t_end=3e-12):
    """Transient: the 200x120 MOSFET at Vg = 1.0 V, Vd = 0.05 V, its drain
    stepped to 0.5 V at t = 0, backward Euler with the reference's adaptive
    dt (transient2d.py:186-306) to t_end: accepted steps per second.  The
    FP64 oracle (NumPy + SciPy SuperLU, the reference's direct solve) steps
    the same device from the same state to cpu_t_end; the GPU's run to that
    time gives the drain current and Newton counts compared, the counts
    with the transient tests' bar: within +-1 of the oracle with the
    equilibrated direct solve (the exact step the GPU's equilibrated Krylov
    solve converges to) and never above the reference-method oracle's + 1."""
    from oracle import transient as otr
    from oracle_build import oracle_device
    from tcadhip import configs
    from tcadhip import transient as tr
    from tcadhip.device import NewtonOptions, device_from_spec
    spec = configs.config2_mosfet()
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions())
    for bias in spec.sweep:
        dev.solve_bias(bias, NewtonOptions())
    psi0, n0, p0 = dev.psi, dev.n, dev.p
    vd = dev.bcs["drain"].V

    def reset():
        dev.bcs["drain"].V = vd
        dev.set_state(psi0, n0, p0)

    # the comparison run, which also warms the path up
    short = tr.solve_transient(dev, {"drain": tr.StepWaveform(vd, 0.5)}, cpu_t_end, dt0)
    reset()
    _sync()
    t0 = time.perf_counter()
    r = tr.solve_transient(dev, {"drain": tr.StepWaveform(vd, 0.5)}, t_end, dt0,
                           keep_history=False)
    _sync()
    dt = time.perf_counter() - t0
    steps = len(r.iterations)
    def oracle_at_start():
        odev = oracle_device(spec)
        odev.psi, odev.n, odev.p = psi0, n0, p0
        for nm, bc in dev.bcs.items():
            obc = odev.bcs[nm]
            if hasattr(bc, "Vg"):
                obc.Vg = bc.Vg
            else:
                obc.V = vd if nm == "drain" else bc.V
        odev._update_surface_mobility(odev.psi)
        return odev

    odev = oracle_at_start()
    t1 = time.perf_counter()
    ores = otr.solve_transient(odev, {"drain": otr.StepWaveform(vd, 0.5)}, cpu_t_end, dt0)
    dt_cpu = time.perf_counter() - t1
    eres = otr.solve_transient(oracle_at_start(), {"drain": otr.StepWaveform(vd, 0.5)},
                               cpu_t_end, dt0, equilibrated=True)
    ng, nr, ne = (list(map(int, x.iterations)) for x in (short, ores, eres))
    g_i = np.asarray(short.terminal_current["drain"])
    o_i = np.asarray(ores.terminal_current["drain"])
    k = min(len(g_i), len(o_i))
    scale = float(np.abs(o_i[:k]).max())
    return {"device": "mosfet_200x120 (Vg 1.0 V), drain 0.05 -> 0.5 V step at t = 0",
            "scheme": "backward Euler, adaptive dt", "t_end_s": t_end, "dt0_s": dt0,
            "steps": steps, "newton_total": int(np.sum(r.iterations)),
            "seconds": round(dt, 3), "steps_per_s": round(steps / dt, 3),
            "compared_to_s": cpu_t_end,
            "times_equal_to_oracle": bool(len(short.times) == len(ores.times) and
                                          np.allclose(short.times, ores.times, rtol=1e-12)),
            "newton_gpu": ng, "newton_oracle": nr, "newton_oracle_equilibrated": ne,
            "newton_within_bar": bool(len(ng) == len(ne) == len(nr) and
                                      all(abs(a - e) <= 1 and a <= r + 1
                                          for a, r, e in zip(ng, nr, ne))),
            "max_rel_drain_err_vs_oracle": float(np.abs(g_i[:k] - o_i[:k]).max() / scale),
            "cpu_oracle_steps_per_s": round(len(ores.iterations) / dt_cpu, 4),
            "cpu_sample": f"FP64 oracle (NumPy + SciPy SuperLU, 1 core), the first "
                          f"{len(ores.iterations)} steps to {cpu_t_end:g} s, {dt_cpu:.1f} s"}


def bench_batch(lanes=2, sample=8, modes=None):
    """256 seeded diodes (40x30..160x100 nodes, -2.0..0.8 V, 0.1 V
    continuation steps), batched, both precisions; the plain FP64 path (one
    device at a time) on an evenly spaced sample.  Unit of work: one device
    converged at its target bias.  Two lanes (half batches on two streams,
    whose launches overlap): 2-4% above one lane in every mode measured,
    three lanes no better than two in the plain modes."""
    from tcadhip import configs, sweep
    from tcadhip.device import NewtonOptions, setup_from_spec
    specs = configs.batch_diodes()
    res = {}
    # plain batch (FP64, FP32-stored matrix), the secant predictor between
    # continuation steps, and the autotuner's choice with the predictor
    modes = modes or {"bicgstab_mixed": NewtonOptions(linsolve="bicgstab_mixed"),
                      "bicgstab": NewtonOptions(),
                      "predictor": (NewtonOptions(), dict(predictor=True)), "tuned": None}
    tuning = None
    if "tuned" in modes:
        # the autotuner on the largest diode (a point of its sweep started as
        # the predicted sweep starts it), its choice applied to every device
        from tcadhip import tune as tuner
        from tcadhip.device import device_from_spec
        big = max(range(len(specs)), key=lambda i: specs[i].N)
        dev = device_from_spec(specs[big])
        k = tuning_start(dev, specs[big].sweep, NewtonOptions(), NewtonOptions(), True)
        # a round solves every open device with one option set, predicted or
        # not: forcing terms (accepted with the predicted start's slack) then
        # apply only to the devices whose start was predicted
        topts, tuning = tuner.tune(dev, specs[big].sweep[k], NewtonOptions(), say=_say,
                                   newton_slack=PREDICTED_SLACK, start="predicted")
        topts.forcing_predicted = topts.forcing
        tuning = dict(tuning, device=big, nodes=int(specs[big].N), point=k)
        modes["tuned"] = (topts, dict(predictor=True))
        del dev
    for mode, mopts in modes.items():
        mopts, kw = mopts if isinstance(mopts, tuple) else (mopts, {})
        setups = [setup_from_spec(s) for s in specs]
        _sync()
        t0 = time.perf_counter()
        r, bss = sweep.run_lanes(setups, [s.sweep for s in specs], mopts, lanes=lanes, **kw)
        _sync()
        dt = time.perf_counter() - t0
        res[mode] = dict(seconds=dt, results=r, rounds=sum(b.rounds for b in bss),
                         rebuilds=sum(b.rebuilds for b in bss))
        _say(f"batch {mode}: {len(specs) / dt:.3f} devices/s")
    idx = list(range(0, len(specs), len(specs) // sample))[:sample]
    _sync()
    t0 = time.perf_counter()
    alone = {}
    for d in idx:
        alone[d] = sweep.BatchedSweep([setup_from_spec(specs[d])], [specs[d].sweep],
                                      NewtonOptions()).run()[0]
    _sync()
    plain = len(idx) / (time.perf_counter() - t0)

    def final(r, names):
        return np.array([r.currents[-1][nm] for nm in names])

    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1))
    cidx = list(range(0, len(specs), max(1, len(specs) // workers)))[:workers]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(workers) as pool:
        cres = pool.map(_cpu_diode, cidx, chunksize=1)
    dt_cpu = time.perf_counter() - t0
    out = {"devices": len(specs), "nodes": int(sum(s.N for s in specs)),
           "continuation_solves": int(sum(len(s.sweep) for s in specs)), "lanes": lanes,
           "plain_fp64_devices_per_s": round(plain, 4),
           "plain_sample": f"{len(idx)} evenly spaced devices solved one at a time",
           "cpu_oracle": {"devices_per_s": round(len(cidx) / dt_cpu, 4), "workers": workers,
                          "mean_seconds_per_device": round(float(np.mean([t for _, t in cres])),
                                                           2),
                          "sample": f"{len(cidx)} evenly spaced devices, FP64 oracle (NumPy + "
                                    f"SciPy SuperLU), one per worker process"}}
    for mode, r in res.items():
        strict = floor_pts = 0
        worst = 0.0
        for d, (ccur, _) in zip(cidx, cres):
            names = [c.name for c in specs[d].contacts]
            got = final(r["results"][d], names)
            ref = np.array(ccur)
            e = float(np.abs(got - ref).max())
            big = float(np.abs(ref).max())
            worst = max(worst, e / big)
            if e <= 1e-6 * big:
                strict += 1
            elif e <= 1e-6 * big + abs(ref.sum()) + abs(got.sum()):
                floor_pts += 1
        e64 = 0.0
        beyond64 = 0
        for d in idx:
            names = [c.name for c in specs[d].contacts]
            got, ref = final(r["results"][d], names), final(alone[d], names)
            e = float(np.abs(got - ref).max())
            e64 = max(e64, e / float(np.abs(ref).max()))
            # reverse-biased diodes: the current is a cancellation below
            # both solves' conservation error (their contact-current sums)
            beyond64 += int(e > 1e-6 * float(np.abs(ref).max()) + abs(ref.sum()) + abs(got.sum()))
        over = None
        if mode != "bicgstab" and "bicgstab" in res:
            # Newton counts against the exact (untuned FP64, plain warm start)
            # batch, device by device: points that needed more than one
            # iteration more (the predictor's starts save iterations)
            over = sum(int(a - b > 1) for x, y in zip(r["results"], res["bicgstab"]["results"])
                       for a, b in zip(x.newton, y.newton))
        out[mode] = {"devices_per_s": round(len(specs) / r["seconds"], 4),
                     "seconds": round(r["seconds"], 3),
                     "newton_total": int(sum(sum(x.newton) for x in r["results"])),
                     "newton_points_above_exact_plus_1": over,
                     "failed": sum(1 for x in r["results"] if x.failed),
                     "rounds": r["rounds"], "compactions": r["rebuilds"],
                     "vs_plain_fp64": round(len(specs) / r["seconds"] / plain, 2),
                     "max_rel_current_err_vs_fp64_alone": e64,
                     "fp64_alone_beyond_conservation_floor": beyond64,
                     "vs_cpu_oracle_devices": len(cidx),
                     "meets_1e-6_no_floor": strict,
                     "under_conservation_floor_only": floor_pts,
                     "max_rel_current_err_vs_oracle": worst}
        mo = modes[mode][0] if isinstance(modes[mode], tuple) else modes[mode]
        if isinstance(modes[mode], tuple):
            out[mode]["sweep"] = modes[mode][1]
        if mode == "tuned":
            out[mode]["options"] = {k: getattr(mo, k) for k in
                                    ("forcing", "forcing_pow", "forcing_slope", "forcing_max",
                                     "forcing_fp32", "tile_rows", "graph", "tile_lag",
                                     "forcing_predicted")}
            out[mode]["autotuner"] = tuning
    return out


def bench_serve(count=16, lanes=2, check=3, budget_gib=32):
    """Serving: seeded Poisson arrivals of mixed requests (2D diode and
    MOSFET sweeps, 3D tri-gate sweeps, 2k-200k nodes, 5-40 points each),
    continuous batching in an HBM-budgeted pool with `lanes` lanes.
    First the saturation throughput (every request present at t = 0), then
    the same stream at 30% and 60% of that rate, where the server keeps up;
    `check` requests re-solved alone must give the same bits.  The server
    starts every continued point from the secant prediction and applies
    quadratic forcing terms (0.1, 1e-3) to the devices whose start was
    predicted (84 -> 102-106 points/s at saturation against the plain
    options, the slopes 0.03-0.1 and linear forcing within 4% of each
    other)."""
    from tcadhip import serve
    from tcadhip.device import NewtonOptions
    sopts = NewtonOptions(forcing=True, forcing_pow=2, forcing_slope=0.1, forcing_max=1e-3,
                          forcing_predicted=True)
    cap = serve.nodes_for_budget(budget_gib << 30)
    base = serve.request_stream(count, rate=1.0, seed=20261203)
    pts_total = sum(len(r.spec.sweep) for r in base)
    out = {"requests": count, "lanes": lanes, "pool_nodes": cap,
           "kinds": {k: sum(r.kind == k for r in base) for k in ("diode", "mosfet", "trigate")},
           "nodes_range": [min(r.spec.N for r in base), max(r.spec.N for r in base)],
           "bias_points": pts_total}

    def run(rate):
        reqs = serve.request_stream(count, rate=rate, seed=20261203) if rate else \
            [serve.Request(r.rid, r.spec, 0.0, r.kind) for r in base]
        S = serve.Server(cap, opts=sopts, lanes=lanes, predictor=True)
        _sync()
        t0 = time.perf_counter()
        served = S.run(reqs, realtime=True)
        _sync()
        dt = time.perf_counter() - t0
        lat = np.array([o.latency for o in served])
        per_pt = np.array([o.latency / max(1, len(o.result.currents)) for o in served])
        return served, reqs, {
            "arrival_rate_per_s": rate, "wall_s": round(dt, 3),
            "points_per_s": round(pts_total / dt, 3),
            "request_latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "request_latency_p99_s": round(float(np.percentile(lat, 99)), 3),
            "latency_per_point_p50_ms": round(1e3 * float(np.percentile(per_pt, 50)), 2),
            "latency_per_point_p99_ms": round(1e3 * float(np.percentile(per_pt, 99)), 2),
            "failed": sum(1 for o in served if o.result.failed),
            "rounds": S.rounds, "batch_rebuilds": S.rebuilds, "max_resident": S.max_resident}
    served, reqs, sat = run(None)
    out["saturation"] = sat
    req_rate = count / sat["wall_s"]
    out["saturation_requests_per_s"] = round(req_rate, 4)
    for frac in (0.3, 0.6):
        _, _, r = run(round(frac * req_rate, 4))
        r["fraction_of_saturation"] = frac
        out[f"load_{int(frac * 100)}pct"] = r
    same = True
    for o, r in list(zip(served, reqs))[:check]:
        a = serve.serve_alone(r.spec, sopts, predictor=True)
        same = same and a.newton == o.result.newton and all(
            x[k] == y[k] for x, y in zip(a.currents, o.result.currents) for k in x)
    # the plain FP64 path on the same requests: one at a time, each point
    # started from the previous converged state (no predictor)
    _sync()
    t1 = time.perf_counter()
    alone_pts = 0
    for r in reqs[:check]:
        alone_pts += len(serve.serve_alone(r.spec, NewtonOptions()).currents)
    _sync()
    out["sweep_start"] = "secant predictor"
    out["options"] = {k: getattr(sopts, k) for k in
                      ("forcing", "forcing_pow", "forcing_slope", "forcing_max",
                       "forcing_predicted")}
    out["alone_equals_mix_on_first"] = check
    out["alone_equals_mix"] = bool(same)
    out["plain_fp64_points_per_s_on_those"] = round(alone_pts / (time.perf_counter() - t1), 4)
    return out


def copy_ceiling(mib=1024, reps=10):
    """On-box bandwidth ceiling: a plain device-to-device copy of `mib` MiB,
    bytes read + written per second (GB/s), timed with HIP events."""
    import torch
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    gbps = 2 * n * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return round(gbps, 1)


def measured_traffic(profile, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(REPO, "profiles", profile)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[0] == f"`{kernel}`":
            try:
                return float(cells[6]) * 1e6
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="bicgstab", choices=("bicgstab", "bicgstab_mixed"),
                    help="linear-solve path of the headline")
    ap.add_argument("--points", type=int, default=None,
                    help="first K Vg points of each sweep only (short profiling runs)")
    ap.add_argument("--probe", type=int, default=4,
                    help="BiCGStab kernel timed in-solver for the roofline: 0 SpMV, 1 s, "
                         "2 second SpMV, 3 update, 4 black tiles (the most GPU time; -1: none)")
    ap.add_argument("--forcing", action="store_true",
                    help="Newton forcing terms (inexact early linear solves)")
    ap.add_argument("--forcing-slope", type=float, default=0.01)
    ap.add_argument("--forcing-max", type=float, default=1e-4)
    ap.add_argument("--forcing-pow", type=int, default=1, choices=(1, 2),
                    help="forcing term linear (1) or quadratic (2) in the last Newton update")
    ap.add_argument("--precond", default="tile", choices=("tile", "node"))
    ap.add_argument("--probe-cuts", default="",
                    help="comma-separated early-exit cuts (us): one more probe step per cut, "
                         "reported as probe_cuts")
    ap.add_argument("--tile-rows", default="auto", choices=("auto", "reg", "mem"),
                    help="tile-inverse rows in registers or read at the point of use "
                         "(untuned runs)")
    ap.add_argument("--set", default="",
                    help="NewtonOptions of the headline's predicted points, key=value,... "
                         "(with --no-tune: a fixed configuration, e.g. the autotuner's choice)")
    ap.add_argument("--start-set", default="",
                    help="NewtonOptions of the points without a prediction, key=value,...")
    ap.add_argument("--no-predictor", action="store_true",
                    help="start each Vg point from the previous converged state (the "
                         "reference's warm start) instead of the secant prediction")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (fixed default solver options)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cap", type=float, default=30.0,
                    help="wall-clock cap [s] of the full-size CPU baseline")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--extras", default="diode,mosfet,ac,transient,batch,serve",
                    help="extra workloads (comma-separated; '' for none)")
    ap.add_argument("--extras-only", action="store_true",
                    help="run the --extras workloads without the headline (profiling runs); "
                         "prints one JSON object of them")
    ap.add_argument("--mosfet-modes", default="bicgstab,bicgstab_mixed,tuned",
                    help="the MOSFET extra's legs (comma-separated)")
    ap.add_argument("--batch-variants", default="",
                    help="only the batched diodes, with these option sets (comma-separated of "
                         "bicgstab, forcing, quad, tuned, forcing_graph, graph, queue, queue2, queue4)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (the reference "
                         "kernel for scaling HBM traffic counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every MI355X box here has one, and this "
              f"workload is not sharded; --gpus {args.gpus} is not supported.", file=sys.stderr)
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
    if args.batch_variants:
        from tcadhip.device import NewtonOptions
        allm = {"bicgstab": NewtonOptions(),
                "forcing": NewtonOptions(forcing=True, forcing_slope=0.1, forcing_max=1e-3),
                "forcing_graph": NewtonOptions(forcing=True, forcing_slope=0.1,
                                               forcing_max=1e-3, graph=True),
                "graph": NewtonOptions(graph=True),
                "quad": NewtonOptions(forcing=True, forcing_pow=2, forcing_slope=0.1,
                                      forcing_max=1e-3),
                "tuned": None,
                "predictor": (NewtonOptions(), dict(predictor=True)),
                "node": NewtonOptions(precond="node"),
                "node_predictor": (NewtonOptions(precond="node"), dict(predictor=True)),
                "predictor_graph": (NewtonOptions(graph=True), dict(predictor=True)),
                "queue": (NewtonOptions(), dict(queue=True)),
                "queue2": (NewtonOptions(), dict(queue=True, queue_depth=2)),
                "queue4": (NewtonOptions(), dict(queue=True, queue_depth=4))}
        out = bench_batch(modes={k: allm[k] for k in args.batch_variants.split(",")})
        print(json.dumps(out), flush=True)
        return
    if args.extras_only:
        out = {}
        for result, key in run_extras(args):
            out[key] = result
        print(json.dumps(out), flush=True)
        return

    _say(f"tri-gate 64x48x48, {args.mode}: warmup {args.warmup}, steps {args.steps}")
    r = bench_trigate(args.steps, args.warmup, args.mode, args.points, args.probe,
                      args.forcing, args.precond, tune=not args.no_tune,
                      forcing_slope=args.forcing_slope, forcing_max=args.forcing_max,
                      forcing_pow=args.forcing_pow, tile_rows=args.tile_rows,
                      probe_cuts=[float(c) for c in args.probe_cuts.split(",") if c],
                      predictor=not args.no_predictor,
                      overrides=parse_options(args.set), start_overrides=parse_options(args.start_set))
    npts = r["points_per_step"]
    value = npts * args.steps / r["seconds"]
    _say(f"tri-gate: {value:.3f} points/s")
    N = r["specs"][0].N
    specs0 = r["specs"][0]
    lat_ms = r["latencies"] * 1e3
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "bias points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.mode == "bicgstab" else "f32 matrix / f64 vectors + f64 residual",
        "data": "synthetic",
        "config": {"workload": "trigate_64x48x48_idvg_vd0.05+0.7_vg0-0.8_34pts", "nodes": N,
                   "unknowns": 3 * N, "bias_points_per_step": npts, "path": args.mode,
                   "precond": "2x1x8-node tile block-Jacobi (FP16 rows, FP32 scales)",
                   "sweep_start": ("plain warm start" if args.no_predictor else
                                   "secant predictor from the last two converged points")},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "solver_options": {k: getattr(r["opts"], k) for k in
                           ("linsolve", "linsolve_rtol", "precond", "forcing", "forcing_pow",
                            "forcing_slope", "forcing_max", "forcing_fp32", "tile_rows", "graph",
                            "tile_lag")},
        "solver_options_unpredicted_points": (
            None if r["opts_start"] is None else
            {k: getattr(r["opts_start"], k) for k in ("forcing", "forcing_pow", "forcing_slope",
                                                      "forcing_max", "forcing_fp32", "tile_rows",
                                                      "graph", "tile_lag")}),
        "autotuner": r["tuning"],
        "newton_per_point": [x[1] for x in r["record"]],
        "newton_exact_per_point": r.get("newton_exact"),
        "newton_within_1_of_exact": (None if r.get("newton_exact") is None else
                                     all(abs(a - b) <= 1 for a, b in
                                         zip([x[1] for x in r["record"]], r["newton_exact"]))),
        "max_rel_drain_err_vs_exact_plain": (
            None if r.get("drain_exact") is None else
            float(max(abs(a - b) / abs(b) for a, b in
                      zip([x[3] for x in r["record"]], r["drain_exact"])))),
        "krylov_per_point": [x[2] for x in r["record"]],
        "equilibrium_newton": r["eq_iterations"],
        "drain_current_A": {"vd0.05": [x[3] for x in r["record"] if "vd0.05" in x[0]],
                            "vd0.7": [x[3] for x in r["record"] if "vd0.7" in x[0]]},
    }
    if "probe" in r and r["probe"]["avg_us"]:
        pr = r["probe"]
        kname = PROBE_KERNELS[pr["kernel"]]
        if r["opts"].tile_rows == "reg":
            kname = kname.replace("_tmem", "_treg")
        shape = tuple(specs0.shape)
        nbytes = krylov_bytes(shape, r["opts"].tile_rb)[pr["kernel"]]
        achieved = nbytes / (pr["avg_us"] * 1e-6) / 1e9
        line["roofline"] = {
            "bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": measured_traffic("trigate_64x48x48_hbm_traffic.md", kname),
            "bytes_per_launch": nbytes, "bytes_per_node": round(nbytes / N, 1),
            "avg_launch_us_in_solver": round(pr["avg_us"], 3),
            "launches_timed": pr["launches"],
            "timing": (f"kernel start/end HIP events (hipExtLaunchKernelGGL) on every launch "
                       f"of the kernel in one untimed step; launches under "
                       f"{probe_min_us(pr['kernel'], shape):.2f} us (early exits after "
                       f"convergence: bytes / (2 x peak)) excluded"),
            "copy_ceiling_GBps": copy_ceiling()}
        if r.get("probe_cuts"):
            line["probe_cuts"] = r["probe_cuts"]
    if not args.no_parity:
        _say("reduced tri-gate vs the reference fixtures")
        line["parity_reduced_24x18x14_vs_reference"] = reduced_trigate_parity(
            r["opts"], not args.no_predictor, r["opts_start"])
    line["cpu_baseline"] = None
    if not args.no_cpu:
        gpu_newton = r["record"][0][1] if r["record"] else 5
        _say(f"CPU baseline at full size (cap {args.cpu_cap:.0f} s)")
        line["cpu_baseline"] = cpu_trigate_baseline(r["devs"][0], r["specs"][0], args.cpu_cap,
                                                    gpu_newton)
    del r
    torch.cuda.empty_cache()
    for result, key in run_extras(args):
        line[key] = result
    print(json.dumps(line), flush=True)


def run_extras(args):
    """The --extras workloads in a fixed order: yields (result, JSON key)."""
    import torch
    extras = [e for e in args.extras.split(",") if e]
    mosfet_modes = tuple(m for m in args.mosfet_modes.split(",") if m)
    for name, fn in (("diode", bench_diode),
                     ("mosfet", lambda: bench_mosfet(modes=mosfet_modes)), ("ac", bench_ac),
                     ("transient", bench_transient), ("batch", bench_batch),
                     ("serve", bench_serve)):
        if name in extras:
            _say(f"extra: {name}")
            key = {"diode": "diode_60x40", "mosfet": "mosfet_200x120",
                   "ac": "ac_mosfet_200x120", "transient": "transient_mosfet_200x120",
                   "batch": "batched_diodes_256", "serve": "serving"}[name]
            yield fn(), key
            torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
