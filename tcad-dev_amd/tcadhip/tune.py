"""Autotuner: picks a device's solver configuration by timing it.

The knobs are the ones whose best value depends on the device and the GPU
rather than on the physics, and that leave the converged result the same
(every candidate refines the converging Newton step to the FP64 rounding
floor, so the terminal currents agree to rounding):

  * Newton forcing terms (TCAD_LIN_FORCING: the linear solves before the
    converging one stop at eta = clamp(c * err^q, rtol, eta_max) instead
    of rtol, err the last Newton update measure), their slope c, power q
    and cap eta_max -- fewer Krylov iterations for, at most, some more
    Newton steps;
  * whether the loose forcing solves run on the FP32-stored matrix (the
    tight and converging ones stay on the FP64 matrix);
  * where the tile-inverse rows live in the BiCGStab kernels: loaded into
    registers before the reductions (fewer waves per SIMD) or read at the
    point of use (full occupancy);
  * whether the BiCGStab iterations go out as replays of a captured HIP
    graph, and whether a Newton step keeps the tile inverses of the step
    before when that step's update was small (tried last, each on the best
    configuration so far).

A candidate is measured by solving one bias point from the same saved
state (psi, n, p restored on the device before each trial), twice, the
faster taken.  Candidates that do not converge, whose currents differ from
the untuned configuration's by more than 1e-9 of the largest current, or
that need more Newton iterations than it (`newton_slack`, 0 by default)
are rejected.  The Newton counts must stay within one of the reference's
at every point of a sweep, and the tuning point is one point: no slack
there keeps a margin elsewhere.  Measured on the 64x48x48 tri-gate sweep
(34 points; exact Newton 143 iterations, 23,922 Krylov, 23.3 points/s):
linear forcing c = 0.1, eta_max = 1e-3 ran 35.0 points/s with 179 Newton
iterations, two points at exact + 2; quadratic forcing c = 0.1,
eta_max = 1e-3 34.3 points/s with 159, every point within exact + 1.
Choices are cached per (dimension, mesh shape, linear-solve path,
preconditioner) for the life of the process.
"""
import time
from dataclasses import replace

import numpy as np

from .device import NewtonOptions, _torch

CANDIDATES = (
    dict(forcing=False, tile_rows="auto"),
    dict(forcing=False, tile_rows="reg"),
    dict(forcing=False, tile_rows="mem"),
    dict(forcing=True, forcing_pow=2, forcing_slope=0.1, forcing_max=1e-3, tile_rows="auto"),
    dict(forcing=True, forcing_pow=2, forcing_slope=0.03, forcing_max=1e-3, tile_rows="auto"),
    dict(forcing=True, forcing_pow=2, forcing_slope=1.0, forcing_max=1e-3, tile_rows="auto"),
    dict(forcing=True, forcing_pow=2, forcing_slope=0.1, forcing_max=1e-3, tile_rows="reg"),
    dict(forcing=True, forcing_pow=1, forcing_slope=0.01, forcing_max=1e-4, tile_rows="auto"),
    dict(forcing=True, forcing_pow=1, forcing_slope=0.01, forcing_max=1e-4, tile_rows="reg"),
    dict(forcing=True, forcing_pow=1, forcing_slope=0.1, forcing_max=1e-3, tile_rows="auto"),
    dict(forcing=True, forcing_pow=2, forcing_slope=0.1, forcing_max=1e-3, forcing_fp32=True,
         tile_rows="auto"),
)

# tried on the best configuration of the first stage, in turn
SECOND_STAGE = (dict(graph=True), dict(tile_lag=1e-2), dict(tile_lag=1e-1))

_KNOBS = ("forcing", "forcing_pow", "forcing_slope", "forcing_max", "forcing_fp32", "tile_rows",
          "graph", "tile_lag")
_cache = {}


def _key(dev, opts):
    return (dev.dim, tuple(dev.shape), opts.linsolve, opts.precond)


def tune(dev, bias, opts=None, candidates=CANDIDATES, repeats=2, say=None, newton_slack=0,
         start="warm"):
    """Time every candidate on one bias point of `dev` (which must hold the
    start every trial restarts from: a converged state, or the secant
    prediction from two, named by `start`, "warm" or "predicted") and
    return (best NewtonOptions, report).  A candidate is accepted when its
    currents agree with the untuned solve's and it needs at most
    newton_slack Newton iterations more.  The device's state and bias are
    left as they were.  Results are cached per device shape, solver and
    start."""
    torch = _torch()
    base = opts or NewtonOptions()
    key = _key(dev, base) + (start,)
    if key in _cache:
        best, report = _cache[key]
        return replace(base, **best), report
    B = dev._gpu()
    saved = [t.clone() for t in (B.psi, B.n, B.p)]
    saved_bias = {nm: (bc.V if hasattr(bc, "V") else bc.Vg) for nm, bc in dev.bcs.items()}
    rows = []
    ref_I = ref_its = None
    try:
        for cand in list(candidates) + list(SECOND_STAGE):
            if not isinstance(cand, dict) or "forcing" not in cand:
                # second stage: the winner so far with one more knob set (its
                # iterations issued as replays of a captured HIP graph; its
                # tile inverses kept over small Newton updates)
                good = [r for r in rows if r["accepted"]]
                if not good:
                    break
                w = min(good, key=lambda r: r["seconds"])
                extra = cand
                cand = {k: w[k] for k in _KNOBS if k in w}
                cand.update(extra)
            o = replace(base, **cand)
            best_t, res = None, None
            for _ in range(repeats):
                for t, s in zip((B.psi, B.n, B.p), saved):
                    t.copy_(s)
                dev._set_voltages(saved_bias)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                dev.solve_bias(bias, o)
                I = dev.terminal_currents()
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                if best_t is None or dt < best_t:
                    best_t = dt
                res = (dev.last_converged, dev.last_iterations, dev.last_krylov_iterations,
                       np.array(list(I.values())))
            conv, its, kry, cur = res
            if ref_I is None:
                ref_I, ref_its = cur, its
            ok = (bool(conv) and its <= ref_its + newton_slack
                  and bool(np.all(np.abs(cur - ref_I) <= 1e-9 * np.abs(ref_I).max() + 1e-300)))
            rows.append(dict(cand, seconds=round(best_t, 5), newton=int(its), krylov=int(kry),
                             accepted=ok))
            if say:
                say(f"tune {cand}: {best_t * 1e3:.2f} ms, Newton {its}, Krylov {kry}"
                    + ("" if ok else " (rejected)"))
    finally:
        for t, s in zip((B.psi, B.n, B.p), saved):
            t.copy_(s)
        dev._set_voltages(saved_bias)
        torch.cuda.synchronize()
    good = [r for r in rows if r["accepted"]]
    win = min(good, key=lambda r: r["seconds"])
    best = {k: win[k] for k in _KNOBS if k in win}
    report = {"candidates": rows, "chosen": best, "point": dict(bias)}
    _cache[key] = (best, report)
    return replace(base, **best), report


def clear_cache():
    _cache.clear()
