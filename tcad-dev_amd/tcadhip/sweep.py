"""Batched bias sweeps: many devices, each with its own list of bias points,
advanced together through shared launches.

Each device follows the reference's sweep semantics -- equilibrium, then
every requested point warm-started from the previous converged state with
the terminal currents taken after each (pytcad/pytcad/mosfet.py:90-103
id_vg_sweep, finfet3d.py:246-263 id_vg_sweep_3d) -- with the step backoff
of pytcad/pytcad/continuation.py:61-120 (adaptive_bias_sweep): when Newton
fails, the device's state is restored to its last CONFIRMED-converged
state and the step toward the next point is halved; a device whose step
shrinks below `min_step` is reported failed, never returned as converged.
A step that succeeds grows by `growth` up to the full distance.

The devices share every kernel launch: one round issues one masked
tcad_solve_bias over all devices that still have points to reach; the
others are masked out and keep their state untouched.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import GPUBatch, NewtonOptions, _torch


@dataclass
class SweepResult:
    """Per-device outcome of a batched sweep."""
    points: list = field(default_factory=list)      # requested bias dicts reached
    currents: list = field(default_factory=list)    # {contact: I} after each point
    newton: list = field(default_factory=list)      # Newton iterations per point (last attempt)
    krylov: list = field(default_factory=list)      # Krylov iterations per point (all attempts)
    steps: int = 0                                  # accepted solves (intermediate included)
    backoffs: int = 0                               # failed attempts
    failed: str = ""                                # reason when the sweep stopped early
    t_done: float = 0.0                             # wall time when the last point finished


def _bias_vector(setup, bias, names):
    return np.array([float(bias.get(nm, _bc_value(setup, nm))) for nm in names])


def _bc_value(setup, name):
    bc = setup.bcs[name]
    return bc.V if hasattr(bc, "V") else bc.Vg


def _apply_bias(setup, names, vec):
    for nm, v in zip(names, vec):
        bc = setup.bcs[nm]
        if hasattr(bc, "V"):
            bc.V = float(v)
        else:
            bc.Vg = float(v)


class BatchedSweep:
    """Run the bias lists `sweeps[d]` (lists of {terminal: V} dicts) of the
    devices `setups[d]` (DeviceSetup, boundary conditions attached) in one
    batch.  `batch` may be passed to work on an existing GPUBatch (e.g. a
    resident pool's); its devices must be `setups` in order."""

    def __init__(self, setups, sweeps, opts=None, batch=None, max_step=None, min_step=None,
                 growth=1.5, shrink=0.5, max_attempts=64):
        self.setups = list(setups)
        self.sweeps = [list(s) for s in sweeps]
        if len(self.sweeps) != len(self.setups):
            raise ValueError("one bias list per device")
        self.opts = opts or NewtonOptions()
        self.B = batch if batch is not None else GPUBatch(self.setups)
        self.max_step = max_step
        self.min_step = min_step
        self.growth, self.shrink = growth, shrink
        self.max_attempts = max_attempts
        nd = len(self.setups)
        # terminals whose bias a sweep moves, per device (contacts and gates)
        self.names = [[nm for nm in s.bcs] for s in self.setups]
        self.results = [SweepResult() for _ in range(nd)]
        self.rounds = 0

    def run(self, equilibrium=True):
        torch = _torch()
        B, opts = self.B, self.opts
        nd = len(self.setups)
        t0 = time.perf_counter()
        # the sweep starts from the terminals' present values (equilibrium
        # is solved at them, as Device2D.solve_equilibrium is)
        cur = [_bias_vector(s, {}, nm) for s, nm in zip(self.setups, self.names)]
        if equilibrium:
            its, conv, status = B.solve_equilibrium(opts)
            for d in range(nd):
                if status[d] != _lib.NEWTON_OK or not conv[d]:
                    self.results[d].failed = "equilibrium did not converge"
        nxt = [0] * nd                        # index of the next requested point
        step = [None] * nd                    # current step length (max-norm, V)
        attempts = [0] * nd
        kry_acc = [0] * nd
        while True:
            todo = [d for d in range(nd)
                    if not self.results[d].failed and nxt[d] < len(self.sweeps[d])]
            if not todo:
                break
            self.rounds += 1
            trial = {}
            for d in todo:
                target = _bias_vector(self.setups[d], self.sweeps[d][nxt[d]], self.names[d])
                dist = float(np.max(np.abs(target - cur[d]))) if target.size else 0.0
                lim = dist
                if self.max_step is not None:
                    lim = min(lim, self.max_step)
                if step[d] is not None:
                    lim = min(lim, step[d])
                frac = 1.0 if dist <= lim or dist == 0.0 else lim / dist
                trial[d] = (cur[d] + frac * (target - cur[d]), frac >= 1.0, lim)
                _apply_bias(self.setups[d], self.names[d], trial[d][0])
            mask = np.zeros(nd, dtype=np.int32)
            mask[todo] = 1
            snap = (B.psi.clone(), B.n.clone(), B.p.clone())
            its, conv, kry, status = B.solve_bias(opts, active=mask)
            reached = []
            for d in todo:
                R = self.results[d]
                kry_acc[d] += int(kry[d])
                vec, final, lim = trial[d]
                if conv[d] and status[d] == _lib.NEWTON_OK:
                    cur[d] = vec
                    R.steps += 1
                    if final:
                        reached.append(d)
                        R.points.append(self.sweeps[d][nxt[d]])
                        R.newton.append(int(its[d]))
                        R.krylov.append(kry_acc[d])
                        kry_acc[d] = 0
                        nxt[d] += 1
                    if step[d] is not None:
                        step[d] = step[d] * self.growth
                        if self.max_step is not None:
                            step[d] = min(step[d], self.max_step)
                    attempts[d] = 0
                else:
                    sl = B.device_slice(d)
                    B.psi[sl].copy_(snap[0][sl])
                    B.n[sl].copy_(snap[1][sl])
                    B.p[sl].copy_(snap[2][sl])
                    _apply_bias(self.setups[d], self.names[d], cur[d])
                    R.backoffs += 1
                    attempts[d] += 1
                    step[d] = lim * self.shrink
                    floor = self.min_step if self.min_step is not None else 1e-4
                    if step[d] < floor or attempts[d] > self.max_attempts:
                        code = int(status[d])
                        why = {_lib.NEWTON_LINEAR_FAIL: "linear solve failed",
                               _lib.NEWTON_NONFINITE: "non-finite Newton step"}.get(
                                   code, "Newton did not converge")
                        R.failed = (f"{why}; step shrank below {floor:g} V before "
                                    f"{self.sweeps[d][nxt[d]]}")
            if reached:
                I = B.terminal_currents()
                torch.cuda.synchronize()
                now = time.perf_counter() - t0
                c = 0
                for d in range(nd):
                    names = [nm for dd, nm in B.contacts if dd == d]
                    if d in reached:
                        self.results[d].currents.append(dict(zip(names, I[c:c + len(names)])))
                        self.results[d].t_done = now
                    c += len(names)
        return self.results
