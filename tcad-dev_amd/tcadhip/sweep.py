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

With `predictor=True` a device's Newton start is not its last converged
state itself but the secant prediction from its last two converged points
(tcad_predict_state: psi linearly, n and p geometrically, within the
Newton update's density clip), when the next bias step is parallel to the
last one -- the tangent predictor of continuation.py:368-404
(arc_length_sweep), applied to the warm-started voltage sweep.  The
converged states are the same (the Newton test and the final-step
refinement decide them); a point typically needs one Newton iteration
fewer.  A predicted start that fails is handled like any failed step
(state restored, step halved) and the retry starts from the converged
state.

The devices share every kernel launch: one round issues one masked
tcad_solve_bias over all devices that still have points to reach; the
others are masked out and keep their state untouched.  Masked devices
still cost their share of every launch (their workgroups start and
return), so once the open devices hold less than `compact` of the
batch's nodes the sweep continues on a batch of the open devices alone,
built over the same planes at the same offsets (no state moves; every
per-device reduction runs over the device's own blocks, so the results
are the bits the full batch would give).
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import GPUBatch, NewtonOptions, _raise_outcome, _torch


@dataclass
class SweepResult:
    """Per-device outcome of a sweep."""
    points: list = field(default_factory=list)      # requested bias dicts reached
    currents: list = field(default_factory=list)    # {contact: I} after each point
    newton: list = field(default_factory=list)      # Newton iterations per point (last attempt)
    krylov: list = field(default_factory=list)      # Krylov iterations per point (all attempts)
    steps: int = 0                                  # accepted solves (intermediate included)
    backoffs: int = 0                               # failed attempts
    failed: str = ""                                # reason when the sweep stopped early
    t_done: float = 0.0                             # wall time when the last point finished


def secant_ratio(prev, cur, nxt, fmax=2.0):
    """Ratio f of the bias step cur -> nxt to the step prev -> cur (vectors
    of terminal voltages) when the two are parallel and point the same way,
    0 otherwise (no prediction); f is capped at fmax."""
    if prev is None:
        return 0.0
    dp, dn = cur - prev, nxt - cur
    a = float(dp @ dp)
    if a == 0.0:
        return 0.0
    f = float(dn @ dp) / a
    if not 0.0 < f <= fmax or np.abs(dn - f * dp).max() > 1e-9 * np.abs(dn).max():
        return 0.0
    return f


def bias_vector(setup, names):
    """Present terminal voltages of a device, in the order of `names`."""
    return np.array([_bc_value(setup, nm) for nm in names])


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


class Track:
    """Continuation state of one device through its list of bias points."""

    def __init__(self, setup, points, max_step=None, min_step=None, growth=1.5, shrink=0.5,
                 max_attempts=64):
        self.setup = setup
        self.points = list(points)
        self.names = list(setup.bcs)
        self.max_step, self.min_step = max_step, min_step
        self.growth, self.shrink = growth, shrink
        self.max_attempts = max_attempts
        self.cur = self._vec({})         # last converged bias (terminals' present values)
        self.prev = None                 # the converged bias before it (secant predictor)
        self.next = 0                    # index of the next requested point
        self.step = None                 # current step length (max-norm, V)
        self.attempts = 0
        self.kry = 0
        self.trial = None
        self.result = SweepResult()

    def _vec(self, bias):
        return np.array([float(bias.get(nm, _bc_value(self.setup, nm))) for nm in self.names])

    @property
    def open(self):
        return not self.result.failed and self.next < len(self.points)

    def propose(self):
        """Set the device's terminals to the next trial bias."""
        target = self._vec(self.points[self.next])
        dist = float(np.max(np.abs(target - self.cur))) if target.size else 0.0
        lim = dist
        if self.max_step is not None:
            lim = min(lim, self.max_step)
        if self.step is not None:
            lim = min(lim, self.step)
        frac = 1.0 if dist <= lim or dist == 0.0 else lim / dist
        self.trial = (self.cur + frac * (target - self.cur), frac >= 1.0, lim)
        _apply_bias(self.setup, self.names, self.trial[0])

    def secant_ratio(self):
        """Ratio of the proposed step to the last converged one (0: no
        prediction)."""
        if self.trial is None:
            return 0.0
        return secant_ratio(self.prev, self.cur, self.trial[0])

    def accept(self, its, kry):
        """The trial converged; returns True when it was a requested point."""
        vec, final, _ = self.trial
        R = self.result
        self.kry += int(kry)
        self.prev = self.cur
        self.cur = vec
        R.steps += 1
        self.attempts = 0
        if self.step is not None:
            self.step = self.step * self.growth
            if self.max_step is not None:
                self.step = min(self.step, self.max_step)
        if final:
            R.points.append(self.points[self.next])
            R.newton.append(int(its))
            R.krylov.append(self.kry)
            self.kry = 0
            self.next += 1
        return final

    def accept_point(self, its, kry):
        """The next requested point converged inside a point queue."""
        target = self._vec(self.points[self.next])
        self.trial = (target, True, 0.0)
        _apply_bias(self.setup, self.names, target)
        return self.accept(its, kry)

    def reject_point(self, status, kry):
        """The next requested point failed inside a point queue (the library
        restored the last converged state): back off from it."""
        target = self._vec(self.points[self.next])
        dist = float(np.max(np.abs(target - self.cur))) if target.size else 0.0
        self.trial = (target, True, dist)
        self.reject(status, kry)

    def reject(self, status, kry):
        """The trial failed: back off (the caller restores the state)."""
        _, _, lim = self.trial
        R = self.result
        self.kry += int(kry)
        self.prev = None
        _apply_bias(self.setup, self.names, self.cur)
        R.backoffs += 1
        self.attempts += 1
        self.step = lim * self.shrink
        floor = self.min_step if self.min_step is not None else 1e-4
        if self.step < floor or self.attempts > self.max_attempts:
            why = {_lib.NEWTON_LINEAR_FAIL: "linear solve failed",
                   _lib.NEWTON_NONFINITE: "non-finite Newton step"}.get(
                       int(status), "Newton did not converge")
            R.failed = (f"{why}; step shrank below {floor:g} V before "
                        f"{self.points[self.next]}")


class Secant:
    """The secant predictor's memory for the devices of a batch: whole
    planes holding each device's state at its previous converged point
    (the snapshot a round takes before it solves)."""

    def __init__(self):
        self.planes = None


def solve_round(B, tracks, opts, secant=None):
    """One masked Newton solve over the open tracks (tracks[d] belongs to
    device d of batch B); returns the indices that reached a requested
    point.  Failed devices get their state restored.  With `secant` (a
    Secant), each open device whose step continues its last one starts
    from the secant prediction."""
    todo = [d for d, t in enumerate(tracks) if t is not None and t.open]
    if not todo:
        return []
    for d in todo:
        tracks[d].propose()
    mask = np.zeros(B.ndev, dtype=np.int32)
    mask[todo] = 1
    with B.on_stream():
        # whole-plane snapshot: three copies, cheaper in launches than one
        # per device; only failed devices' ranges are restored from it
        snap = (B.psi.clone(), B.n.clone(), B.p.clone())
        if secant is not None and secant.planes is not None:
            ratios = np.zeros(B.ndev)
            for d in todo:
                ratios[d] = tracks[d].secant_ratio()
            if (ratios > 0).any():
                B.predict(secant.planes, ratios)
        its, conv, kry, status = B.solve_bias(opts, active=mask)
        reached = []
        for d in todo:
            t = tracks[d]
            if conv[d] and status[d] == _lib.NEWTON_OK:
                if t.accept(its[d], kry[d]):
                    reached.append(d)
            else:
                sl = B.device_slice(d)
                B.psi[sl].copy_(snap[0][sl])
                B.n[sl].copy_(snap[1][sl])
                B.p[sl].copy_(snap[2][sl])
                t.reject(status[d], kry[d])
        if secant is not None:
            # each converged device's previous point is the snapshot's
            # state; a failed device's prediction is off until it converges
            # twice more (Track.reject clears its previous bias)
            secant.planes = snap
        B.stream_sync()
    return reached


def device_sweep(dev, biases, opts=None, predictor=True, on_point=None, opts_start=None):
    """The reference's sweep loop (id_vg_sweep, mosfet.py:90-103;
    id_vg_sweep_3d, finfet3d.py:246-263) on one solved GPU device: every
    bias point warm-started from the previous converged one, with
    `predictor` from the secant prediction of the last two (when the bias
    step continues the last one).  A predicted start that does not converge
    is solved again from the previous converged state, as the reference
    starts.  opts_start (default opts) solves the points that start from
    the previous converged state (Newton options tuned on predicted starts,
    such as forcing terms, may cost iterations from a farther start).
    on_point(k, its, kry) is called after each point (the device holds its
    converged state then).  Returns per point (Newton iterations, Krylov
    iterations), both attempts counted."""
    opts = opts or NewtonOptions()
    opts_start = opts_start or opts
    B = dev._gpu()
    names = list(dev._s.bcs)

    cur, prev, prev_state = bias_vector(dev._s, names), None, None
    out = []
    for k, bias in enumerate(biases):
        dev._set_voltages(bias)
        nxt = bias_vector(dev._s, names)
        f = secant_ratio(prev, cur, nxt) if predictor else 0.0
        snap = None
        if predictor:
            with B.on_stream():
                snap = (B.psi.clone(), B.n.clone(), B.p.clone())
                if f > 0.0:
                    B.predict(prev_state, np.array([f]))
        its, conv, kry, status = B.solve_bias(opts if f > 0.0 else opts_start)
        its_tot, kry_tot = int(its[0]), int(kry[0])
        if f > 0.0 and not (conv[0] and status[0] == _lib.NEWTON_OK):
            with B.on_stream():
                B.psi.copy_(snap[0])
                B.n.copy_(snap[1])
                B.p.copy_(snap[2])
            its, conv, kry, status = B.solve_bias(opts_start)
            its_tot += int(its[0])
            kry_tot += int(kry[0])
        dev.last_iterations = its_tot
        dev.last_krylov_iterations = kry_tot
        dev.last_converged = bool(conv[0])
        _raise_outcome(int(status[0]), "device_sweep")
        if not dev.last_converged:
            raise RuntimeError(f"device_sweep: Newton did not converge at {bias}")
        out.append((its_tot, kry_tot))
        if on_point is not None:
            on_point(k, its_tot, kry_tot)
        prev, cur, prev_state = cur, nxt, snap
    B.stream_sync()
    return out


def record_currents(B, tracks, reached, now):
    """Terminal currents of the devices in `reached` after their point."""
    if not reached:
        return
    I = B.terminal_currents()
    c = 0
    per_dev = {}
    for d, nm in B.contacts:
        per_dev.setdefault(d, []).append((nm, I[c]))
        c += 1
    for d in reached:
        tracks[d].result.currents.append(dict(per_dev.get(d, [])))
        tracks[d].result.t_done = now


class BatchedSweep:
    """Run the bias lists `sweeps[d]` (lists of {terminal: V} dicts) of the
    devices `setups[d]` (DeviceSetup, boundary conditions attached) in one
    batch.  `batch` may be passed to work on an existing GPUBatch; its
    devices must be `setups` in order."""

    def __init__(self, setups, sweeps, opts=None, batch=None, max_step=None, min_step=None,
                 growth=1.5, shrink=0.5, max_attempts=64, compact=0.5, queue=False,
                 queue_depth=None, predictor=False):
        self.setups = list(setups)
        self.compact = compact
        # secant predictor between a device's rounds (not across queue
        # passes, whose points the library runs itself)
        self.secant = Secant() if predictor else None
        # queue=True: point queue (tcad_solve_queue): every device runs
        # through up to queue_depth of its requested points per library call
        # without a barrier between devices; a device that fails a point
        # falls back to the rounds below (step backoff).  Only without a step
        # limit: the queue holds the requested points themselves.  Off by
        # default: on a bandwidth-bound batch the rounds, which shrink the
        # batch to the devices still open, were measured faster (256 diodes:
        # 33.3 devices/s in rounds, 30.1 with one queue pass, 32.4 with two
        # points per pass).
        self.queue = queue and max_step is None
        self.queue_depth = queue_depth
        if len(sweeps) != len(self.setups):
            raise ValueError("one bias list per device")
        self.opts = opts or NewtonOptions()
        self.B = batch if batch is not None else GPUBatch(self.setups)
        self.tracks = [Track(s, pts, max_step, min_step, growth, shrink, max_attempts)
                       for s, pts in zip(self.setups, sweeps)]
        self.results = [t.result for t in self.tracks]
        self.rounds = 0

    def run(self, equilibrium=True):
        torch = _torch()
        B, opts = self.B, self.opts
        t0 = time.perf_counter()
        if equilibrium:
            # solved at the terminals' present values, as
            # Device2D.solve_equilibrium is
            its, conv, status = B.solve_equilibrium(opts)
            for d, t in enumerate(self.tracks):
                if status[d] != _lib.NEWTON_OK or not conv[d]:
                    t.result.failed = "equilibrium did not converge"
        ids = list(range(len(self.tracks)))
        self.rebuilds = 0
        while any(t.open for t in self.tracks):
            open_ids = [i for i in ids if self.tracks[i].open]
            if (self.compact and len(open_ids) < len(ids)
                    and self._nodes(open_ids) < self.compact * self._nodes(ids)):
                B, ids = self._subbatch(open_ids), open_ids
                self.rebuilds += 1
            tracks = [self.tracks[i] for i in ids]
            if self.queue and any(t.open and t.step is None for t in tracks):
                # devices in step backoff sit this pass out
                self.queue_pass(B, tracks, t0)
                continue
            self.rounds += 1
            reached = solve_round(B, tracks, opts, self.secant)
            record_currents(B, tracks, reached, time.perf_counter() - t0)
        return self.results

    def queue_pass(self, B, tracks, t0):
        """One tcad_solve_queue call over the open tracks (tracks[d] belongs
        to device d of batch B), up to queue_depth points each."""
        lists = []
        for t in tracks:
            if t is None or not t.open or t.step is not None:
                lists.append([])
                continue
            pts = t.points[t.next:]
            lists.append(pts[:self.queue_depth] if self.queue_depth else pts)
        if not any(lists):
            return
        self.rounds += 1
        with B.on_stream():
            res = B.solve_queue(self.opts, lists)
        t_now = time.perf_counter() - t0
        for t, pl, (done, its, kry, status, cur) in zip(tracks, lists, res):
            if not pl:
                continue
            for k in range(done):
                t.accept_point(its[k], kry[k])
                t.result.currents.append(cur[k])
                t.result.t_done = t_now
            # the predictor's planes do not follow the queue's points
            t.prev = None
            if done < len(pl):
                t.reject_point(status, kry[done] if done < len(kry) else 0)

    def _nodes(self, idx):
        return sum(self.setups[i].N for i in idx)

    def _subbatch(self, idx):
        """A batch of the devices idx over the full batch's planes, at
        their offsets there."""
        full = self.B
        planes = {k: full._t[k] for k in GPUBatch.FIELD_KEYS}
        planes.update(psi=full.psi, n=full.n, p=full.p)
        return GPUBatch([self.setups[i] for i in idx],
                        offsets=[int(full.offsets[i]) for i in idx], planes=planes,
                        stream=full.stream)


def run_lanes(setups, sweeps, opts=None, lanes=1, max_step=None, **sweep_kw):
    """BatchedSweep over `lanes` independent batches, each on its own HIP
    stream driven by its own host thread, so that their launches overlap
    on the GPU.  Devices are grouped by node count (similar sizes share a
    batch, so a batch waits less for its slowest device).  Returns the
    per-device results in the input order and the per-lane sweeps."""
    import threading
    torch = _torch()
    order = np.argsort([s.N for s in setups], kind="stable")
    groups = [list(g) for g in np.array_split(order, lanes) if len(g)]
    results = [None] * len(setups)
    sweeps_out = [None] * len(groups)
    errors = []

    def lane(k, idx):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                B = GPUBatch([setups[i] for i in idx], stream=stream)
                bs = BatchedSweep([setups[i] for i in idx], [sweeps[i] for i in idx], opts,
                                  batch=B, max_step=max_step, **sweep_kw)
                res = bs.run()
            for i, r in zip(idx, res):
                results[i] = r
            sweeps_out[k] = bs
        except Exception as e:             # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=lane, args=(k, g)) for k, g in enumerate(groups)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return results, sweeps_out
