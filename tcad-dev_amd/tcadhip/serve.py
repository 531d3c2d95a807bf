"""Serving loop: continuous batching of bias-sweep requests on one GPU.

A request is one device (DeviceSpec) and its list of bias points, solved
as id_vg_sweep does (pytcad/pytcad/mosfet.py:90-103,
finfet3d.py:246-263): equilibrium, then every point warm-started from the
previous converged state, terminal currents after each, with the step
backoff of continuation.py:61-120 when Newton fails.

Requests are admitted as they arrive, while others are being solved:

  * every admitted device lives in a resident pool (DevicePool): its
    static fields and its state (psi, n, p) occupy a fixed node range of
    pool-owned planes in HBM from admission to retirement, so the warm
    start of every continuation step stays on the device;
  * admission is by HBM budget: a request is admitted when a free node
    range of its size exists in the pool (first fit, in arrival order,
    smaller requests may pass a larger one that does not fit yet);
  * the work is cut into rounds.  A round is one masked Newton solve
    (tcad_solve_bias) over every resident device with points left, so
    requests that arrived at different times share every launch.  When
    devices are admitted, the library batch is rebuilt over the resident
    devices' node ranges (tcad_batch_create with explicit node offsets);
    the state does not move.  New devices get their equilibrium (masked
    to them) before joining the next round;
  * a device retires when its last point is done (or it failed): its
    node range returns to the pool at once, and it leaves the batch at
    the next rebuild (until then it is masked out of every solve).

Results do not depend on what else is resident: the kernels never mix
devices and every per-device reduction runs in a fixed order over that
device's own blocks, so a request gives the same bits alone and in the
mix.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import GPUBatch, NewtonOptions, _torch, setup_from_spec
from .sweep import Secant, SweepResult, Track, record_currents, solve_round

# library workspace per node of a batch (upper bound over 2D/3D): the
# stencil-block matrix in FP64 and FP32 (39 entries in 3D), its node block
# inverse in both, 16 three-plane FP64 vectors, the continuity residual
# plane, the equilibrium system (7 entries, rhs, 6 vectors, 16 planes of
# tile-inverse rows), the tile inverse rows of the coupled solve (36 planes
# of 4 FP16 + 3 FP32 scales) and the per-node contact and gate index tables
WORKSPACE_BYTES_PER_NODE = (39 * 12 + 9 * 12 + 16 * 3 * 8 + 8 + 7 * 8 + 8 + 6 * 8 + 16 * 8
                            + 36 * 8 + 3 * 4 + 2 * 4)
# pool planes per node: 15 static fields + psi, n, p (FP64)
POOL_BYTES_PER_NODE = 18 * 8


def nodes_for_budget(budget_bytes):
    """Pool capacity (nodes) that an HBM budget allows, counting both the
    pool planes and the batch workspace over the same node range."""
    return int(budget_bytes // (POOL_BYTES_PER_NODE + WORKSPACE_BYTES_PER_NODE))


class RangeAllocator:
    """First-fit allocator of node ranges [start, start + n) in [0, capacity)."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]          # sorted (start, length)

    def alloc(self, n):
        for i, (s, L) in enumerate(self.free):
            if L >= n:
                if L == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, L - n)
                return s
        return None

    def release(self, off, n):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for s, L in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + L)
            else:
                merged.append((s, L))
        self.free = merged

    @property
    def used(self):
        return self.capacity - sum(L for _, L in self.free)


class DevicePool(RangeAllocator):
    """Resident node planes (static fields and state) of fixed capacity."""

    def __init__(self, capacity):
        super().__init__(capacity)
        torch = _torch()
        keys = GPUBatch.FIELD_KEYS + ("psi", "n", "p")
        self.planes = {k: torch.zeros(self.capacity, dtype=torch.float64, device="cuda")
                       for k in keys}

    def write(self, setup, off):
        """Static fields of `setup` into its node range; state cleared."""
        torch = _torch()
        f = setup.node_fields()
        for k in GPUBatch.FIELD_KEYS:
            self.planes[k][off:off + setup.N].copy_(torch.from_numpy(f[k]))
        for k in ("psi", "n", "p"):
            self.planes[k][off:off + setup.N].zero_()


@dataclass
class Request:
    rid: int
    spec: object
    arrival: float                   # seconds after the start of the run
    kind: str = ""
    keep_state: bool = False         # hand back the final (psi, n, p)


@dataclass
class Served:
    """Outcome of one request."""
    rid: int
    kind: str
    nodes: int
    points: int
    arrival: float
    admitted: float = 0.0
    done: float = 0.0
    result: SweepResult = field(default_factory=SweepResult)
    state: tuple = None              # final scaled (psi, n, p) on the host, if asked
    @property
    def latency(self):
        return self.done - self.arrival


class _Lane:
    """One continuous-batching loop: its own pool share, HIP stream (None =
    the default stream) and, with several lanes, its own host thread."""

    def __init__(self, capacity, server, stream):
        self.server = server
        self.stream = stream
        with self._ctx():
            self.pool = DevicePool(capacity)
        self.queue = []                 # routed here, not admitted yet
        self.members, self.live = [], []
        self.B = None
        self.rounds = self.rebuilds = self.max_resident = 0
        self.load = 0                   # outstanding node-points routed here
        # the secant predictor's previous states live in whole pool-sized
        # planes (the round snapshots), indexed by each device's offset
        self.secant = Secant() if server.predictor else None

    def _ctx(self):
        import contextlib
        torch = _torch()
        return contextlib.nullcontext() if self.stream is None else torch.cuda.stream(self.stream)

    @property
    def busy(self):
        return bool(self.queue or self.live)

    def step(self, now, out, lock=None):
        """Admit what fits, then one round over the resident devices.
        Returns the number of devices that reached a point.  `lock` guards
        the queue and load against the dispatcher thread."""
        import contextlib
        guard = lock if lock is not None else contextlib.nullcontext()
        srv = self.server
        new = []
        with guard:
            routed = list(self.queue)
        for r in routed:
            off = self.pool.alloc(r.spec.N)
            if off is None:
                continue
            setup = setup_from_spec(r.spec)
            self.pool.write(setup, off)
            slot = dict(req=r, setup=setup, off=off,
                        track=Track(setup, r.spec.sweep, max_step=srv.max_step))
            out[r.rid] = Served(r.rid, r.kind, r.spec.N, len(r.spec.sweep), r.arrival,
                                admitted=now(), result=slot["track"].result)
            self.live.append(slot)
            new.append(slot)
            with guard:
                self.queue.remove(r)
        if not self.live:
            self.B, self.members = None, []
            return 0
        if new:
            self.B = None                         # release the old workspace first
            self.members = list(self.live)
            self.B = GPUBatch([m["setup"] for m in self.members],
                              offsets=[m["off"] for m in self.members], planes=self.pool.planes,
                              stream=self.stream)
            self.rebuilds += 1
            self.max_resident = max(self.max_resident, len(self.members))
            mask = np.array([1 if m in new else 0 for m in self.members], dtype=np.int32)
            its, conv, status = self.B.solve_equilibrium(srv.eq_opts, active=mask)
            for d, m in enumerate(self.members):
                if mask[d] and (status[d] != _lib.NEWTON_OK or not conv[d]):
                    m["track"].result.failed = "equilibrium did not converge"
        tracks = [m["track"] for m in self.members]
        reached = solve_round(self.B, tracks, srv.opts, self.secant)
        self.rounds += 1
        record_currents(self.B, tracks, reached, now())
        keep = []
        for m in self.live:
            if m["track"].open:
                keep.append(m)
                continue
            out[m["req"].rid].done = now()
            if m["req"].keep_state:
                sl = slice(m["off"], m["off"] + m["setup"].N)
                with self._ctx():
                    out[m["req"].rid].state = tuple(
                        self.pool.planes[k][sl].cpu().numpy().reshape(m["setup"].shape)
                        for k in ("psi", "n", "p"))
            self.pool.release(m["off"], m["setup"].N)
            with guard:
                self.load -= m["req"].spec.N * len(m["req"].spec.sweep)
        self.live = keep
        return len(reached)


class Server:
    """Continuous-batching solver.  `opts` is used for every bias point;
    equilibrium uses `eq_opts`.  With lanes > 1 the pool is split into
    that many lanes, each an independent loop on its own HIP stream and
    host thread, so that their launches overlap on the GPU; an arriving
    request goes to the lane with the least outstanding work (node count
    x points) that can hold it."""

    def __init__(self, capacity_nodes, opts=None, eq_opts=None, max_step=None, lanes=1,
                 predictor=False):
        self.opts = opts or NewtonOptions()
        self.eq_opts = eq_opts or NewtonOptions(max_iter=400)
        self.max_step = max_step
        # secant predictor between a request's continuation steps
        # (tcadhip.sweep; a request's results alone and in the mix stay
        # the same bits: the prediction of a device reads only its own range)
        self.predictor = predictor
        torch = _torch()
        self.lanes = [_Lane(int(capacity_nodes) // lanes, self,
                            None if lanes == 1 else torch.cuda.Stream())
                      for _ in range(lanes)]

    @property
    def rounds(self):
        return sum(L.rounds for L in self.lanes)

    @property
    def rebuilds(self):
        return sum(L.rebuilds for L in self.lanes)

    @property
    def max_resident(self):
        return max(L.max_resident for L in self.lanes)

    def _route(self, r):
        fits = [L for L in self.lanes if r.spec.N <= L.pool.capacity]
        if not fits:
            raise ValueError(f"request {r.rid}: {r.spec.N} nodes exceed the lane capacity "
                             f"{self.lanes[0].pool.capacity}")
        L = min(fits, key=lambda L: L.load)
        L.load += r.spec.N * len(r.spec.sweep)
        L.queue.append(r)

    def run(self, requests, realtime=True, log=None):
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if not any(r.spec.N <= L.pool.capacity for L in self.lanes):
                raise ValueError(f"request {r.rid}: {r.spec.N} nodes exceed the lane "
                                 f"capacity {self.lanes[0].pool.capacity}")
        out = {}
        if len(self.lanes) == 1:
            self._run_single(pending, realtime, out, log)
        else:
            if not realtime:
                raise ValueError("several lanes run in real time only")
            self._run_lanes(pending, out)
        return [out[r.rid] for r in sorted(requests, key=lambda r: r.rid)]

    def _run_single(self, pending, realtime, out, log):
        L = self.lanes[0]
        t0 = time.perf_counter()
        vt = [0.0]                                # virtual clock when not real time

        def now():
            return time.perf_counter() - t0 if realtime else vt[0]

        while pending or L.busy:
            while pending and pending[0].arrival <= now():
                self._route(pending.pop(0))
            if not L.busy:
                wait = pending[0].arrival - now()
                if realtime and wait > 0:
                    time.sleep(wait)
                elif not realtime:
                    vt[0] = max(vt[0], pending[0].arrival)
                continue
            t_round = time.perf_counter()
            reached = L.step(now, out)
            if not realtime:
                vt[0] += time.perf_counter() - t_round
            if log is not None and reached:
                log(f"t={now():.2f}s live={len(L.live)} queue={len(L.queue)} "
                    f"pool={L.pool.used}/{L.pool.capacity}")

    def _run_lanes(self, pending, out):
        import threading
        torch = _torch()
        lock = threading.Condition()
        state = {"arrivals_done": False}
        errors = []
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def lane_loop(L):
            try:
                with torch.cuda.stream(L.stream):
                    while True:
                        with lock:
                            while not L.busy and not state["arrivals_done"]:
                                lock.wait(0.05)
                            if not L.busy and state["arrivals_done"]:
                                return
                        L.step(now, out, lock)
            except Exception as e:                     # re-raised by the caller
                errors.append(e)
                with lock:
                    state["arrivals_done"] = True
                    lock.notify_all()

        threads = [threading.Thread(target=lane_loop, args=(L,)) for L in self.lanes]
        for t in threads:
            t.start()
        try:
            for r in pending:
                wait = r.arrival - now()
                if wait > 0:
                    time.sleep(wait)
                if errors:
                    break
                with lock:
                    self._route(r)
                    lock.notify_all()
        finally:
            with lock:
                state["arrivals_done"] = True
                lock.notify_all()
            for t in threads:
                t.join()
        if errors:
            raise errors[0]


def request_stream(count, rate, seed, mix=("diode", "mosfet", "trigate"), max_nodes=200_000,
                   min_points=5, max_points=40, small=False):
    """Seeded Poisson arrivals (rate per second) of mixed sweep requests:
    2D diodes (forward/reverse bias ramps), 2D MOSFET Id-Vg sweeps and 3D
    tri-gate Id-Vg sweeps, 5-40 points each.  `small` shrinks the meshes
    (for tests)."""
    from . import configs
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        kind = mix[int(rng.integers(len(mix)))]
        k = int(rng.integers(min_points, max_points + 1))
        if kind == "diode":
            nx, ny = (int(rng.integers(16, 41)), int(rng.integers(10, 25))) if small else \
                (int(rng.integers(50, 161)), int(rng.integers(40, 101)))
            Na = float(10 ** rng.uniform(15, 19))
            Nd = float(10 ** rng.uniform(15, 19))
            vt = float(rng.uniform(-2.0, 0.8))
            biases = [vt * (s + 1) / k for s in range(k)]
            spec = configs.diode_2d(nx, ny, xj_frac=float(rng.uniform(0.3, 0.7)), Na=Na, Nd=Nd,
                                    biases=biases,
                                    models=dict(doping_mobility=True, srh=True, auger=True,
                                                bgn=True, surface_mobility=False),
                                    material={}, name=f"req{rid}_diode_{nx}x{ny}")
        elif kind == "mosfet":
            nx, ny = (int(rng.integers(20, 41)), int(rng.integers(14, 25))) if small else \
                (int(rng.integers(100, 201)), int(rng.integers(60, 121)))
            Vd = float(rng.choice([0.05, 0.5]))
            spec = configs.mosfet_2d(nx, ny, Vd=Vd, vg_list=np.linspace(0.0, 1.0, k),
                                     name=f"req{rid}_mosfet_{nx}x{ny}")
        else:
            if small:
                nx, ny, nz = int(rng.integers(10, 17)), int(rng.integers(8, 13)), \
                    int(rng.integers(6, 11))
            else:
                while True:
                    nx, ny, nz = int(rng.integers(32, 65)), int(rng.integers(24, 57)), \
                        int(rng.integers(24, 57))
                    if nx * ny * nz <= max_nodes:
                        break
            Vd = float(rng.choice([0.05, 0.7]))
            spec = configs.trigate_3d(nx, ny, nz, Vd=Vd, vg_list=np.linspace(0.0, 0.8, k),
                                      name=f"req{rid}_trigate_{nx}x{ny}x{nz}")
        out.append(Request(rid, spec, t, kind))
    return out


def serve_alone(spec, opts=None, eq_opts=None, predictor=False):
    """The same request on its own (a batch of one), for comparison."""
    from .sweep import BatchedSweep
    setup = setup_from_spec(spec)
    opts = opts or NewtonOptions()
    eq_opts = eq_opts or NewtonOptions(max_iter=400)
    B = GPUBatch([setup])
    its, conv, status = B.solve_equilibrium(eq_opts)
    bs = BatchedSweep([setup], [spec.sweep], opts, batch=B, predictor=predictor)
    return bs.run(equilibrium=False)[0]
