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
from .sweep import SweepResult, Track, record_currents, solve_round

# library workspace per node of a batch (upper bound over 2D/3D): the
# stencil-block matrix in FP64 and FP32 (39 entries in 3D), its block
# inverse in both, 15 three-plane FP64 vectors, the equilibrium system
WORKSPACE_BYTES_PER_NODE = 39 * 12 + 9 * 12 + 15 * 3 * 8 + 8 + 7 * 8 + 8 + 5 * 8
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

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    """Continuous-batching solver over one DevicePool.  `opts` is used for
    every bias point; equilibrium uses `eq_opts`."""

    def __init__(self, capacity_nodes, opts=None, eq_opts=None, max_step=None):
        self.pool = DevicePool(capacity_nodes)
        self.opts = opts or NewtonOptions()
        self.eq_opts = eq_opts or NewtonOptions(max_iter=400)
        self.max_step = max_step
        self.rounds = 0
        self.rebuilds = 0
        self.max_resident = 0

    def run(self, requests, realtime=True, log=None):
        torch = _torch()
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if r.spec.N > self.pool.capacity:
                raise ValueError(f"request {r.rid}: {r.spec.N} nodes exceed the pool "
                                 f"capacity {self.pool.capacity}")
        queue, members, live, out = [], [], [], {}
        B = None
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        vt = 0.0                      # virtual clock when not real time

        def now():
            return clock() if realtime else vt

        while pending or queue or live:
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.pop(0))
            # admission by HBM budget, arrival order, first fit
            new = []
            for r in list(queue):
                off = self.pool.alloc(r.spec.N)
                if off is None:
                    continue
                setup = setup_from_spec(r.spec)
                self.pool.write(setup, off)
                slot = dict(req=r, setup=setup, off=off,
                            track=Track(setup, r.spec.sweep, max_step=self.max_step))
                out[r.rid] = Served(r.rid, r.kind, r.spec.N, len(r.spec.sweep), r.arrival,
                                    admitted=t, result=slot["track"].result)
                live.append(slot)
                new.append(slot)
                queue.remove(r)
            if not live:
                B, members = None, []
                if pending:
                    wait = pending[0].arrival - now()
                    if realtime and wait > 0:
                        time.sleep(wait)
                    elif not realtime:
                        vt = max(vt, pending[0].arrival)
                continue
            if new:
                B = None                      # release the old workspace first
                members = list(live)
                B = GPUBatch([s["setup"] for s in members],
                             offsets=[s["off"] for s in members], planes=self.pool.planes)
                self.rebuilds += 1
                self.max_resident = max(self.max_resident, len(members))
                mask = np.array([1 if s in new else 0 for s in members], dtype=np.int32)
                its, conv, status = B.solve_equilibrium(self.eq_opts, active=mask)
                for d, s in enumerate(members):
                    if mask[d] and (status[d] != _lib.NEWTON_OK or not conv[d]):
                        s["track"].result.failed = "equilibrium did not converge"
            tracks = [s["track"] for s in members]
            t_round = time.perf_counter()
            reached = solve_round(B, tracks, self.opts)
            self.rounds += 1
            if not realtime:
                vt += time.perf_counter() - t_round
            record_currents(B, tracks, reached, now())
            keep = []
            for s in live:
                if s["track"].open:
                    keep.append(s)
                    continue
                out[s["req"].rid].done = now()
                self.pool.release(s["off"], s["setup"].N)
            live = keep
            if log is not None and reached:
                log(f"t={now():.2f}s live={len(live)} queue={len(queue)} "
                    f"pool={self.pool.used}/{self.pool.capacity}")
        return [out[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


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


def serve_alone(spec, opts=None, eq_opts=None):
    """The same request on its own (a batch of one), for comparison."""
    from .sweep import BatchedSweep
    setup = setup_from_spec(spec)
    opts = opts or NewtonOptions()
    eq_opts = eq_opts or NewtonOptions(max_iter=400)
    B = GPUBatch([setup])
    its, conv, status = B.solve_equilibrium(eq_opts)
    bs = BatchedSweep([setup], [spec.sweep], opts, batch=B)
    return bs.run(equilibrium=False)[0]
