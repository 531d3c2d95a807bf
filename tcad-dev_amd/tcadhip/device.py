"""GPU-backed Device2D / Device3D with the reference's operator API.

Mirrors pytcad/pytcad/device2d.py (Device2D: __init__ :168-453,
add_contact/add_gate :672-680, solve_equilibrium :1308-1373,
_residual_jacobian :1402-1920, solve_bias :1923-2191, terminal_current
:2194-2227) and pytcad/pytcad/device3d.py (same methods, :437-2306) for
the structured Boltzmann path: one semiconductor, ohmic contacts, Robin
gates, SRH/Auger, BGN, Caughey-Thomas mobility, lagged Lombardi surface
mobility (2D).  Every Newton step runs on the GPU through libtcadhip.so;
the host side here only computes the one-time coefficient arrays
(the reference's __init__ arithmetic) and moves scalars.

State (psi, n, p) stays resident on the GPU between bias points (the
warm start of solve_bias); the `psi`, `n`, `p` attributes copy it to
the host on access.
"""
import ctypes
import warnings
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import physics as ph
from .mesh import control_volume_widths

_MODEL_FLAGS = ("doping_mobility", "srh", "auger", "bgn", "surface_mobility")


@dataclass
class NewtonOptions:
    """pytcad/pytcad/device.py:539-628 for the fields this path reads.

    linsolve: "auto" | "direct" | "bicgstab" -> FP64 BiCGStab with the
    block-Jacobi preconditioner `precond` names (no direct factorisation
    runs on the GPU); "bicgstab_mixed" -> BiCGStab on the FP32-stored
    matrix inside FP64 iterative refinement.  linsolve_rtol is the relative residual of the
    (row-equilibrated) Newton system.  forcing: inexact early Newton
    solves (target clamp(0.01 * previous update, linsolve_rtol, 1e-4));
    the converging step is refined to the FP64 rounding floor either way.
    Forcing is off by default: on the 200x120 MOSFET sweep it added a
    Newton iteration per point and cost more Krylov work than it saved.
    precond: "tile" (default; every 16-node tile -- 4 x 4 in 2D, 2 x 1 x 8
    in 3D -- gets the exact inverse of its 48 x 48 diagonal block, stored
    as FP16 rows with a FP32 scale per row) or "node" (the reference's 3x3
    node block-Jacobi for every device)."""
    max_iter: int = 100
    tol_update: float = 1e-8
    max_dpsi: float = 5.0
    linsolve: str = "auto"
    linsolve_rtol: float = 1e-10
    lin_maxiter: int = 20000
    forcing: bool = False
    precond: str = "tile"
    verbose: bool = False
    # run-time tuning (tcad_set_tuning): forcing eta = clamp(forcing_slope *
    # err, linsolve_rtol, forcing_max); forcing_fp32: loose forcing solves on
    # the FP32-stored matrix; tile_rows "auto" | "reg" | "mem"; graph: BiCGStab
    # iterations replayed from a captured HIP graph; ir_continue: refinement
    # sweeps continue the previous sweep's BiCGStab ("fp32": only inner
    # solves on the FP32-stored matrix, "all", "never"); fp32_inner: the
    # per-sweep residual reduction of inner solves on the FP32 matrix;
    # forcing_pow: eta = forcing_slope * err^forcing_pow; tile_lag: Newton
    # steps after an update below it keep the tile inverses of the step before;
    # linear_fallback: an unconverged Newton solve is redone, for that
    # iteration, by GMRES + ILU(0) on the device's CSR system
    # (device3d.py:2154-2168's per-iteration fallback); forcing_predicted:
    # forcing terms only for devices whose start the secant predictor moved
    # (the others solve every Newton step to linsolve_rtol); tile_rb:
    # red-black tile ordering of the tile preconditioner (the black tiles'
    # inverses applied after the red ones', to the residual less their
    # couplings to them: forward block Gauss-Seidel) instead of tile
    # block-Jacobi
    forcing_slope: float = 0.01
    forcing_max: float = 1e-4
    forcing_fp32: bool = False
    tile_rows: str = "auto"
    graph: bool = False
    ir_continue: str = "fp32"
    fp32_inner: float = 1e-6
    forcing_pow: int = 1
    tile_lag: float = 0.0
    linear_fallback: bool = True
    forcing_predicted: bool = False
    tile_rb: bool = True

    def apply_tuning(self, handle):
        L = _lib.lib()
        rows = {"auto": 0, "reg": 1, "mem": 2}
        if self.tile_rows not in rows:
            raise ValueError(f"unknown tile_rows {self.tile_rows!r}")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_TILE_ROWS, float(rows[self.tile_rows])),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FORCING_SLOPE, float(self.forcing_slope)),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FORCING_MAX, float(self.forcing_max)),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FORCING_FP32,
                                     1.0 if self.forcing_fp32 else 0.0), "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_GRAPH, 1.0 if self.graph else 0.0),
                   "tcad_set_tuning")
        cont = {"never": 0, "fp32": 1, "all": 2}
        if self.ir_continue not in cont:
            raise ValueError(f"unknown ir_continue {self.ir_continue!r}")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_IR_CONTINUE,
                                     float(cont[self.ir_continue])), "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FP32_INNER, float(self.fp32_inner)),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FORCING_POW, float(self.forcing_pow)),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_TILE_LAG, float(self.tile_lag)),
                   "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_LINEAR_FALLBACK,
                                     1.0 if self.linear_fallback else 0.0), "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_FORCING_PREDICTED,
                                     1.0 if self.forcing_predicted else 0.0), "tcad_set_tuning")
        _lib.check(L.tcad_set_tuning(handle, _lib.TUNE_TILE_RB, 1.0 if self.tile_rb else 0.0),
                   "tcad_set_tuning")

    def lin_mode(self):
        m = self.linsolve
        if m in ("auto", "direct", "bicgstab", "gmres"):
            mode = _lib.LIN_BICGSTAB_F64
        elif m in ("bicgstab_mixed", "mixed"):
            mode = _lib.LIN_BICGSTAB_MIXED
        else:
            raise ValueError(f"unknown linsolve {m!r}")
        if self.precond not in ("tile", "node"):
            raise ValueError(f"unknown precond {self.precond!r}")
        return (mode | (_lib.LIN_FORCING if self.forcing else 0)
                | (_lib.PRECOND_NODE if self.precond == "node" else 0))


class DirichletBC:
    def __init__(self, idx, V):
        self.idx = idx
        self.V = float(V)


class GateBC:
    def __init__(self, idx, kappa, Vfb, Vg, normal_axis):
        self.idx = idx
        self.kappa = float(kappa)
        self.Vfb = float(Vfb)
        self.Vg = float(Vg)
        self.normal_axis = normal_axis


def _hmean(lo, hi):
    return 2.0 * lo * hi / (lo + hi)


class DeviceSetup:
    """One device's coefficient arrays, computed once on the host with the
    reference's __init__ arithmetic (device2d.py:319-445,
    device3d.py:537-695), in array order (Ny,Nx) / (Nz,Ny,Nx)."""

    def __init__(self, coords, doping, Ntotal=None, T=300.0, material=ph.SILICON,
                 models=None):
        models = models or ph.Models()
        if isinstance(models, dict):
            models = ph.Models(**models)
        self.dim = len(coords)
        if self.dim not in (2, 3):
            raise ValueError("coords must be [x, y] or [x, y, z]")
        if self.dim == 3 and models.surface_mobility:
            raise NotImplementedError("surface_mobility is a 2D model (device.py:431-434)")
        self.coords = [np.asarray(c, dtype=float) for c in coords]
        self.shape = tuple(c.size for c in self.coords[::-1])
        self.N = int(np.prod(self.shape))
        self.models = models
        self.mat = material
        self.T = float(T)
        self.doping = np.asarray(doping, dtype=float).reshape(self.shape)
        self.Ntot = (np.abs(self.doping) if Ntotal is None
                     else np.asarray(Ntotal, dtype=float).reshape(self.shape))
        if self.Ntot.max() > 1e19:
            warnings.warn("Doping exceeds ~1e19 cm^-3: Boltzmann statistics used here "
                          "overestimate the carrier density.")
        self.VT = ph.thermal_voltage(T)
        self.eps = material.eps_r * ph.EPS0
        self.ni = material.ni(T)
        self.Ns = max(float(np.abs(self.doping).max()), self.ni)
        self.LD = np.sqrt(self.eps * self.VT / (ph.Q * self.Ns))
        self.J0 = ph.Q * ph.D0_REF * self.Ns / self.LD
        self.R0 = ph.D0_REF * self.Ns / self.LD ** 2
        self.h = [np.diff(c / self.LD) for c in self.coords]
        self.dVw = [control_volume_widths(h) for h in self.h]
        if self.dim == 2:
            self.dV = np.outer(self.dVw[1], self.dVw[0])
        else:
            self.dV = (self.dVw[2][:, None, None] * self.dVw[1][None, :, None]
                       * self.dVw[0][None, None, :])
        self.C = self.doping / self.Ns
        self.nie = ph.nie_effective(self.Ntot, material, T, models.bgn)
        self.nie_s = self.nie / self.Ns
        if models.doping_mobility:
            self.mu_n0 = ph.mobility_caughey_thomas(self.Ntot, material, T, "n")
            self.mu_p0 = ph.mobility_caughey_thomas(self.Ntot, material, T, "p")
        else:
            self.mu_n0 = np.full(self.shape, material.mu_n_max)
            self.mu_p0 = np.full(self.shape, material.mu_p_max)
        self.tau_n = ph.lifetime_scharfetter(self.Ntot, material.tau_n0, material.tau_Nref)
        self.tau_p = ph.lifetime_scharfetter(self.Ntot, material.tau_p0, material.tau_Nref)
        # node-shaped edge arrays: entry k is the edge k -> k + axis stride
        self.an, self.ap, self.dln = [], [], []
        for a in range(self.dim):
            ax = self.dim - 1 - a
            lo = [slice(None)] * self.dim
            hi = [slice(None)] * self.dim
            lo[ax] = slice(None, -1)
            hi[ax] = slice(1, None)
            lo, hi = tuple(lo), tuple(hi)
            hb = self.h[a].reshape([-1 if b == ax else 1 for b in range(self.dim)])
            dn = _hmean(self.mu_n0[lo], self.mu_n0[hi]) * self.VT / ph.D0_REF
            dp = _hmean(self.mu_p0[lo], self.mu_p0[hi]) * self.VT / ph.D0_REF
            for arr, val in ((self.an, dn / hb), (self.ap, dp / hb),
                             (self.dln, np.log(self.nie_s[hi] / self.nie_s[lo]))):
                full = np.zeros(self.shape)
                full[lo] = val
                arr.append(full)
        self.bcs = OrderedDict()

    def flat(self, idx):
        return np.ravel_multi_index(tuple(np.asarray(i, dtype=np.int64) for i in idx),
                                    self.shape).astype(np.int32)

    def gate_face_weight(self, bc):
        """device3d.py:939-947; 2D gates are y-normal, weight dVx[i]."""
        a = "xyz".index(bc.normal_axis)
        if a >= self.dim:
            raise ValueError(f"normal_axis {bc.normal_axis!r} invalid in {self.dim}D")
        others = [b for b in range(self.dim) if b != a]
        w = self.dVw[others[0]][bc.idx[self.dim - 1 - others[0]]]
        for b in others[1:]:
            w = w * self.dVw[b][bc.idx[self.dim - 1 - b]]
        return w

    def contact_values(self, bc, V):
        """contacts.py:21-37 at the contact's nodes: (psi0, n0, p0, ln(n0/nie))."""
        return ph.ohmic_values(self.C[bc.idx], self.nie_s[bc.idx], V, self.VT)

    def contacts(self):
        return [(nm, bc) for nm, bc in self.bcs.items() if isinstance(bc, DirichletBC)]

    def gates(self):
        return [(nm, bc) for nm, bc in self.bcs.items() if isinstance(bc, GateBC)]

    def description(self, keep, node_offset=-1):
        """A DeviceDesc; host arrays it points to are appended to `keep`.
        node_offset < 0 packs the device after the previous one."""
        def arr(a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        D = _lib.DeviceDesc()
        D.dim = self.dim
        D.nx = self.coords[0].size
        D.ny = self.coords[1].size
        D.nz = self.coords[2].size if self.dim == 3 else 1
        hs = [arr(h) for h in self.h] + [None] * (3 - self.dim)
        dvs = [arr(v) for v in self.dVw] + [None] * (3 - self.dim)
        D.h_hx, D.h_hy, D.h_hz = (_lib.dptr(x) for x in hs)
        D.h_dvx, D.h_dvy, D.h_dvz = (_lib.dptr(x) for x in dvs)
        D.VT, D.Ns, D.LD, D.J0, D.R0 = self.VT, self.Ns, self.LD, self.J0, self.R0
        D.Cn_auger, D.Cp_auger = self.mat.Cn_auger, self.mat.Cp_auger
        D.srh, D.auger = int(self.models.srh), int(self.models.auger)
        D.surface_mobility = int(self.models.surface_mobility and self.dim == 2)
        D.T = self.T
        if D.surface_mobility:
            D.h_mun_row0 = _lib.dptr(arr(self.mu_n0[0, :]))
            D.h_mun_row1 = _lib.dptr(arr(self.mu_n0[1, :]))
            D.h_mup_row0 = _lib.dptr(arr(self.mu_p0[0, :]))
            D.h_mup_row1 = _lib.dptr(arr(self.mu_p0[1, :]))
        cs = self.contacts()
        ptr, nodes, lnr, n0, p0 = [0], [], [], [], []
        for _, bc in cs:
            kk = self.flat(bc.idx)
            _, nn0, pp0, ll = self.contact_values(bc, 0.0)
            nodes.append(kk); lnr.append(ll); n0.append(nn0); p0.append(pp0)
            ptr.append(ptr[-1] + kk.size)
        D.ncontacts = len(cs)
        cat = (lambda xs, dt: arr(np.concatenate(xs) if xs else np.zeros(0), dt))
        D.h_contact_ptr = _lib.iptr(arr(ptr, np.int32))
        D.h_contact_nodes = _lib.iptr(cat(nodes, np.int32))
        D.h_contact_lnr = _lib.dptr(cat(lnr, np.float64))
        D.h_contact_n0 = _lib.dptr(cat(n0, np.float64))
        D.h_contact_p0 = _lib.dptr(cat(p0, np.float64))
        gs = self.gates()
        ptr, nodes, kw, psib = [0], [], [], []
        for _, bc in gs:
            kk = self.flat(bc.idx)
            nodes.append(kk)
            kw.append(np.broadcast_to(bc.kappa * self.gate_face_weight(bc), kk.shape))
            psib.append(np.arcsinh(self.C[bc.idx] / (2.0 * self.nie_s[bc.idx])))
            ptr.append(ptr[-1] + kk.size)
        D.ngates = len(gs)
        D.h_gate_ptr = _lib.iptr(arr(ptr, np.int32))
        D.h_gate_nodes = _lib.iptr(cat(nodes, np.int32))
        D.h_gate_kw = _lib.dptr(cat(kw, np.float64))
        D.h_gate_psib = _lib.dptr(cat(psib, np.float64))
        D.node_offset = int(node_offset)
        return D

    def node_fields(self):
        """Per-node static arrays (flattened) keyed as in StaticFields."""
        f = {"C": self.C, "nie_s": self.nie_s, "nie": self.nie, "tau_n": self.tau_n,
             "tau_p": self.tau_p, "dV": self.dV}
        for a in range(3):
            z = np.zeros(self.shape)
            f[f"an{a}"] = self.an[a] if a < self.dim else z
            f[f"ap{a}"] = self.ap[a] if a < self.dim else z
            f[f"dln{a}"] = self.dln[a] if a < self.dim else z
        return {k: np.ascontiguousarray(v, dtype=np.float64).ravel() for k, v in f.items()}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("tcadhip needs a HIP (ROCm) GPU; torch.cuda.is_available() is False")
    return torch


class GPUBatch:
    """A libtcadhip batch: the devices of `setups` share every launch.

    By default the devices are packed into planes this object allocates.
    With `planes` (a dict of CUDA float64 tensors for FIELD_KEYS and
    psi, n, p) and `offsets` (first node of each device), the batch works
    in place on caller-owned planes, where each device's static fields
    are already written at its range: a resident pool re-batches its
    devices this way without moving their state."""

    FIELD_KEYS = ("C", "nie_s", "nie", "tau_n", "tau_p", "dV",
                  "an0", "an1", "an2", "ap0", "ap1", "ap2", "dln0", "dln1", "dln2")

    def __init__(self, setups, offsets=None, planes=None, stream=None):
        torch = _torch()
        L = _lib.lib()
        self.setups = list(setups)
        # every launch of this batch goes to `stream` (a torch.cuda.Stream,
        # None = the default stream); torch work on the same planes belongs
        # on it too (see on_stream)
        self.stream = stream
        if planes is None:
            self.offsets = np.cumsum([0] + [s.N for s in self.setups])[:-1]
            host = {k: np.concatenate([s.node_fields()[k] for s in self.setups])
                    for k in self.FIELD_KEYS}
            self._t = {k: torch.from_numpy(v).to("cuda") for k, v in host.items()}
            N = sum(s.N for s in self.setups)
            self.psi = torch.zeros(N, dtype=torch.float64, device="cuda")
            self.n = torch.zeros_like(self.psi)
            self.p = torch.zeros_like(self.psi)
        else:
            self.offsets = np.asarray(offsets, dtype=np.int64)
            self._t = {k: planes[k] for k in self.FIELD_KEYS}
            self.psi, self.n, self.p = planes["psi"], planes["n"], planes["p"]
        self.N = int(max(o + s.N for o, s in zip(self.offsets, self.setups)))
        torch.cuda.current_stream().synchronize()          # fields uploaded
        F = _lib.StaticFields()
        F.d_C, F.d_nie_s, F.d_nie = (self._t[k].data_ptr() for k in ("C", "nie_s", "nie"))
        F.d_tau_n, F.d_tau_p, F.d_dV = (self._t[k].data_ptr() for k in ("tau_n", "tau_p", "dV"))
        for a in range(3):
            F.d_an[a] = self._t[f"an{a}"].data_ptr()
            F.d_ap[a] = self._t[f"ap{a}"].data_ptr()
            F.d_dlnnie[a] = self._t[f"dln{a}"].data_ptr()
        keep = []
        descs = (_lib.DeviceDesc * len(self.setups))(
            *[s.description(keep, int(o)) for s, o in zip(self.setups, self.offsets)])
        self._handle = ctypes.c_void_p()
        sh = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _lib.check(L.tcad_batch_create(descs, len(self.setups), ctypes.byref(F), sh,
                                       ctypes.byref(self._handle)), "tcad_batch_create")
        self._keep = keep
        self.contacts = [(d, nm) for d, s in enumerate(self.setups) for nm, _ in s.contacts()]
        self.gates = [(d, nm) for d, s in enumerate(self.setups) for nm, _ in s.gates()]
        self.ndev = len(self.setups)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().tcad_batch_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def handle(self):
        return self._handle

    def stream_sync(self):
        """Wait for the work queued on this batch's stream."""
        torch = _torch()
        (self.stream or torch.cuda.current_stream()).synchronize()

    def on_stream(self):
        """Context that makes this batch's stream torch's current stream."""
        import contextlib
        torch = _torch()
        return contextlib.nullcontext() if self.stream is None else torch.cuda.stream(self.stream)

    def set_bias(self):
        Vc = np.array([s.bcs[nm].V for d, nm in self.contacts for s in [self.setups[d]]],
                      dtype=np.float64)
        Vg = np.array([s.bcs[nm].Vg for d, nm in self.gates for s in [self.setups[d]]],
                      dtype=np.float64)
        Vfb = np.array([s.bcs[nm].Vfb for d, nm in self.gates for s in [self.setups[d]]],
                       dtype=np.float64)
        _lib.check(_lib.lib().tcad_set_bias(
            self._handle, _lib.dptr(Vc) if Vc.size else None,
            _lib.dptr(Vg) if Vg.size else None,
            _lib.dptr(Vfb) if Vfb.size else None), "tcad_set_bias")

    def _mask(self, active):
        if active is None:
            return None
        m = np.ascontiguousarray(active, dtype=np.int32)
        if m.shape != (self.ndev,):
            raise ValueError(f"active mask must have {self.ndev} entries")
        return m

    def solve_equilibrium(self, opts, active=None):
        """Equilibrium of the selected devices; per device (iterations,
        converged, outcome code)."""
        its = np.zeros(self.ndev, dtype=np.int32)
        conv = np.zeros(self.ndev, dtype=np.int32)
        status = np.zeros(self.ndev, dtype=np.int32)
        m = self._mask(active)
        self.set_bias()
        _lib.check(_lib.lib().tcad_solve_equilibrium(
            self._handle, opts.max_iter, opts.tol_update, opts.max_dpsi, 1, _lib.iptr(m),
            self.psi.data_ptr(), self.n.data_ptr(), self.p.data_ptr(),
            _lib.iptr(its), _lib.iptr(conv), _lib.iptr(status)), "tcad_solve_equilibrium")
        return its, conv, status

    def solve_bias(self, opts, active=None):
        """Coupled Newton of the selected devices; per device (iterations,
        converged, Krylov iterations, outcome code)."""
        its = np.zeros(self.ndev, dtype=np.int32)
        conv = np.zeros(self.ndev, dtype=np.int32)
        kry = np.zeros(self.ndev, dtype=np.int32)
        status = np.zeros(self.ndev, dtype=np.int32)
        m = self._mask(active)
        self.set_bias()
        opts.apply_tuning(self._handle)
        _lib.check(_lib.lib().tcad_solve_bias(
            self._handle, opts.max_iter, opts.tol_update, opts.max_dpsi, opts.lin_mode(),
            opts.linsolve_rtol, opts.lin_maxiter, _lib.iptr(m), self.psi.data_ptr(),
            self.n.data_ptr(), self.p.data_ptr(), _lib.iptr(its), _lib.iptr(conv),
            _lib.iptr(kry), _lib.iptr(status)), "tcad_solve_bias")
        return its, conv, kry, status

    def predict(self, prev, ratios):
        """Secant predictor of every device's next Newton start
        (tcad_predict_state): prev = (psi, n, p) planes holding each
        device's state at its previous converged point, the present state in
        self.psi/n/p; ratios[d] = step to the next point / step from the
        previous one (<= 0: device d is not touched)."""
        r = np.ascontiguousarray(ratios, dtype=np.float64)
        if r.shape != (self.ndev,):
            raise ValueError(f"one ratio per device ({self.ndev})")
        _lib.check(_lib.lib().tcad_predict_state(
            self._handle, prev[0].data_ptr(), prev[1].data_ptr(), prev[2].data_ptr(),
            self.psi.data_ptr(), self.n.data_ptr(), self.p.data_ptr(), _lib.dptr(r)),
            "tcad_predict_state")

    def solve_queue(self, opts, point_lists):
        """Every device's list of bias points (dicts of terminal voltages,
        each point keeping the terminals it does not name at their previous
        values) solved in one tcad_solve_queue call: a device goes on to
        its next point as soon as it converges at the present one, and stops
        at its first failure with its last converged state restored.
        Returns per device (points converged, Newton iterations per point,
        Krylov iterations per point, outcome code of the point it stopped at,
        currents per converged point as {contact: I})."""
        nd = self.ndev
        if len(point_lists) != nd:
            raise ValueError(f"one point list per device ({nd})")
        npmax = max((len(pl) for pl in point_lists), default=0)
        if npmax == 0:
            return [(0, [], [], _lib.NEWTON_OK, []) for _ in range(nd)]
        nc, ng = len(self.contacts), len(self.gates)
        V = np.zeros((npmax, max(nc, 1)))
        G = np.zeros((npmax, max(ng, 1)))
        cidx = {dn: i for i, dn in enumerate(self.contacts)}
        gidx = {dn: i for i, dn in enumerate(self.gates)}
        for d, (s, pl) in enumerate(zip(self.setups, point_lists)):
            vals = {nm: (bc.V if hasattr(bc, "V") else bc.Vg) for nm, bc in s.bcs.items()}
            for k, pt in enumerate(pl):
                vals.update({nm: float(v) for nm, v in pt.items()})
                for nm, v in vals.items():
                    if (d, nm) in cidx:
                        V[k, cidx[(d, nm)]] = v
                    elif (d, nm) in gidx:
                        G[k, gidx[(d, nm)]] = v
        npts = np.array([len(pl) for pl in point_lists], dtype=np.int32)
        done = np.zeros(nd, dtype=np.int32)
        its = np.zeros(nd * npmax, dtype=np.int32)
        kry = np.zeros(nd * npmax, dtype=np.int32)
        status = np.zeros(nd, dtype=np.int32)
        I = np.zeros(npmax * max(nc, 1))
        self.set_bias()
        opts.apply_tuning(self._handle)
        V = np.ascontiguousarray(V)
        G = np.ascontiguousarray(G)
        _lib.check(_lib.lib().tcad_solve_queue(
            self._handle, opts.max_iter, opts.tol_update, opts.max_dpsi, opts.lin_mode(),
            opts.linsolve_rtol, opts.lin_maxiter, npmax, _lib.iptr(npts), _lib.dptr(V.ravel()),
            _lib.dptr(G.ravel()), self.psi.data_ptr(), self.n.data_ptr(), self.p.data_ptr(),
            _lib.iptr(done), _lib.iptr(its), _lib.iptr(kry), _lib.iptr(status), _lib.dptr(I)),
            "tcad_solve_queue")
        I = I.reshape(npmax, max(nc, 1))
        out = []
        for d in range(nd):
            cur = [{nm: float(I[k, cidx[(d, nm)]]) for dd, nm in self.contacts if dd == d}
                   for k in range(int(done[d]))]
            k1 = min(int(done[d]) + 1, len(point_lists[d]))
            out.append((int(done[d]), its[d * npmax:d * npmax + k1].tolist(),
                        kry[d * npmax:d * npmax + k1].tolist(), int(status[d]), cur))
        return out

    def terminal_currents(self):
        out = np.zeros(max(1, len(self.contacts)), dtype=np.float64)
        _lib.check(_lib.lib().tcad_terminal_currents(
            self._handle, self.psi.data_ptr(), self.n.data_ptr(), self.p.data_ptr(),
            _lib.dptr(out)), "tcad_terminal_currents")
        return out[:len(self.contacts)]

    def y_parameters(self, freqs, rtol, maxiter, active=None):
        """Small-signal Y(f) of the selected devices at their current
        state and bias: (Y [ndev, nfreq, P, P] complex in each device's
        port order -- its contacts, then its gates -- with P the largest
        port count, per-device outcome codes, Krylov iterations)."""
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        P = max(len(s.contacts()) + len(s.gates()) for s in self.setups)
        Y = np.zeros((self.ndev, len(f), P, P, 2), dtype=np.float64)
        status = np.zeros(self.ndev, dtype=np.int32)
        iters = np.zeros(self.ndev, dtype=np.int64)
        m = self._mask(active)
        self.set_bias()
        _lib.check(_lib.lib().tcad_y_parameters(
            self._handle, _lib.iptr(m), len(f), _lib.dptr(f), float(rtol), int(maxiter),
            self.psi.data_ptr(), self.n.data_ptr(), self.p.data_ptr(), _lib.dptr(Y),
            _lib.iptr(status), iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
            "tcad_y_parameters")
        return Y[..., 0] + 1j * Y[..., 1], status, iters

    def device_slice(self, d):
        o = int(self.offsets[d])
        return slice(o, o + self.setups[d].N)


def _raise_outcome(code, what):
    """A failed linear solve raises, as the reference's solve_linear does
    (linsolve.py:973-982); so does a non-finite assembly or step."""
    if code == _lib.NEWTON_LINEAR_FAIL:
        raise _lib.LinearSolveError(f"{what}: the Newton linear solve did not reach a "
                                    "relative residual of 1e-6")
    if code == _lib.NEWTON_NONFINITE:
        raise _lib.LinearSolveError(f"{what}: non-finite or singular Newton system")


class _GPUDevice:
    """Shared implementation of the GPU Device2D / Device3D."""

    def __init__(self, coords, doping, Ntotal=None, T=300.0, material=ph.SILICON,
                 models=None):
        self._s = DeviceSetup(coords, doping, Ntotal, T, material, models)
        for k in ("dim", "shape", "N", "VT", "Ns", "LD", "J0", "R0", "C", "nie_s", "nie",
                  "mu_n0", "mu_p0", "tau_n", "tau_p", "dV", "eps", "models", "doping",
                  "Ntot"):
            setattr(self, k, getattr(self._s, k))
        self.bcs = self._s.bcs
        self._batch = None
        self._solved = False
        self.last_converged = None
        self.last_iterations = None
        self.last_krylov_iterations = None
        self.last_eq_iterations = None

    # -- construction ----------------------------------------------------
    def _invalidate(self):
        self._batch = None
        self._solved = False

    def _gpu(self):
        if self._batch is None:
            self._batch = GPUBatch([self._s])
        return self._batch

    def _set_voltages(self, voltages):
        for name, V in (voltages or {}).items():
            bc = self.bcs[name]
            if isinstance(bc, DirichletBC):
                bc.V = float(V)
            else:
                bc.Vg = float(V)

    # -- solves -------------------------------------------------------------
    def solve_equilibrium(self, opts=None):
        opts = opts or NewtonOptions()
        B = self._gpu()
        its, conv, status = B.solve_equilibrium(opts)
        self.last_eq_iterations = int(its[0])
        _raise_outcome(int(status[0]), "solve_equilibrium")
        if not conv[0]:
            warnings.warn("equilibrium Poisson solve did not converge.")
        self._solved = True
        return self

    def solve_bias(self, voltages=None, opts=None):
        opts = opts or NewtonOptions()
        if not self._solved:
            self.solve_equilibrium(opts)
        self._set_voltages(voltages)
        its, conv, kry, status = self._gpu().solve_bias(opts)
        self.last_iterations = int(its[0])
        self.last_converged = bool(conv[0])
        self.last_krylov_iterations = int(kry[0])
        _raise_outcome(int(status[0]), "solve_bias")
        if not self.last_converged:
            warnings.warn("Newton did not converge")
        return self

    def terminal_current(self, name):
        bc = self.bcs[name]
        if not isinstance(bc, DirichletBC):
            raise ValueError(f"terminal_current: '{name}' is not a Dirichlet contact")
        if not self._solved:
            raise RuntimeError("terminal_current: solve the device first")
        B = self._gpu()
        I = B.terminal_currents()
        return float(I[[nm for _, nm in B.contacts].index(name)])

    def terminal_currents(self):
        """All contacts' currents in contact order (one GPU reduction)."""
        B = self._gpu()
        return dict(zip([nm for _, nm in B.contacts], B.terminal_currents()))

    # -- state ---------------------------------------------------------------
    def _state(self, t):
        if not self._solved:
            return None
        return t.cpu().numpy().reshape(self.shape)

    @property
    def psi(self):
        return self._state(self._gpu().psi)

    @property
    def n(self):
        return self._state(self._gpu().n)

    @property
    def p(self):
        return self._state(self._gpu().p)

    def set_state(self, psi, n, p):
        torch = _torch()
        B = self._gpu()
        for t, v in ((B.psi, psi), (B.n, n), (B.p, p)):
            t.copy_(torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).ravel()))
        self._solved = True

    @property
    def psi_V(self):
        return self.psi * self.VT

    @property
    def n_cm3(self):
        return self.n * self.Ns

    @property
    def p_cm3(self):
        return self.p * self.Ns

    def current_density_edges(self):
        """Electron and hole current densities [A/cm^2] on the mesh edges
        at the current state and bias, per axis: {"x": (Jn, Jp), ...}
        with the reference's edge shapes, e.g. (Ny, Nx-1) along x in 2D
        (the Jn_x/Jp_x ... that solve_bias leaves on the device,
        device2d.py:2188-2190, device3d.py:2266-2270)."""
        torch = _torch()
        if not self._solved:
            raise RuntimeError("current_density_edges: solve the device first")
        B = self._gpu()
        B.set_bias()
        N = self.N
        J = torch.zeros(2 * self.dim * N, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().tcad_edge_currents(B.handle, B.psi.data_ptr(), B.n.data_ptr(),
                                                 B.p.data_ptr(), J.data_ptr()),
                   "tcad_edge_currents")
        Jh = J.cpu().numpy().reshape(self.dim, 2, N) * self.J0
        out = {}
        for a, name in enumerate("xyz"[:self.dim]):
            ax = self.dim - 1 - a                    # array axis of this mesh axis
            sl = [slice(None)] * self.dim
            sl[ax] = slice(None, -1)
            out[name] = tuple(Jh[a, c].reshape(self.shape)[tuple(sl)] for c in (0, 1))
        return out

    # -- residual / Jacobian boundary (tests and inspection) -----------------
    def _update_surface_mobility(self, psi):
        """Apply the lagged Lombardi update at `psi` (device2d.py:612-642)."""
        torch = _torch()
        B = self._gpu()
        t = torch.from_numpy(np.ascontiguousarray(psi, dtype=np.float64).ravel()).to("cuda")
        _lib.check(_lib.lib().tcad_surface_mobility(B.handle, t.data_ptr()),
                   "tcad_surface_mobility")

    def _edge_shaped(self, Jh, carrier):
        """Node-shaped edge planes -> the reference's per-axis edge arrays,
        e.g. (Ny, Nx-1) along x in 2D."""
        out = []
        for a in range(self.dim):
            ax = self.dim - 1 - a
            sl = [slice(None)] * self.dim
            sl[ax] = slice(None, -1)
            out.append(Jh[a, carrier].reshape(self.shape)[tuple(sl)])
        return out

    def _residual_jacobian(self, psi, n, p, voltages, theta=None, current_rows_for=None):
        """The reference's operator: device2d.py:1402-1920 returns
        (F, J, Jn_x, Jn_y, Jp_x, Jp_y, F_n, F_p), device3d.py:1485-1978
        (F, J, Jn_x, Jn_y, Jn_z, Jp_x, Jp_y, Jp_z, F_n, F_p).  F is the
        interleaved residual after the contact and gate rows, J the raw
        (not eliminated, not scaled) Jacobian as scipy CSR with unit rows
        at contact unknowns, the edge currents scaled SG fluxes in edge
        shapes, F_n and F_p the raw continuity residuals.  Sets
        `_dirichlet_rows` (the eliminated rows, in the reference's order)
        and, for `current_rows_for` node indices, `_raw_current_rows` =
        (rows, cols, vals) of their raw continuity rows 3k+1, 3k+2 before
        the contact rows replace them (device2d.py:1751-1758).  Assembled
        on the GPU by tcad_residual_jacobian; theta (the energy-balance
        temperature) is not part of this path."""
        if theta is not None:
            raise NotImplementedError("energy balance (theta) is not part of this path")
        torch = _torch()
        self._set_voltages(voltages)
        B = self._gpu()
        B.set_bias()
        N = self.N
        dev = [torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).ravel()).to("cuda")
               for v in (psi, n, p)]
        ne = 9 + 10 * max(s.dim for s in B.setups)
        F = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        Fn = torch.zeros(N, dtype=torch.float64, device="cuda")
        Fp = torch.zeros(N, dtype=torch.float64, device="cuda")
        Jr = torch.zeros(ne * N, dtype=torch.float64, device="cuda")
        L = _lib.lib()
        _lib.check(L.tcad_residual_jacobian(B.handle, dev[0].data_ptr(), dev[1].data_ptr(),
                                            dev[2].data_ptr(), F.data_ptr(), Fn.data_ptr(),
                                            Fp.data_ptr(), Jr.data_ptr()),
                   "tcad_residual_jacobian")
        Jq = torch.zeros(2 * self.dim * N, dtype=torch.float64, device="cuda")
        _lib.check(L.tcad_edge_currents(B.handle, dev[0].data_ptr(), dev[1].data_ptr(),
                                        dev[2].data_ptr(), Jq.data_ptr()), "tcad_edge_currents")
        Jh = Jq.cpu().numpy().reshape(self.dim, 2, N)
        Fh = F.cpu().numpy().reshape(3, N).T.ravel()
        rows, cols, vals = raw_stencil_coo(Jr.cpu().numpy().reshape(ne, N), self._s)
        contact = np.unique(np.concatenate(
            [np.ravel_multi_index(bc.idx, self.shape) for bc in self.bcs.values()
             if isinstance(bc, DirichletBC)] or [np.zeros(0, dtype=np.int64)]))
        if current_rows_for is not None:
            k_req = np.asarray(current_rows_for, dtype=int)
            sel = np.isin(rows, np.concatenate([3 * k_req + 1, 3 * k_req + 2]))
            self._raw_current_rows = (rows[sel].copy(), cols[sel].copy(), vals[sel].copy())
        if self.dim == 3:
            # device3d.py:1910-1922: component blocks, psi rows then n then p
            strip = np.concatenate([3 * contact, 3 * contact + 1, 3 * contact + 2])
        else:
            strip = np.unique(np.concatenate([3 * contact, 3 * contact + 1, 3 * contact + 2]))
        keep = ~np.isin(rows, strip)
        import scipy.sparse as sp
        J = sp.csr_matrix((np.concatenate([vals[keep], np.ones(strip.size)]),
                           (np.concatenate([rows[keep], strip]),
                            np.concatenate([cols[keep], strip]))), shape=(3 * N, 3 * N))
        self._dirichlet_rows = strip
        Fn_h = Fn.cpu().numpy().reshape(self.shape)
        Fp_h = Fp.cpu().numpy().reshape(self.shape)
        return (Fh, J, *self._edge_shaped(Jh, 0), *self._edge_shaped(Jh, 1), Fn_h, Fp_h)

    def eliminated_system(self, psi, n, p, voltages):
        """The Newton system the GPU solves, at (psi, n, p): (F, Jd, Jn,
        Jp, Fnp) with F as in _residual_jacobian, Jd the Dirichlet-
        eliminated Jacobian (what eliminate_csr returns, dirichlet.py:
        161-191) as scipy CSR, per-axis scaled SG edge currents
        (node-shaped, last entry along each axis 0) and the raw
        pre-Dirichlet F_n + F_p per node."""
        torch = _torch()
        self._set_voltages(voltages)
        B = self._gpu()
        B.set_bias()
        N = self.N
        dev = [torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).ravel()).to("cuda")
               for v in (psi, n, p)]
        F = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        Fnp = torch.zeros(N, dtype=torch.float64, device="cuda")
        L = _lib.lib()
        _lib.check(L.tcad_assemble(B.handle, dev[0].data_ptr(), dev[1].data_ptr(),
                                   dev[2].data_ptr(), F.data_ptr(), Fnp.data_ptr()),
                   "tcad_assemble")
        ne = ctypes.c_int32()
        _lib.check(L.tcad_export_jacobian(B.handle, None, None, None, ctypes.byref(ne)),
                   "tcad_export_jacobian")
        blocks = np.zeros(ne.value * N)
        rs = np.zeros(3 * N)
        _lib.check(L.tcad_export_jacobian(B.handle, _lib.dptr(blocks), _lib.dptr(rs), None,
                                          None), "tcad_export_jacobian")
        Jd = stencil_to_csr(blocks.reshape(ne.value, N), rs,
                            np.asarray(n, dtype=float).ravel(),
                            np.asarray(p, dtype=float).ravel(), self._s)
        J = torch.zeros(2 * self.dim * N, dtype=torch.float64, device="cuda")
        _lib.check(L.tcad_edge_currents(B.handle, dev[0].data_ptr(), dev[1].data_ptr(),
                                        dev[2].data_ptr(), J.data_ptr()), "tcad_edge_currents")
        Jh = J.cpu().numpy().reshape(self.dim, 2, N)
        Fh = F.cpu().numpy().reshape(3, N).T.ravel()
        return (Fh, Jd, [Jh[a, 0].reshape(self.shape) for a in range(self.dim)],
                [Jh[a, 1].reshape(self.shape) for a in range(self.dim)],
                Fnp.cpu().numpy().reshape(self.shape))


_PATTERN = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 2)]


def raw_stencil_coo(blocks, setup):
    """COO (rows, cols, vals) of the raw Jacobian rows exported in
    stencil-block form (tcad_residual_jacobian): every node's 3x3 block
    and, per existing neighbour, the five coupling entries; interleaved
    unknown numbering 3k + c."""
    N = setup.N
    idx = np.arange(N).reshape(setup.shape)
    k = np.arange(N)
    rows, cols, vals = [], [], []
    for r in range(3):
        for c in range(3):
            rows.append(3 * k + r); cols.append(3 * k + c); vals.append(blocks[3 * r + c])
    for a in range(setup.dim):
        ax = setup.dim - 1 - a
        for side in (0, 1):
            dirn = 2 * a + side
            sl_self = [slice(None)] * setup.dim
            sl_nb = [slice(None)] * setup.dim
            if side == 0:
                sl_self[ax] = slice(1, None); sl_nb[ax] = slice(None, -1)
            else:
                sl_self[ax] = slice(None, -1); sl_nb[ax] = slice(1, None)
            ks = idx[tuple(sl_self)].ravel()
            kn = idx[tuple(sl_nb)].ravel()
            for m, (r, c) in enumerate(_PATTERN):
                rows.append(3 * ks + r); cols.append(3 * kn + c)
                vals.append(blocks[9 + 5 * dirn + m][ks])
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def stencil_to_csr(blocks, rowscale, n, p, setup):
    """Rebuild the unscaled eliminated Jacobian Jd = R^-1 (A_s) Cc^-1 from
    the exported stencil blocks (interleaved 3k+c numbering)."""
    import scipy.sparse as sp
    N = setup.N
    shape = setup.shape
    idx = np.arange(N).reshape(shape)
    colscale = np.stack([np.ones(N), n, p], axis=0)      # (3, N)
    rows, cols, vals = [], [], []
    k = np.arange(N)
    for r in range(3):
        for c in range(3):
            rows.append(3 * k + r); cols.append(3 * k + c)
            vals.append(blocks[3 * r + c] / rowscale[r * N + k] / colscale[c, k])
    pattern = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 2)]
    for a in range(setup.dim):
        ax = setup.dim - 1 - a
        for side in (0, 1):          # 0: -axis neighbour, 1: +axis neighbour
            dirn = 2 * a + side
            sl_self = [slice(None)] * setup.dim
            sl_nb = [slice(None)] * setup.dim
            if side == 0:
                sl_self[ax] = slice(1, None); sl_nb[ax] = slice(None, -1)
            else:
                sl_self[ax] = slice(None, -1); sl_nb[ax] = slice(1, None)
            ks = idx[tuple(sl_self)].ravel()
            kn = idx[tuple(sl_nb)].ravel()
            for m, (r, c) in enumerate(pattern):
                v = blocks[9 + 5 * dirn + m][ks] / rowscale[r * N + ks] / colscale[c, kn]
                rows.append(3 * ks + r); cols.append(3 * kn + c); vals.append(v)
    rows = np.concatenate(rows); cols = np.concatenate(cols); vals = np.concatenate(vals)
    keep = vals != 0.0
    return sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(3 * N, 3 * N))


class Device2D(_GPUDevice):
    """GPU twin of pytcad.device2d.Device2D on a structured mesh
    (device2d.py:138-453)."""

    def __init__(self, mesh, doping, Ntotal=None, T=300.0, material=ph.SILICON, models=None):
        super().__init__([mesh.x, mesh.y], doping, Ntotal, T, material, models)
        self.mesh = mesh
        self.Ny, self.Nx = self.shape

    def add_contact(self, name, i, j, V=0.0):
        """device2d.py:672-674."""
        i, j = np.broadcast_arrays(np.atleast_1d(np.asarray(i, dtype=int)),
                                   np.atleast_1d(np.asarray(j, dtype=int)))
        self.bcs[name] = DirichletBC((j.copy(), i.copy()), V)
        self._invalidate()
        return self.bcs[name]

    def add_gate(self, name, i, j, tox_cm, Vfb, Vg=0.0):
        """device2d.py:676-680: kappa = eps_ox LD / (eps_si tox)."""
        i, j = np.broadcast_arrays(np.atleast_1d(np.asarray(i, dtype=int)),
                                   np.atleast_1d(np.asarray(j, dtype=int)))
        kappa = ph.EPS_OX_R * ph.EPS0 * self.LD / (self.eps * tox_cm)
        self.bcs[name] = GateBC((j.copy(), i.copy()), kappa, Vfb, Vg, "y")
        self._invalidate()
        return self.bcs[name]


class Device3D(_GPUDevice):
    """GPU twin of pytcad.device3d.Device3D on a structured mesh
    (device3d.py:411-724)."""

    def __init__(self, mesh, doping, Ntotal=None, T=300.0, material=ph.SILICON, models=None):
        super().__init__([mesh.x, mesh.y, mesh.z], doping, Ntotal, T, material, models)
        self.mesh = mesh
        self.Nz, self.Ny, self.Nx = self.shape

    def add_contact(self, name, i, j, k, V=0.0):
        """device3d.py:748-750."""
        i, j, k = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=int))
                                        for a in (i, j, k)))
        self.bcs[name] = DirichletBC((k.copy(), j.copy(), i.copy()), V)
        self._invalidate()
        return self.bcs[name]

    def add_gate(self, name, i, j, k, tox_cm, Vfb, Vg=0.0, normal_axis="z"):
        """device3d.py:752-756."""
        if normal_axis not in ("x", "y", "z"):
            raise ValueError(f"normal_axis must be 'x', 'y', or 'z', got {normal_axis!r}")
        i, j, k = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=int))
                                        for a in (i, j, k)))
        kappa = ph.EPS_OX_R * ph.EPS0 * self.LD / (self.eps * tox_cm)
        self.bcs[name] = GateBC((k.copy(), j.copy(), i.copy()), kappa, Vfb, Vg, normal_axis)
        self._invalidate()
        return self.bcs[name]


def device_from_spec(spec):
    """Build a GPU Device2D/Device3D from a configs.DeviceSpec."""
    from .mesh import Mesh2D, Mesh3D
    mat = ph.SILICON.with_overrides(spec.material)
    models = ph.Models(**spec.models)
    if spec.dim == 2:
        dev = Device2D(Mesh2D(*spec.coords), spec.doping, spec.Ntotal, T=spec.T, material=mat,
                       models=models)
        for c in spec.contacts:
            dev.add_contact(c.name, c.idx[1], c.idx[0], V=c.V)
        for g in spec.gates:
            dev.add_gate(g.name, g.idx[1], g.idx[0], g.tox_cm, g.Vfb, g.Vg)
    else:
        dev = Device3D(Mesh3D(*spec.coords), spec.doping, spec.Ntotal, T=spec.T, material=mat,
                       models=models)
        for c in spec.contacts:
            dev.add_contact(c.name, c.idx[2], c.idx[1], c.idx[0], V=c.V)
        for g in spec.gates:
            dev.add_gate(g.name, g.idx[2], g.idx[1], g.idx[0], g.tox_cm, g.Vfb, g.Vg,
                         normal_axis=g.normal_axis)
    return dev


def setup_from_spec(spec):
    """The DeviceSetup (boundary conditions attached) of a DeviceSpec, for
    building batches."""
    return device_from_spec(spec)._s
