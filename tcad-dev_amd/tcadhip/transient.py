"""Transient drift-diffusion on the GPU: the theta-scheme time stepping of
pytcad/pytcad/transient2d.py and transient3d.py, driving one Newton step
per time step through tcad_transient_step.

Same contract as the reference's solve_transient: the device steps from
its current state to t_end; each contact follows its waveform (a contact
not named keeps its voltage; gates are not time-varying); dt grows by
`growth` after a step that took <= max(3, max_iter // 4) Newton
iterations, shrinks by `shrink` and retries after a failed step, and the
run raises once dt falls below dt_min; output_times are landed on
exactly; every contact's current (the raw F_n + F_p at its nodes, the
conduction current) is recorded after each accepted step.  The state,
the old-time copies and the old residuals stay on the GPU; only the
contact voltages go up and the currents come back per step.
"""
import numpy as np

from . import _lib
from .device import DirichletBC, NewtonOptions, _torch


class Waveform:
    """Base class: value(t) -> bias [V] at time t [s]
    (pytcad/pytcad/transient.py:45-50)."""

    def value(self, t):
        raise NotImplementedError


class StepWaveform(Waveform):
    """v0 for t < t_step, v1 from t_step on (transient.py:52-59)."""

    def __init__(self, v0, v1, t_step=0.0):
        self.v0, self.v1, self.t_step = float(v0), float(v1), float(t_step)

    def value(self, t):
        return self.v1 if t >= self.t_step else self.v0


class RampWaveform(Waveform):
    """Linear from v0 at t0 to v1 at t1, constant outside (transient.py:62-76)."""

    def __init__(self, v0, v1, t0, t1):
        if t1 <= t0:
            raise ValueError("RampWaveform requires t1 > t0")
        self.v0, self.v1, self.t0, self.t1 = float(v0), float(v1), float(t0), float(t1)

    def value(self, t):
        if t <= self.t0:
            return self.v0
        if t >= self.t1:
            return self.v1
        return self.v0 + (t - self.t0) / (self.t1 - self.t0) * (self.v1 - self.v0)


class PulseWaveform(Waveform):
    """v_pulse on [t_start, t_start + width), v_base elsewhere (transient.py:79-92)."""

    def __init__(self, v_base, v_pulse, t_start, width):
        if width <= 0.0:
            raise ValueError("PulseWaveform requires width > 0")
        self.v_base, self.v_pulse = float(v_base), float(v_pulse)
        self.t_start, self.width = float(t_start), float(width)

    def value(self, t):
        return self.v_pulse if self.t_start <= t < self.t_start + self.width else self.v_base


class ConstantWaveform(Waveform):
    """A fixed bias (transient.py:95-102)."""

    def __init__(self, v):
        self.v = float(v)

    def value(self, t):
        return self.v


def as_waveform(v):
    return v if isinstance(v, Waveform) else ConstantWaveform(v)


class TransientResult:
    """times [s]; *_hist: one snapshot per accepted step (row 0 the
    initial state); terminal_current {contact: [...]} in the device's
    current unit (A/cm in 2D, A in 3D); dt_hist [s]; iterations: Newton
    iterations of each accepted step (transient2d.py:53-67)."""

    def __init__(self, times, psi_hist, n_hist, p_hist, terminal_current, dt_hist, iterations):
        self.times = np.asarray(times, dtype=float)
        self.psi_hist = np.asarray(psi_hist, dtype=float)
        self.n_hist = np.asarray(n_hist, dtype=float)
        self.p_hist = np.asarray(p_hist, dtype=float)
        self.terminal_current = {k: np.asarray(v, dtype=float)
                                 for k, v in terminal_current.items()}
        self.dt_hist = np.asarray(dt_hist, dtype=float)
        self.iterations = np.asarray(iterations, dtype=int)


def solve_transient(device, waveforms, t_end, dt0, theta=1.0, opts=None, dt_min=None,
                    dt_max=None, growth=1.5, shrink=0.5, output_times=None, keep_history=True):
    """Step a tcadhip Device2D/Device3D from its current state to t_end
    (transient2d.py:186-306, transient3d.py:234-363)."""
    torch = _torch()
    opts = opts or NewtonOptions()
    if not device._solved:
        raise RuntimeError("solve_transient needs an initial condition: call "
                           "solve_equilibrium() or solve_bias() first")
    B = device._gpu()
    L = _lib.lib()
    wf = {name: as_waveform(v) for name, v in waveforms.items()}
    contact_names = [nm for nm, bc in device.bcs.items() if isinstance(bc, DirichletBC)]
    for nm in wf:
        if nm not in contact_names:
            raise ValueError(f"solve_transient: '{nm}' is not a Dirichlet contact")

    def set_voltages(t):
        for nm in contact_names:
            if nm in wf:
                device.bcs[nm].V = wf[nm].value(t)
        B.set_bias()

    t0 = device.Ns / device.R0
    dt_s = dt0 / t0
    dt_min_s = (dt_min if dt_min is not None else dt0 / 1024.0) / t0
    dt_max_s = (dt_max if dt_max is not None else dt0 * 64.0) / t0
    order = [nm for _, nm in B.contacts]
    idx = [order.index(nm) for nm in contact_names]
    psi, n, p = B.psi, B.n, B.p
    n_old, p_old = torch.empty_like(n), torch.empty_like(p)
    Fn_old = Fp_old = None
    if theta != 1.0:
        Fn_old, Fp_old = torch.empty_like(n), torch.empty_like(p)
    save = (psi.clone(), n.clone(), p.clone())

    def snapshot():
        return [t.cpu().numpy().reshape(device.shape) for t in (psi, n, p)]

    t = 0.0
    set_voltages(0.0)
    times, dt_hist, its = [t], [], []
    hist = [snapshot()] if keep_history else []
    currents = {nm: [] for nm in contact_names}

    def record():
        I = B.terminal_currents()
        for nm, k in zip(contact_names, idx):
            currents[nm].append(float(I[k]))

    record()
    out_times = sorted(output_times) if output_times else None
    out_idx = 0
    mode = opts.lin_mode()
    cnt = (np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32))
    while t < t_end - 1e-15:
        step = dt_s
        t_new = t + step * t0
        if out_times is not None and out_idx < len(out_times) and t_new > out_times[out_idx]:
            t_new = out_times[out_idx]
            step = (t_new - t) / t0
        if t_new > t_end:
            t_new = t_end
            step = (t_new - t) / t0
        if theta != 1.0:
            set_voltages(t)
            _lib.check(L.tcad_continuity_residuals(B.handle, psi.data_ptr(), n.data_ptr(),
                                                   p.data_ptr(), Fn_old.data_ptr(),
                                                   Fp_old.data_ptr()),
                       "tcad_continuity_residuals")
        n_old.copy_(n)
        p_old.copy_(p)
        for dst, src in zip(save, (psi, n, p)):
            dst.copy_(src)
        set_voltages(t_new)
        dt_arr = np.array([step], dtype=np.float64)
        _lib.check(L.tcad_transient_step(
            B.handle, opts.max_iter, opts.tol_update, opts.max_dpsi, mode, opts.linsolve_rtol,
            opts.lin_maxiter, None, _lib.dptr(dt_arr), float(theta), n_old.data_ptr(),
            p_old.data_ptr(), None if Fn_old is None else Fn_old.data_ptr(),
            None if Fp_old is None else Fp_old.data_ptr(), psi.data_ptr(), n.data_ptr(),
            p.data_ptr(), _lib.iptr(cnt[0]), _lib.iptr(cnt[1]), _lib.iptr(cnt[2])),
            "tcad_transient_step")
        nit, ok, status = int(cnt[0][0]), bool(cnt[1][0]), int(cnt[2][0])
        if status == _lib.NEWTON_LINEAR_FAIL:
            # the reference's _newton_step works on copies (transient2d.py:139):
            # a failed step leaves the device at the start of the step
            for dst, src in zip((psi, n, p), save):
                dst.copy_(src)
            set_voltages(t)
            raise _lib.LinearSolveError("solve_transient: the Newton linear solve did not "
                                        "reach a relative residual of 1e-6")
        if ok:
            t = t_new
            times.append(t)
            dt_hist.append(step * t0)
            its.append(nit)
            record()
            if keep_history:
                hist.append(snapshot())
            if out_times is not None and out_idx < len(out_times) \
                    and abs(t - out_times[out_idx]) < 1e-15 * max(1.0, t):
                out_idx += 1
            dt_s = min(step * growth, dt_max_s) if nit <= max(3, opts.max_iter // 4) else step
        else:
            for dst, src in zip((psi, n, p), save):
                dst.copy_(src)
            dt_s = step * shrink
            if dt_s < dt_min_s:
                set_voltages(t)
                raise RuntimeError(f"solve_transient stalled at t={t:.3e}s approaching "
                                   f"t={t_new:.3e}s: dt shrank below dt_min="
                                   f"{dt_min_s * t0:.3e}s without Newton convergence")
    set_voltages(t)
    if not keep_history:
        hist = [snapshot()]
    ph, nh, qh = zip(*hist)
    return TransientResult(times, ph, nh, qh, currents, dt_hist, its)
