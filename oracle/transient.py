"""Transient (theta-scheme) drift-diffusion steps on the structured 2D/3D
devices, restated from pytcad/pytcad/transient2d.py and transient3d.py.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by: golden fixtures written by tests/golden/make_golden_transient.py
from the reference's transient2d.solve_transient (a diode with a voltage
step, backward Euler and Crank-Nicolson, with adaptive dt): times, dt
history, every contact's current per step, Newton iteration counts per
accepted step and the final state must match bit for bit
(tests/test_oracle_transient.py).  The 3D variant differs only in the
early exit when the line search rejects every step (transient3d.py:216)
and in the current unit (J0 LD^2); it is pinned by z-invariance against
the pinned 2D restatement.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from .dd import DirichletBC, equilibrated_direct
from .physics import NewtonOptions


# -- waveforms: pytcad/pytcad/transient.py:45-106 ---------------------------
class Waveform:
    def value(self, t):
        raise NotImplementedError


class StepWaveform(Waveform):
    """transient.py:52-59: v0 for t < t_step, v1 for t >= t_step."""

    def __init__(self, v0, v1, t_step=0.0):
        self.v0, self.v1, self.t_step = float(v0), float(v1), float(t_step)

    def value(self, t):
        return self.v1 if t >= self.t_step else self.v0


class RampWaveform(Waveform):
    """transient.py:62-76: linear from v0 to v1 over [t0, t1]."""

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
    """transient.py:79-92: v_pulse on [t_start, t_start + width)."""

    def __init__(self, v_base, v_pulse, t_start, width):
        if width <= 0.0:
            raise ValueError("PulseWaveform requires width > 0")
        self.v_base, self.v_pulse = float(v_base), float(v_pulse)
        self.t_start, self.width = float(t_start), float(width)

    def value(self, t):
        return self.v_pulse if self.t_start <= t < self.t_start + self.width else self.v_base


class ConstantWaveform(Waveform):
    """transient.py:95-102."""

    def __init__(self, v):
        self.v = float(v)

    def value(self, t):
        return self.v


def as_waveform(v):
    """transient.py:105-106."""
    return v if isinstance(v, Waveform) else ConstantWaveform(v)


class TransientResult:
    """transient2d.py:53-67 / transient3d.py:73-87, plus the Newton
    iteration count of every accepted step (the reference prints it only
    when verbose)."""

    def __init__(self, times, psi_hist, n_hist, p_hist, terminal_current, dt_hist, iterations):
        self.times = np.asarray(times, dtype=float)
        self.psi_hist = np.asarray(psi_hist, dtype=float)
        self.n_hist = np.asarray(n_hist, dtype=float)
        self.p_hist = np.asarray(p_hist, dtype=float)
        self.terminal_current = {k: np.asarray(v, dtype=float)
                                 for k, v in terminal_current.items()}
        self.dt_hist = np.asarray(dt_hist, dtype=float)
        self.iterations = np.asarray(iterations, dtype=int)


def time_scale(device):
    """transient2d.py:49-50: t = t_s * Ns / R0."""
    return device.Ns / device.R0


def non_contact_flat_index(device):
    """transient2d.py:94-101 / transient3d.py:108-115."""
    mask = np.zeros(device.dV.shape, dtype=bool)
    for bc in device.bcs.values():
        if isinstance(bc, DirichletBC):
            mask[bc.idx] = True
    return np.arange(device.N)[~mask.ravel()]


def step_residual_jacobian(device, psi, n, p, voltages_new, n_old, p_old, F_old_n, F_old_p,
                           dV, dt_s, theta, k_free):
    """transient2d.py:104-135 / transient3d.py:118-149: the bias residual
    plus the storage term on the non-contact continuity rows; theta != 1
    mixes in the old-time continuity residual and scales those rows."""
    F_new, J_new, _, _, F_n_raw, F_p_raw = device._residual_jacobian(psi, n, p, voltages_new)
    F = F_new.copy()
    idx_n = 3 * k_free + 1
    idx_p = 3 * k_free + 2
    dVi = dV.ravel()[k_free]
    n_f, p_f = n.ravel()[k_free], p.ravel()[k_free]
    n0_f, p0_f = n_old.ravel()[k_free], p_old.ravel()[k_free]
    if theta != 1.0:
        F[idx_n] = theta * F[idx_n] + (1.0 - theta) * F_old_n.ravel()[k_free]
        F[idx_p] = theta * F[idx_p] + (1.0 - theta) * F_old_p.ravel()[k_free]
        row_scale = np.ones(F.shape[0])
        row_scale[idx_n] = theta
        row_scale[idx_p] = theta
        J = sp.diags(row_scale) @ J_new
    else:
        J = J_new
    F[idx_n] -= dVi / dt_s * (n_f - n0_f)
    F[idx_p] += dVi / dt_s * (p_f - p0_f)
    rows = np.concatenate([idx_n, idx_p])
    vals = np.concatenate([-dVi / dt_s, dVi / dt_s])
    J = J + sp.csr_matrix((vals, (rows, rows)), shape=J.shape)
    return F, J, F_n_raw, F_p_raw


def newton_step(device, psi0, n0, p0, voltages_new, F_old_n, F_old_p, dV, dt_s, theta, k_free,
                opts, equilibrated=False, stats=None):
    """transient2d.py:138-183 / transient3d.py:152-219: Newton on the step
    residual with a direct solve, psi clip, Armijo backtracking on
    0.5 |F|^2 (up to 40 halvings), density clip; err from the clipped
    (not damped) psi update and the damped density updates.  In 3D a step
    whose line search rejects everything ends the Newton loop.
    equilibrated=True (not a reference method) solves the same system
    after column scaling (1, n, p) and row equilibration: the exact step
    the GPU's equilibrated Krylov solve converges to, where the
    reference's SuperLU on the unscaled Jacobian loses digits.
    stats (test instrumentation, not in the reference): a dict that
    receives the merit 0.5 |F|^2 at the step's start ("start") and at the
    state the loop ends on ("end")."""
    psi, n, p = psi0.copy(), n0.copy(), p0.copy()
    shape = psi.shape
    for it in range(opts.max_iter):
        F, J, _, _ = step_residual_jacobian(device, psi, n, p, voltages_new, n0, p0, F_old_n,
                                            F_old_p, dV, dt_s, theta, k_free)
        du = equilibrated_direct(J, -F, n, p) if equilibrated else spsolve(J.tocsc(), -F)
        dpsi = np.clip(du[0::3].reshape(shape), -opts.max_dpsi, opts.max_dpsi)
        dn = du[1::3].reshape(shape)
        dp = du[2::3].reshape(shape)
        base = 0.5 * float(np.dot(F, F))
        if stats is not None:
            stats.setdefault("start", base)
            stats["end"] = base
        lam = 1.0
        for _ in range(40):
            psi_t = psi + lam * dpsi
            n_t = np.clip(n + lam * dn, 0.1 * n, 10.0 * n)
            p_t = np.clip(p + lam * dp, 0.1 * p, 10.0 * p)
            Ft, *_ = step_residual_jacobian(device, psi_t, n_t, p_t, voltages_new, n0, p0,
                                            F_old_n, F_old_p, dV, dt_s, theta, k_free)
            merit = 0.5 * float(np.dot(Ft, Ft))
            if np.isfinite(merit) and merit <= base * (1.0 - 1e-4 * lam):
                break
            lam *= 0.5
        else:
            lam = 0.0
        if stats is not None and lam > 0.0:
            stats["end"] = merit
        n_prev, p_prev = n, p
        psi = psi + lam * dpsi
        n = np.clip(n + lam * dn, 0.1 * n, 10.0 * n)
        p = np.clip(p + lam * dp, 0.1 * p, 10.0 * p)
        rel_n = (np.abs(n - n_prev) / np.maximum(n_prev, 1e-10)).max()
        rel_p = (np.abs(p - p_prev) / np.maximum(p_prev, 1e-10)).max()
        err = max(float(np.abs(dpsi).max()), float(rel_n), float(rel_p))
        if err < opts.tol_update:
            return psi, n, p, True, it + 1
        if device.ndim == 3 and lam == 0.0:
            return psi, n, p, False, it + 1
    return psi, n, p, False, opts.max_iter


def solve_transient(device, waveforms, t_end, dt0, theta=1.0, opts=None, dt_min=None,
                    dt_max=None, growth=1.5, shrink=0.5, output_times=None, equilibrated=False):
    """transient2d.py:186-306 / transient3d.py:234-363: step the device
    from its current state to t_end with adaptive dt (grow by `growth`
    after a step that took <= max(3, max_iter // 4) Newton iterations,
    shrink by `shrink` and retry after a failed step, stop below dt_min),
    landing exactly on output_times; every contact's current after each
    accepted step."""
    opts = opts or NewtonOptions()
    if device.psi is None:
        raise RuntimeError("solve_transient needs an initial condition")
    wf = {name: as_waveform(v) for name, v in waveforms.items()}
    contact_names = [nm for nm, bc in device.bcs.items() if isinstance(bc, DirichletBC)]

    def voltages_at(t):
        return {nm: (wf[nm].value(t) if nm in wf else device.bcs[nm].V) for nm in contact_names}

    dV = device.dV
    k_free = non_contact_flat_index(device)
    t0 = time_scale(device)
    dt_s = dt0 / t0
    dt_min_s = (dt_min if dt_min is not None else dt0 / 1024.0) / t0
    dt_max_s = (dt_max if dt_max is not None else dt0 * 64.0) / t0
    ldpow = device.LD ** (device.ndim - 1)
    t = 0.0
    psi, n, p = device.psi.copy(), device.n.copy(), device.p.copy()
    times, psi_h, n_h, p_h, dt_h, its = [t], [psi.copy()], [n.copy()], [p.copy()], [], []
    currents = {nm: [] for nm in contact_names}
    merits = []

    def record(psi_, n_, p_, volts):
        *_, F_n, F_p = device._residual_jacobian(psi_, n_, p_, volts)
        for nm in contact_names:
            idx = device.bcs[nm].idx
            # (sum * J0) * LD^(dim-1): the reference's evaluation order
            currents[nm].append(float((F_n[idx] + F_p[idx]).sum()) * device.J0 * ldpow)

    record(psi, n, p, voltages_at(0.0))
    out_times = sorted(output_times) if output_times else None
    out_idx = 0
    while t < t_end - 1e-15:
        step = dt_s
        t_new = t + step * t0
        if out_times is not None and out_idx < len(out_times) and t_new > out_times[out_idx]:
            t_new = out_times[out_idx]
            step = (t_new - t) / t0
        if t_new > t_end:
            t_new = t_end
            step = (t_new - t) / t0
        v_new = voltages_at(t_new)
        F_old_n = F_old_p = None
        if theta != 1.0:
            *_, F_old_n, F_old_p = device._residual_jacobian(psi, n, p, voltages_at(t))
        st = {}
        psi1, n1, p1, ok, nit = newton_step(device, psi, n, p, v_new, F_old_n, F_old_p, dV, step,
                                            theta, k_free, opts, equilibrated, st)
        if ok:
            merits.append((st["start"], st["end"]))
            t = t_new
            psi, n, p = psi1, n1, p1
            times.append(t)
            psi_h.append(psi.copy())
            n_h.append(n.copy())
            p_h.append(p.copy())
            dt_h.append(step * t0)
            its.append(nit)
            record(psi, n, p, v_new)
            if out_times is not None and out_idx < len(out_times) \
                    and abs(t - out_times[out_idx]) < 1e-15 * max(1.0, t):
                out_idx += 1
            dt_s = min(step * growth, dt_max_s) if nit <= max(3, opts.max_iter // 4) else step
        else:
            dt_s = step * shrink
            if dt_s < dt_min_s:
                raise RuntimeError(f"solve_transient stalled at t={t:.3e}s: dt below dt_min")
    device.psi, device.n, device.p = psi, n, p
    final = voltages_at(t)
    for nm in contact_names:
        device.bcs[nm].V = final[nm]
    res = TransientResult(times, psi_h, n_h, p_h, currents, dt_h, its)
    # test instrumentation: merit 0.5 |F|^2 at the start and the end of every
    # accepted step
    res.start_merit = np.array([m[0] for m in merits])
    res.end_merit = np.array([m[1] for m in merits])
    return res
