"""Small-signal AC analysis on the GPU: the N-port admittance Y(f) of
pytcad/pytcad/ac2d.py and ac3d.py, through tcad_y_parameters.

Same contract as the reference's y_parameters: the device must hold a
converged DC operating point (solve_equilibrium / solve_bias first); every
boundary condition is a port, in device.bcs order; Y[f, i, k] = dI_i/dV_k
in A/V (per cm of depth in 2D) at each frequency [Hz]; ohmic rows are
conduction-current responses, gate rows displacement currents
(ac2d.py:265-342).  The reference factors J + j w C per frequency; here
each port's complex system is solved iteratively on the GPU (FP64 complex
BiCGStab, restarted on the true residual down to `rtol` or until it stops
falling), so Y matches the direct solve to the accuracy of that residual.
"""
import numpy as np

from . import _lib
from .device import DirichletBC, GateBC


class YParamResult:
    """freqs [Hz], Y [A/V] of shape (len(freqs), P, P), port_names in
    device.bcs order; iterations: Krylov iterations spent
    (ac2d.py:249-262)."""

    def __init__(self, freqs, Y, port_names, iterations=0):
        self.freqs = np.asarray(freqs, dtype=float)
        self.Y = np.asarray(Y, dtype=complex)
        self.port_names = tuple(port_names)
        self.iterations = int(iterations)

    def __repr__(self):
        return (f"YParamResult(ports={list(self.port_names)}, "
                f"n_freqs={len(self.freqs)}, f=[{self.freqs[0]:.3g}, {self.freqs[-1]:.3g}] Hz)")


def y_parameters(device, freqs, rtol=1e-16, maxiter=20000):
    """N-port Y(f) of a tcadhip Device2D/Device3D at its current state
    (ac2d.py:265-342, ac3d.py:187-277).  rtol: target true residual of
    each complex solve relative to its forcing; the default asks for the
    FP64 floor (sweeps stop once the residual stops falling), since an
    ohmic entry at a heavily doped contact cancels by up to 1e11."""
    if not device._solved:
        raise RuntimeError("y_parameters needs a converged DC operating point: call "
                           "solve_equilibrium() or solve_bias() first")
    for nm, bc in device.bcs.items():
        if not isinstance(bc, (DirichletBC, GateBC)):
            raise TypeError(f"unsupported port type {type(bc).__name__} for '{nm}'")
    freqs = np.atleast_1d(np.asarray(freqs, dtype=float))
    B = device._gpu()
    Y, status, iters = B.y_parameters(freqs, rtol, maxiter)
    if int(status[0]) == _lib.NEWTON_LINEAR_FAIL:
        raise _lib.LinearSolveError("y_parameters: a complex linear solve stalled above a "
                                    "relative residual of 1e-6")
    # the library's port order: contacts, then gates
    lib_order = [nm for nm, _ in B.setups[0].contacts()] + [nm for nm, _ in B.setups[0].gates()]
    names = list(device.bcs.keys())
    perm = [lib_order.index(nm) for nm in names]
    Yd = Y[0][:, perm][:, :, perm]
    return YParamResult(freqs, Yd, names, int(iters[0]))


def cutoff_frequency(yres, port_in, port_out):
    """|h21| = |Y[out, in] / Y[in, in]| crossing 1, by log-log
    interpolation between the bracketing frequencies; None where the sweep
    does not bracket the crossing (ac2d.py:345-404)."""
    def idx(q):
        return yres.port_names.index(q) if isinstance(q, str) else int(q)

    i_in, i_out = idx(port_in), idx(port_out)
    freqs = yres.freqs
    Y_in = yres.Y[:, i_in, i_in]
    Y_out = yres.Y[:, i_out, i_in]
    with np.errstate(divide="ignore", invalid="ignore"):
        h21 = np.where(np.abs(Y_in) > 0, Y_out / Y_in, np.inf)
    mag = np.abs(h21)
    if not np.any(np.isfinite(mag)) or mag[0] < 1.0 or mag[-1] >= 1.0:
        return None
    below = np.nonzero(mag < 1.0)[0]
    i1 = int(below[0])
    if i1 == 0:
        return None
    i0 = i1 - 1
    lf0, lf1 = np.log(freqs[i0]), np.log(freqs[i1])
    lm0, lm1 = np.log(mag[i0]), np.log(mag[i1])
    if lm0 == lm1:
        return float(freqs[i0])
    return float(np.exp(lf0 + (0.0 - lm0) / (lm1 - lm0) * (lf1 - lf0)))
