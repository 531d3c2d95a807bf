"""Small-signal AC N-port admittance of the structured 2D/3D devices,
restated from pytcad/pytcad/ac2d.py and ac3d.py.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by: golden fixtures written by tests/golden/make_golden_ac.py from
the reference's ac2d.y_parameters (a diode and a MOSFET with its gate
port, several frequencies): Y must match bit for bit
(tests/test_oracle_ac.py); the 3D restatement differs only in the node
indexing and the area unit (J0 LD^2) and is pinned by z-invariance
against the pinned 2D one.

y_parameters(..., refine=True) is not a reference method: the same
systems with the solution refined in extended precision (80-bit long
double residuals and solution) and the observations summed in extended
precision.  An ohmic entry at a heavily doped contact cancels by up to
1e11, and the reference's FP64 direct solve resolves it only to a few
1e-15 of its terms; the refined Y is the value those entries approach.
Pinned by the exact identities of the discretisation that the FP64 path
only meets to its rounding (tests/test_oracle_ac.py): with ohmic ports
only, raising every contact by the same voltage shifts psi uniformly and
moves no current, so each row of Y sums to zero at every frequency; and by
agreement with the bit-pinned FP64 path on entries that do not cancel.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu

from .dd import DirichletBC, GateBC


def time_scale(device):
    """ac2d.py:79-83."""
    return device.Ns / device.R0


def storage_matrix(device):
    """ac2d.py:107-124 / ac3d.py: the coefficient of j w_s: -dV on the
    electron rows, +dV on the hole rows, zero at ohmic-contact nodes
    (S_n = S_p = 0, ac2d.py:85-104)."""
    N = device.N
    dVf = device.dV.ravel()
    excl = [np.ravel_multi_index(bc.idx, device.shape).astype(int)
            for bc in device.bcs.values() if isinstance(bc, DirichletBC)]
    excl = np.unique(np.concatenate(excl)) if excl else np.array([], dtype=int)
    keep = np.setdiff1d(np.arange(N), excl, assume_unique=False)
    rows = np.concatenate([3 * keep + 1, 3 * keep + 2])
    vals = np.concatenate([-dVf[keep], dVf[keep]])
    return sp.csr_matrix((vals, (rows, rows)), shape=(3 * N, 3 * N))


def ohmic_current_sensitivities(device, psi, n, p, voltages, node_sets):
    """ac2d.py:168-194: per ohmic node set, the sum over its nodes of the
    raw continuity rows 3k+1 and 3k+2 of the Jacobian (one assembly)."""
    N = device.N
    device._residual_jacobian(psi, n, p, voltages,
                              current_rows_for=np.unique(np.concatenate(node_sets)))
    r, c, v = device._raw_current_rows
    del device._raw_current_rows
    in_state = c < 3 * N
    out = []
    for kk in node_sets:
        sel = in_state & np.isin(r // 3, kk)
        S = np.zeros(3 * N)
        np.add.at(S, c[sel], v[sel])
        out.append(S)
    return out


class YParamResult:
    """ac2d.py:249-262: freqs [Hz], Y [A/V] shape (len(freqs), P, P),
    port_names in device.bcs order."""

    def __init__(self, freqs, Y, port_names):
        self.freqs = np.asarray(freqs, dtype=float)
        self.Y = np.asarray(Y, dtype=complex)
        self.port_names = tuple(port_names)


def _refined_solve(A, lu, b, steps):
    """Iterative refinement with extended-precision residuals; returns
    the solution in long double."""
    Acsr = A.tocsr()
    rows = np.repeat(np.arange(Acsr.shape[0]), np.diff(Acsr.indptr))
    data = Acsr.data.astype(np.clongdouble)
    bl = b.astype(np.clongdouble)
    x = lu.solve(b).astype(np.clongdouble)
    for _ in range(steps):
        Ax = np.zeros(Acsr.shape[0], dtype=np.clongdouble)
        np.add.at(Ax, rows, data * x[Acsr.indices])
        r = bl - Ax
        x = x + lu.solve(r.astype(complex)).astype(np.clongdouble)
    return x


def y_parameters(device, freqs, refine=False, refine_steps=4):
    """ac2d.py:265-342 / ac3d.py:187-277: every bc is a port; ohmic
    forcing b[3m] = 1 on the contact's psi rows, gate forcing
    b[3m] = -kappa w; ohmic observation S_i . du_k, gate observation
    j w_s sum kappa w (delta_ik - du_k[3m]); one LU of J + j w_s C per
    frequency; Y * J0 * LD^(dim-1) / VT.  refine=True: see the module
    docstring."""
    if device.psi is None:
        raise RuntimeError("y_parameters needs a converged DC operating point")
    N = device.N
    psi0, n0, p0 = device.psi, device.n, device.p
    voltages = {nm: bc.V for nm, bc in device.bcs.items() if isinstance(bc, DirichletBC)}
    _, J0, *_ = device._residual_jacobian(psi0, n0, p0, voltages)
    Cmat = storage_matrix(device)
    t0 = time_scale(device)
    J0c = J0.tocsr().astype(complex)
    names = list(device.bcs.keys())
    P = len(names)
    b_ports, kinds, node_sets, weights, S_ohmic = [], [], [], [], []
    for nm in names:
        bc = device.bcs[nm]
        kk = np.ravel_multi_index(bc.idx, device.shape).astype(int)
        node_sets.append(kk)
        b = np.zeros(3 * N, dtype=complex)
        if isinstance(bc, DirichletBC):
            kinds.append("ohmic")
            b[3 * kk] = 1.0
            weights.append(None)
        elif isinstance(bc, GateBC):
            kinds.append("gate")
            w = bc.kappa * device._gate_face_weight(bc)
            b[3 * kk] = -w
            weights.append(w)
        else:
            raise TypeError(f"unsupported port type {type(bc).__name__}")
        S_ohmic.append(None)
        b_ports.append(b)
    ohmic = [i for i, kd in enumerate(kinds) if kd == "ohmic"]
    if ohmic:
        for i, S in zip(ohmic, ohmic_current_sensitivities(device, psi0, n0, p0, voltages,
                                                           [node_sets[i] for i in ohmic])):
            S_ohmic[i] = S
    Y = np.empty((len(freqs), P, P), dtype=complex)
    ldpow = device.LD if device.ndim == 2 else device.LD ** 2
    if refine:
        Yl = np.empty((len(freqs), P, P), dtype=np.clongdouble)
        for kf, f in enumerate(freqs):
            omega_s = 2.0 * np.pi * f * t0
            A = (J0c + 1j * omega_s * Cmat).tocsc()
            lu = splu(A)
            du = [_refined_solve(A, lu, b_ports[k], refine_steps) for k in range(P)]
            wl = np.longdouble(omega_s)
            for k in range(P):
                for i in range(P):
                    if kinds[i] == "ohmic":
                        Yl[kf, i, k] = np.sum(S_ohmic[i].astype(np.longdouble) * du[k])
                    else:
                        delta = 1.0 if i == k else 0.0
                        s = np.sum(weights[i].astype(np.longdouble)
                                   * (delta - du[k][3 * node_sets[i]]))
                        Yl[kf, i, k] = 1j * wl * s
        scale = np.longdouble(device.J0) * np.longdouble(ldpow) / np.longdouble(device.VT)
        return YParamResult(freqs, (Yl * scale).astype(complex), names)
    for kf, f in enumerate(freqs):
        omega_s = 2.0 * np.pi * f * t0
        lu = splu((J0c + 1j * omega_s * Cmat).tocsc())
        du = [lu.solve(b_ports[k]) for k in range(P)]
        for k in range(P):
            for i in range(P):
                if kinds[i] == "ohmic":
                    Y[kf, i, k] = S_ohmic[i].astype(complex) @ du[k]
                else:
                    delta = 1.0 if i == k else 0.0
                    Y[kf, i, k] = 1j * omega_s * np.sum(
                        weights[i] * (delta - du[k][3 * node_sets[i]]))
    return YParamResult(freqs, Y * device.J0 * ldpow / device.VT, names)


def cutoff_frequency(yres, port_in, port_out):
    """ac2d.py:345-404: |h21| = |Y[out,in] / Y[in,in]| crossing 1 by
    log-log interpolation; None where the sweep does not bracket it."""
    idx = (lambda q: yres.port_names.index(q) if isinstance(q, str) else int(q))
    i_in, i_out = idx(port_in), idx(port_out)
    freqs = yres.freqs
    Y_in = yres.Y[:, i_in, i_in]
    Y_out = yres.Y[:, i_out, i_in]
    with np.errstate(divide="ignore", invalid="ignore"):
        h21 = np.where(np.abs(Y_in) > 0, Y_out / Y_in, np.inf)
    mag = np.abs(h21)
    if not np.any(np.isfinite(mag)) or mag[0] < 1.0 or mag[-1] >= 1.0:
        return None
    i1 = int(np.where(mag < 1.0)[0][0])
    i0 = i1 - 1
    if i0 < 0:
        return None
    lf0, lf1 = np.log(freqs[i0]), np.log(freqs[i1])
    lm0, lm1 = np.log(mag[i0]), np.log(mag[i1])
    if lm0 == lm1:
        return float(freqs[i0])
    return float(np.exp(lf0 + (0.0 - lm0) / (lm1 - lm0) * (lf1 - lf0)))
