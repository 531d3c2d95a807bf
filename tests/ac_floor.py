"""Rounding floor of each small-signal admittance entry (test helper).

An ohmic row of Y is the sum over the contact's raw current rows of
S_i[col] du_k[col] (ac2d.py:316-319).  At a heavily doped contact the
coefficients S_i are up to 1e11 times the result (a p+ anode's
hole-current sensitivities to potential cancel between the node and its
neighbours), so an FP64 solution du_k, correct to a few eps of its largest
component, moves the entry by a few eps * sum|S_i| * max|du_k|: the
reference's direct solve lands within 1e-14 of that scale of the exact
value, which is fixed by current conservation there.  floor[f, i, k] is
sum|S_i| * max|du_k| in A/V, from the oracle's own direct solves; a gate
row uses the same bound on its displacement-current terms.
"""
import numpy as np
from scipy.sparse.linalg import splu

from oracle import ac as oac
from oracle.dd import DirichletBC, GateBC


def observation_floor(device, freqs):
    N = device.N
    volts = {nm: bc.V for nm, bc in device.bcs.items() if isinstance(bc, DirichletBC)}
    _, J0, *_ = device._residual_jacobian(device.psi, device.n, device.p, volts)
    C = oac.storage_matrix(device)
    t0 = oac.time_scale(device)
    names = list(device.bcs)
    nodes = [np.ravel_multi_index(device.bcs[nm].idx, device.shape).astype(int) for nm in names]
    ohmic = [i for i, nm in enumerate(names) if isinstance(device.bcs[nm], DirichletBC)]
    S = dict(zip(ohmic, oac.ohmic_current_sensitivities(
        device, device.psi, device.n, device.p, volts, [nodes[i] for i in ohmic])))
    b = []
    w = {}
    for i, nm in enumerate(names):
        bc = device.bcs[nm]
        v = np.zeros(3 * N, dtype=complex)
        if isinstance(bc, DirichletBC):
            v[3 * nodes[i]] = 1.0
        else:
            assert isinstance(bc, GateBC)
            w[i] = bc.kappa * device._gate_face_weight(bc)
            v[3 * nodes[i]] = -w[i]
        b.append(v)
    ldpow = device.LD if device.ndim == 2 else device.LD ** 2
    scale = device.J0 * ldpow / device.VT
    P = len(names)
    out = np.zeros((len(freqs), P, P))
    for f, fr in enumerate(freqs):
        om = 2.0 * np.pi * fr * t0
        lu = splu((J0.tocsr().astype(complex) + 1j * om * C).tocsc())
        for k in range(P):
            du = lu.solve(b[k])
            for i in range(P):
                big = np.abs(du).max()
                if i in S:
                    out[f, i, k] = np.abs(S[i]).sum() * big * scale
                else:
                    out[f, i, k] = om * np.abs(w[i]).sum() * max(1.0, big) * scale
    return out
