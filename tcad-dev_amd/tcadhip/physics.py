"""Host-side material models and scaling used once per device to build
the per-node / per-edge coefficient arrays that the HIP kernels read.

Everything here runs once at device construction (setup, not the hot
path): the per-Newton-iteration arithmetic (Bernoulli, SRH/Auger,
Lombardi surface mobility, SG fluxes) is in the HIP kernels.

Formulas follow pytcad/pytcad/materials.py (Semiconductor :18-118,
mobility_caughey_thomas :271-290, bandgap_narrowing_slotboom :435-453,
nie_effective :456-465, lifetime_scharfetter :468-471),
pytcad/pytcad/contacts.py:21-37 (ohmic_values),
pytcad/pytcad/constants.py:11-21 and pytcad/pytcad/moscap.py:60-90
(EPS_OX_R, flatband_voltage).
"""
from dataclasses import dataclass, replace

import numpy as np

Q = 1.602176634e-19
KB = 1.380649e-23
KB_EV = 8.617333262e-5
EPS0 = 8.8541878128e-14
D0_REF = 1.0
EPS_OX_R = 3.9
# Lombardi CVT constants, pytcad/pytcad/materials.py:370-387
CVT_B_N = 2.5e8
CVT_B_P = 5.0e7
CVT_DELTA_N = 5.82e14
CVT_DELTA_P = 2.05e14


def thermal_voltage(T):
    return KB * T / Q


@dataclass(frozen=True)
class Semiconductor:
    eps_r: float = 11.7
    chi: float = 4.05
    Eg0: float = 1.1700
    varshni_alpha: float = 4.730e-4
    varshni_beta: float = 636.0
    Nc300: float = 2.86e19
    Nv300: float = 3.10e19
    mu_n_min: float = 92.0
    mu_n_max: float = 1360.0
    mu_n_Nref: float = 1.3e17
    mu_n_alpha: float = 0.91
    mu_n_Texp: float = -2.33
    mu_p_min: float = 47.7
    mu_p_max: float = 495.0
    mu_p_Nref: float = 6.3e16
    mu_p_alpha: float = 0.76
    mu_p_Texp: float = -2.23
    tau_n0: float = 1.0e-5
    tau_p0: float = 3.0e-6
    tau_Nref: float = 5.0e16
    Cn_auger: float = 2.8e-31
    Cp_auger: float = 9.9e-32
    bgn_E0: float = 6.92e-3
    bgn_N0: float = 1.3e17

    def Eg(self, T):
        return self.Eg0 - self.varshni_alpha * T**2 / (T + self.varshni_beta)

    def Nc(self, T):
        return self.Nc300 * (T / 300.0) ** 1.5

    def Nv(self, T):
        return self.Nv300 * (T / 300.0) ** 1.5

    def ni(self, T):
        return np.sqrt(self.Nc(T) * self.Nv(T)) * np.exp(-self.Eg(T) / (2.0 * KB_EV * T))

    def with_overrides(self, overrides):
        return replace(self, **overrides) if overrides else self


SILICON = Semiconductor()


@dataclass(frozen=True)
class Models:
    """Physics flags of pytcad/pytcad/device.py:373-536 honoured by this
    path.  Flags the reference offers but this path does not implement
    are refused by the device constructors rather than ignored."""
    doping_mobility: bool = True
    srh: bool = True
    auger: bool = True
    bgn: bool = True
    surface_mobility: bool = False


def mobility_caughey_thomas(N, mat, T, carrier):
    N = np.maximum(np.asarray(N, dtype=float), 1.0)
    if carrier == "n":
        mu_max = mat.mu_n_max * (T / 300.0) ** mat.mu_n_Texp
        mu_min, Nref, a = mat.mu_n_min, mat.mu_n_Nref, mat.mu_n_alpha
    else:
        mu_max = mat.mu_p_max * (T / 300.0) ** mat.mu_p_Texp
        mu_min, Nref, a = mat.mu_p_min, mat.mu_p_Nref, mat.mu_p_alpha
    return mu_min + (mu_max - mu_min) / (1.0 + (N / Nref) ** a)


def nie_effective(N, mat, T, use_bgn=True):
    ni = mat.ni(T)
    N = np.asarray(N, dtype=float)
    if not use_bgn:
        return np.full_like(N, ni)
    Nc = np.maximum(N, 1.0)
    x = np.log(Nc / mat.bgn_N0)
    dEg = mat.bgn_E0 * (x + np.sqrt(x * x + 0.5) - np.sqrt(0.5))
    dEg = np.where(Nc > mat.bgn_N0, dEg, 0.0)
    return ni * np.exp(dEg / (2.0 * KB_EV * T))


def lifetime_scharfetter(N, tau0, Nref):
    N = np.maximum(np.asarray(N, dtype=float), 1.0)
    return tau0 / (1.0 + N / Nref)


def ohmic_values(C, nie, V, VT):
    """Returns (psi0, n0, p0, ln(n0/nie)); psi0 = V/VT + ln(n0/nie)."""
    C = np.asarray(C, dtype=float)
    nie = np.asarray(nie, dtype=float)
    root = np.sqrt(C * C + 4.0 * nie * nie)
    n0_if_n = 0.5 * (C + root)
    p0_if_p = 0.5 * (-C + root)
    is_n = C >= 0.0
    n0 = np.where(is_n, n0_if_n, nie * nie / np.maximum(p0_if_p, 1e-300))
    p0 = np.where(is_n, nie * nie / np.maximum(n0_if_n, 1e-300), p0_if_p)
    lnr = np.log(n0 / nie)
    return V / VT + lnr, n0, p0, lnr


def flatband_voltage(Nsub, tox_cm, gate="n+poly", T=300.0, mat=SILICON):
    """V_FB = phi_ms for a MOS gate on substrate doping Nsub (Qf = 0)."""
    VT = thermal_voltage(T)
    nie = float(nie_effective(abs(Nsub), mat, T, True))
    Ns = max(abs(float(Nsub)), mat.ni(T))
    psi_b = np.arcsinh((Nsub / Ns) / (2.0 * (nie / Ns)))
    chi, Eg = mat.chi, mat.Eg(T)
    phi_m = {"n+poly": chi, "p+poly": chi + Eg, "Al": 4.10}.get(gate)
    if phi_m is None:
        phi_m = float(gate)
    return phi_m - (chi + 0.5 * Eg - psi_b * VT)
