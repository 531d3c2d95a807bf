"""Oracle physics: pointwise FP64 kernels and material parameters.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Each function restates the reference function named in its docstring.
Pinning (tests/test_oracle_physics.py): Bernoulli/dBernoulli against an
mpmath 50-digit evaluation and the identities B(-x) = x + B(x),
B'(-x) = -1 - B'(x), B(0) = 1; recombination against R = 0 at
np = n_ie^2 and a finite-difference derivative check; ohmic values
against n0 - p0 = C and n0 p0 = n_ie^2; the mobility / BGN / lifetime
fits against hand-evaluated closed forms; control-volume widths against
sum(dV) = L.  The whole set is further pinned end-to-end by the golden
fixtures in tests/golden/ (generated from the reference itself).
"""
from dataclasses import dataclass, replace

import numpy as np

# pytcad/pytcad/constants.py:11-14
Q = 1.602176634e-19
KB = 1.380649e-23
KB_EV = 8.617333262e-5
EPS0 = 8.8541878128e-14
# pytcad/pytcad/kernels.py:29
D0_REF = 1.0
# pytcad/pytcad/moscap.py:60
EPS_OX_R = 3.9


def thermal_voltage(T):
    """pytcad/pytcad/constants.py:19-21."""
    return KB * T / Q


@dataclass(frozen=True)
class Semiconductor:
    """Silicon parameter set, pytcad/pytcad/materials.py:18-118 (only the
    fields the drift-diffusion path reads)."""
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
        """materials.py:96-102 (Varshni)."""
        return self.Eg0 - self.varshni_alpha * T**2 / (T + self.varshni_beta)

    def Nc(self, T):
        """materials.py:104-105."""
        return self.Nc300 * (T / 300.0) ** 1.5

    def Nv(self, T):
        """materials.py:107-108."""
        return self.Nv300 * (T / 300.0) ** 1.5

    def ni(self, T):
        """materials.py:110-118."""
        return np.sqrt(self.Nc(T) * self.Nv(T)) * np.exp(
            -self.Eg(T) / (2.0 * KB_EV * T))


SILICON = Semiconductor()


@dataclass(frozen=True)
class Models:
    """The physics flags of pytcad/pytcad/device.py:373-536 that the
    structured drift-diffusion path reads (everything else stays off)."""
    doping_mobility: bool = True
    srh: bool = True
    auger: bool = True
    bgn: bool = True
    surface_mobility: bool = False


@dataclass(frozen=True)
class NewtonOptions:
    """pytcad/pytcad/device.py:539-628 (the fields the path reads)."""
    max_iter: int = 100
    tol_update: float = 1e-8
    max_dpsi: float = 5.0
    linsolve: str = "direct"
    linsolve_rtol: float = 1e-10
    block_size: int = 3
    # 3D iterative solves fall back to direct on LinearSolveError
    # (device3d.py:2154-2168); off where a direct solve is out of reach
    direct_fallback: bool = True


def mobility_caughey_thomas(N, mat, T, carrier):
    """materials.py:271-290."""
    N = np.maximum(np.asarray(N, dtype=float), 1.0)
    if carrier == "n":
        mu_max = mat.mu_n_max * (T / 300.0) ** mat.mu_n_Texp
        mu_min, Nref, a = mat.mu_n_min, mat.mu_n_Nref, mat.mu_n_alpha
    else:
        mu_max = mat.mu_p_max * (T / 300.0) ** mat.mu_p_Texp
        mu_min, Nref, a = mat.mu_p_min, mat.mu_p_Nref, mat.mu_p_alpha
    return mu_min + (mu_max - mu_min) / (1.0 + (N / Nref) ** a)


# materials.py:370-387 (Lombardi CVT constants)
CVT_B_N = 2.5e8
CVT_B_P = 5.0e7
CVT_DELTA_N = 5.82e14
CVT_DELTA_P = 2.05e14


def mobility_cvt(E_eff, mu_ct, carrier, T):
    """materials.py:389-429."""
    E_eff = np.maximum(np.asarray(E_eff, dtype=float), 1.0)
    if carrier == "n":
        B, delta = CVT_B_N, CVT_DELTA_N
    else:
        B, delta = CVT_B_P, CVT_DELTA_P
    mu_ph = B / (T * E_eff ** (1.0 / 3.0))
    mu_sr = delta / E_eff ** 2
    inv_mu = (1.0 / np.maximum(mu_ct, 1.0)
              + 1.0 / np.maximum(mu_ph, 1.0)
              + 1.0 / np.maximum(mu_sr, 1.0))
    return 1.0 / inv_mu


def bandgap_narrowing_slotboom(N, mat):
    """materials.py:435-453."""
    N = np.maximum(np.asarray(N, dtype=float), 1.0)
    x = np.log(N / mat.bgn_N0)
    dEg = mat.bgn_E0 * (x + np.sqrt(x * x + 0.5) - np.sqrt(0.5))
    return np.where(N > mat.bgn_N0, dEg, 0.0)


def nie_effective(N, mat, T, use_bgn=True):
    """materials.py:456-465."""
    ni = mat.ni(T)
    if not use_bgn:
        return np.full_like(np.asarray(N, dtype=float), ni)
    dEg = bandgap_narrowing_slotboom(N, mat)
    return ni * np.exp(dEg / (2.0 * KB_EV * T))


def lifetime_scharfetter(N, tau0, Nref):
    """materials.py:468-471."""
    N = np.maximum(np.asarray(N, dtype=float), 1.0)
    return tau0 / (1.0 + N / Nref)


def recombination(n, p, nie, tau_n, tau_p, mat, auger=True):
    """materials.py:474-518, Boltzmann branch (np_eq=None): SRH with
    mid-gap traps plus optional Auger, with exact derivatives."""
    ni2 = nie * nie
    excess = n * p - ni2
    den = tau_p * (n + nie) + tau_n * (p + nie)
    R = excess / den
    dRdn = (p * den - excess * tau_p) / den**2
    dRdp = (n * den - excess * tau_n) / den**2
    if auger:
        C = mat.Cn_auger * n + mat.Cp_auger * p
        R = R + C * excess
        dRdn = dRdn + mat.Cn_auger * excess + C * p
        dRdp = dRdp + mat.Cp_auger * excess + C * n
    return R, dRdn, dRdp


def bernoulli(x):
    """B(x) = x/expm1(x), kernels.py:88-96."""
    x = np.clip(np.asarray(x, dtype=float), -700.0, 700.0)
    small = np.abs(x) < 1e-4
    xs = np.where(small, 1.0, x)
    return np.where(small, 1.0 - x / 2.0 + x * x / 12.0, xs / np.expm1(xs))


def dbernoulli(x):
    """dB/dx, kernels.py:99-109."""
    x = np.clip(np.asarray(x, dtype=float), -700.0, 700.0)
    small = np.abs(x) < 1e-4
    xs = np.where(small, 1.0, x)
    B = np.where(small, 1.0, xs / np.expm1(xs))
    em = -np.expm1(-xs)
    em = np.where(np.abs(em) < 1e-300, 1e-300, em)
    full = B * (1.0 / xs - 1.0 / em)
    series = -0.5 + x / 6.0 - x**3 / 180.0
    return np.where(small, series, full)


def ohmic_values(C, nie, V, VT):
    """contacts.py:21-37: majority carrier from the neutrality quadratic,
    minority from mass action; psi0 = V/VT + ln(n0/nie)."""
    C = np.asarray(C, dtype=float)
    nie = np.asarray(nie, dtype=float)
    root = np.sqrt(C * C + 4.0 * nie * nie)
    n0_if_n = 0.5 * (C + root)
    p0_if_p = 0.5 * (-C + root)
    is_n = C >= 0.0
    n0 = np.where(is_n, n0_if_n, nie * nie / np.maximum(p0_if_p, 1e-300))
    p0 = np.where(is_n, nie * nie / np.maximum(n0_if_n, 1e-300), p0_if_p)
    psi0 = V / VT + np.log(n0 / nie)
    return psi0, n0, p0


def control_volume_widths(h):
    """mesh2d.py:15-28: interior nodes take half of each adjacent
    interval, endpoints one half-interval."""
    h = np.asarray(h, dtype=float)
    dv = np.zeros(h.size + 1)
    dv[1:-1] = 0.5 * (h[:-1] + h[1:])
    dv[0] = 0.5 * h[0]
    dv[-1] = 0.5 * h[-1]
    return dv


def flatband_voltage(Nsub, tox_cm, gate="n+poly", T=300.0, mat=SILICON):
    """moscap.py:63-90 with Qf = 0."""
    VT = thermal_voltage(T)
    ni = mat.ni(T)
    nie = float(nie_effective(abs(Nsub), mat, T, True))
    Ns = max(abs(float(Nsub)), ni)
    C = Nsub / Ns
    nie_s = nie / Ns
    psi_b = np.arcsinh(C / (2.0 * nie_s))
    chi, Eg = mat.chi, mat.Eg(T)
    if gate == "n+poly":
        phi_m = chi
    elif gate == "p+poly":
        phi_m = chi + Eg
    elif gate == "Al":
        phi_m = 4.10
    else:
        phi_m = float(gate)
    phi_semi = chi + 0.5 * Eg - psi_b * VT
    return phi_m - phi_semi


__all__ = [n for n in dir() if not n.startswith("_")] + ["replace"]
