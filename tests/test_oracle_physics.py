"""Oracle physics pinned by closed forms and high-precision references
(nothing here retypes the oracle's formula)."""
import mpmath
import numpy as np
import pytest

from oracle import physics as ph

mpmath.mp.dps = 50


def _B_exact(x):
    x = mpmath.mpf(float(x))
    return 1 if x == 0 else x / mpmath.expm1(x)


def _dB_exact(x):
    x = mpmath.mpf(float(x))
    if x == 0:
        return mpmath.mpf(-0.5)
    e = mpmath.exp(x)
    return (e - 1 - x * e) / (e - 1) ** 2


@pytest.mark.parametrize("scale", [1e-6, 1e-3, 1.0, 10.0, 100.0, 600.0])
def test_bernoulli_vs_mpmath(scale):
    """The reference's scale sweep (pytcad/tests/test_physics_kernels_accel_parity.py:24-46)."""
    rng = np.random.default_rng(int(scale * 1000) + 1)
    x = np.concatenate([rng.uniform(-1, 1, 200) * scale, [1e-4, -1e-4, 0.0, 9.999e-5]])
    B = ph.bernoulli(x)
    dB = ph.dbernoulli(x)
    for xi, b, db in zip(x, B, dB):
        be = float(_B_exact(xi))
        dbe = float(_dB_exact(xi))
        assert abs(b - be) <= 1e-12 * max(abs(be), 1e-300) + 1e-300
        # the small-|x| series drops the x^4 term: |err| <= x^4/720 + rounding
        assert abs(db - dbe) <= 1e-9 * max(abs(dbe), 1e-300) + abs(xi) ** 4 / 700 + 1e-300


def test_bernoulli_identities():
    x = np.linspace(-50, 50, 2001)
    assert np.allclose(ph.bernoulli(-x), x + ph.bernoulli(x), rtol=1e-12, atol=1e-12)
    assert np.allclose(ph.dbernoulli(-x), -1.0 - ph.dbernoulli(x), rtol=1e-10, atol=1e-10)
    assert ph.bernoulli(np.array([0.0]))[0] == 1.0


def test_recombination_zero_at_equilibrium_and_fd_derivatives():
    rng = np.random.default_rng(3)
    nie = rng.uniform(1e9, 1e11, 50)
    n = rng.lognormal(30, 5, 50)
    p = nie * nie / n
    R, _, _ = ph.recombination(n, p, nie, 1e-7, 3e-7, ph.SILICON, auger=True)
    assert np.all(np.abs(R) <= 1e-12 * (n * p / (1e-7 * n)))
    p2 = rng.lognormal(30, 5, 50)
    R, dRdn, dRdp = ph.recombination(n, p2, nie, 1e-7, 3e-7, ph.SILICON, auger=True)
    for d, arg in ((dRdn, 0), (dRdp, 1)):
        h = 1e-6 * (n if arg == 0 else p2)
        a = (n + h, p2) if arg == 0 else (n, p2 + h)
        b_ = (n - h, p2) if arg == 0 else (n, p2 - h)
        fd = (ph.recombination(*a, nie, 1e-7, 3e-7, ph.SILICON)[0]
              - ph.recombination(*b_, nie, 1e-7, 3e-7, ph.SILICON)[0]) / (2 * h)
        assert np.allclose(d, fd, rtol=1e-5, atol=0)


def test_ohmic_values_neutrality_and_mass_action():
    C = np.array([1.0, 1e-3, -1.0, -3e-4, 1e-9])
    nie = np.full(5, 1e-7)
    psi0, n0, p0 = ph.ohmic_values(C, nie, 0.3, 0.025)
    assert np.allclose(n0 - p0, C, rtol=1e-12, atol=1e-22)
    assert np.allclose(n0 * p0, nie * nie, rtol=1e-12)
    assert np.allclose(psi0, 0.3 / 0.025 + np.arcsinh(C / (2 * nie)), rtol=1e-12)


def test_control_volume_widths_sum():
    h = np.array([0.3, 0.1, 0.7, 0.2])
    assert abs(ph.control_volume_widths(h).sum() - h.sum()) < 1e-15


def test_material_closed_forms():
    si = ph.SILICON
    assert abs(ph.thermal_voltage(300.0) - 0.025851999786435) < 1e-12
    assert abs(si.ni(300.0) - np.sqrt(2.86e19 * 3.10e19)
               * np.exp(-(1.17 - 4.73e-4 * 9e4 / 936) / (2 * 8.617333262e-5 * 300))) < 1e-3
    mu = ph.mobility_caughey_thomas(np.array([1.3e17]), si, 300.0, "n")[0]
    assert abs(mu - (92 + (1360 - 92) / 2)) < 1e-9
    assert ph.bandgap_narrowing_slotboom(np.array([1e16]), si)[0] == 0.0
    assert abs(ph.lifetime_scharfetter(5e16, 1e-5, 5e16) - 0.5e-5) < 1e-20
