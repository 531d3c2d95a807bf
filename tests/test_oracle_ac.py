"""AC oracle (oracle/ac.py) pinned against the reference:

* golden fixtures written by tests/golden/make_golden_ac.py from the
  reference's ac2d.y_parameters (forward- and reverse-biased diode, a
  MOSFET with its gate port): the complex Y(f) must match bit for bit;
* DC limit: at low frequency the ohmic block of Y is the DC small-signal
  conductance, checked against a finite difference of two terminal-current
  solves;
* a z-invariant 3D device has the 2D Y scaled by the depth.
"""
import warnings

import numpy as np
import pytest

from golden.fixtures import load
from oracle import ac as oac
from oracle import dd
from oracle import physics as oph
from oracle_build import oracle_device

FIXTURES = ["ac_diode_fwd.npz", "ac_diode_rev.npz", "ac_mosfet.npz"]


def dc_device(name):
    spec, d = load(name)
    dev = oracle_device(spec)
    opts = oph.NewtonOptions()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev.solve_equilibrium(opts)
        bias = {str(k): float(v) for k, v in zip(d["bias_keys"], d["bias_vals"])}
        if bias:
            dev.solve_bias(bias, opts)
    return spec, d, dev


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_ac_matches_reference(name):
    spec, d, dev = dc_device(name)
    res = oac.y_parameters(dev, d["freqs"])
    assert list(res.port_names) == [str(p) for p in d["port_names"]]
    assert np.array_equal(res.Y.real, d["Y_re"])
    assert np.array_equal(res.Y.imag, d["Y_im"])


def test_low_frequency_conductance_is_dc_derivative():
    spec, d, dev = dc_device("ac_diode_fwd.npz")
    res = oac.y_parameters(dev, [1.0])
    G = res.Y[0, 0, 0].real                      # dI_anode / dV_anode
    V = float(d["bias_vals"][list(d["bias_keys"]).index("anode")])
    dV = 1e-4
    I = []
    for v in (V - dV, V + dV):
        dev.solve_bias({"anode": v}, oph.NewtonOptions())
        I.append(dev.terminal_current("anode"))
    G_fd = (I[1] - I[0]) / (2 * dV)
    assert abs(G - G_fd) <= 1e-4 * abs(G_fd)
    assert abs(res.Y[0, 0, 0].imag) <= 1e-6 * abs(G)


def test_ac_3d_z_invariant_equals_2d():
    x = np.linspace(0, 1e-4, 14)
    y = np.linspace(0, 0.3e-4, 6)
    dop = np.tile(np.where(x < 0.45e-4, -2e17, 5e17), (6, 1))
    m = oph.Models(doping_mobility=True, srh=True, auger=True, bgn=True)
    d2 = dd.Device2D(x, y, dop, models=m)
    d2.add_contact("a", 0, np.arange(6))
    d2.add_contact("c", 13, np.arange(6))
    nz = 3
    z = np.linspace(0, 0.2e-4, nz)
    d3 = dd.Device3D(x, y, z, np.tile(dop, (nz, 1, 1)), models=m)
    J, K = np.meshgrid(np.arange(6), np.arange(nz), indexing="ij")
    d3.add_contact("a", 0, J.ravel(), K.ravel())
    d3.add_contact("c", 13, J.ravel(), K.ravel())
    for dv in (d2, d3):
        dv.solve_equilibrium()
        dv.solve_bias({"a": 0.3})
    f = [1e6, 1e10]
    Y2 = oac.y_parameters(d2, f).Y
    Y3 = oac.y_parameters(d3, f).Y
    assert np.allclose(Y3, Y2 * z[-1], rtol=1e-6, atol=1e-9 * np.abs(Y2).max() * z[-1])


@pytest.mark.parametrize("name", FIXTURES)
def test_refined_oracle_pinned(name):
    """The extended-precision Y (refine=True) is within the reference's own
    rounding of the fixture, and -- for the ohmic-only diodes, where a
    common voltage on every contact moves no current -- meets the row-sum
    identity sum_k Y[i, k] = 0 to ~1e-17 of the rows' rounding floor,
    which the FP64 direct solve (the reference's) misses by up to 50x."""
    from ac_floor import observation_floor
    spec, d, dev = dc_device(name)
    freqs = d["freqs"]
    Yr = oac.y_parameters(dev, freqs, refine=True).Y
    Yfix = d["Y_re"] + 1j * d["Y_im"]
    floor = observation_floor(dev, freqs)
    for f in range(len(freqs)):
        tol = 1e-6 * np.abs(Yfix[f]).max() + 1e-12 * floor[f]
        assert np.all(np.abs(Yr[f] - Yfix[f]) <= tol), f
    if "gate" in [str(p) for p in d["port_names"]]:
        return
    rowfloor = floor.sum(axis=2)
    assert np.all(np.abs(Yr.sum(axis=2)) <= 3e-17 * rowfloor)
    if name == "ac_diode_rev.npz":
        assert np.any(np.abs(Yfix.sum(axis=2)) > 3e-17 * rowfloor)
