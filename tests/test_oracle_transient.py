"""Transient oracle (oracle/transient.py) pinned against the reference:

* golden fixtures written by tests/golden/make_golden_transient.py from
  the reference's transient2d.solve_transient -- the oracle must give the
  same time steps, dt history, Newton iteration counts, per-step contact
  currents and final state, bit for bit;
* a z-invariant 3D device steps exactly like the pinned 2D oracle
  (psi in every z-plane, I_3D = I_2D * Lz, same steps and counts);
* with the contact voltages held at their DC values the transient stays
  at the DC state (the storage term vanishes at n = n_old, p = p_old).
"""
import warnings

import numpy as np
import pytest

from golden.fixtures import load
from oracle import dd
from oracle import physics as oph
from oracle import transient as otr
from oracle_build import oracle_device

FIXTURES = ["transient_diode_be.npz", "transient_diode_cn.npz", "transient_mosfet_be.npz"]


def waveforms_from(d):
    """Rebuild the fixture's waveform from its description."""
    desc = [str(v) for v in d["wf_desc"]]
    name, kind, vals = desc[0], desc[1], [float(v) for v in desc[2:]]
    if kind == "step":
        return {name: otr.StepWaveform(*vals)}
    if kind == "ramp":
        return {name: otr.RampWaveform(*vals)}
    raise ValueError(kind)


def run_oracle(name):
    spec, d = load(name)
    dev = oracle_device(spec)
    opts = oph.NewtonOptions()
    dev.solve_equilibrium(opts)
    pre = {str(k): float(v) for k, v in zip(d["prebias_keys"], d["prebias_vals"])}
    if pre:
        dev.solve_bias(pre, opts)
    out_t = [float(v) for v in d["output_times"]] or None
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        res = otr.solve_transient(dev, waveforms_from(d), float(d["t_end"]), float(d["dt0"]),
                                  theta=float(d["theta"]), opts=opts, output_times=out_t)
    return spec, d, dev, res


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_transient_matches_reference(name):
    spec, d, dev, res = run_oracle(name)
    assert np.array_equal(res.times, d["times"])
    assert np.array_equal(res.dt_hist, d["dt_hist"])
    assert np.array_equal(res.iterations, d["iterations"])
    cur = np.array([res.terminal_current[c.name] for c in spec.contacts]).T
    assert np.array_equal(cur, d["currents"])
    assert np.array_equal(res.psi_hist[-1], d["psi_final"])
    assert np.array_equal(res.n_hist[-1], d["n_final"])
    assert np.array_equal(res.p_hist[-1], d["p_final"])


def _diode2d(nx=16, ny=6):
    x = np.linspace(0, 1e-4, nx)
    y = np.linspace(0, 0.3e-4, ny)
    dop = np.tile(np.where(x < 0.45e-4, -2e17, 5e17), (ny, 1))
    m = oph.Models(doping_mobility=True, srh=True, auger=True, bgn=True)
    d = dd.Device2D(x, y, dop, models=m)
    d.add_contact("a", 0, np.arange(ny))
    d.add_contact("c", nx - 1, np.arange(ny))
    return d, x, y, dop, m


def test_transient_3d_z_invariant_equals_2d():
    d2, x, y, dop, m = _diode2d()
    nz = 3
    z = np.linspace(0, 0.2e-4, nz)
    d3 = dd.Device3D(x, y, z, np.tile(dop, (nz, 1, 1)), models=m)
    J, K = np.meshgrid(np.arange(y.size), np.arange(nz), indexing="ij")
    d3.add_contact("a", 0, J.ravel(), K.ravel())
    d3.add_contact("c", x.size - 1, J.ravel(), K.ravel())
    res = []
    for dev in (d2, d3):
        dev.solve_equilibrium()
        # a small step: the line search's merit 0.5 |F|^2 weights the PDE
        # rows by the z box widths, so z-invariance of the trajectory holds
        # only where the full (unclipped) Newton step is accepted
        res.append(otr.solve_transient(dev, {"a": otr.StepWaveform(0.0, 0.05, 0.0)}, 2e-11,
                                       1e-12))
    r2, r3 = res
    assert np.array_equal(r2.times, r3.times)
    # the update test err < 1e-8 can fall either side in the last
    # iteration with the 3D rounding: counts within +-1
    assert np.all(np.abs(r2.iterations - r3.iterations) <= 1)
    for k in range(nz):
        assert np.abs(r3.psi_hist[-1][k] - r2.psi_hist[-1]).max() <= 1e-9
    Ia2, Ic2 = r2.terminal_current["a"], r2.terminal_current["c"]
    floor = 10 * np.abs(Ia2 + Ic2) * z[-1]
    assert np.all(np.abs(r3.terminal_current["a"] - Ia2 * z[-1])
                  <= 1e-7 * np.abs(Ia2 * z[-1]) + floor + 1e-30)


def test_transient_at_dc_stays_put():
    d2, *_ = _diode2d()
    d2.solve_equilibrium()
    d2.solve_bias({"a": 0.3})
    psi0, n0 = d2.psi.copy(), d2.n.copy()
    I0 = d2.terminal_current("a")
    res = otr.solve_transient(d2, {"a": 0.3}, 1e-11, 1e-12)
    assert len(res.times) > 3
    assert np.abs(res.psi_hist[-1] - psi0).max() <= 1e-9
    assert (np.abs(res.n_hist[-1] - n0) / n0).max() <= 1e-7
    assert np.allclose(res.terminal_current["a"], I0, rtol=1e-6,
                       atol=10 * abs(I0 + d2.terminal_current("c")))
