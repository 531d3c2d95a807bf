"""The oracle against the golden fixtures generated from the reference
(tests/golden/make_golden.py for Device2D, tests/golden/make_golden_3d.py
for Device3D).  The oracle reproduces the reference bit for bit on these
inputs, so the bar is np.array_equal."""
import numpy as np
import pytest
import scipy.sparse as sp

from golden.fixtures import load
from oracle_build import oracle_device, oracle_sweep

SWEEPS = ["diode_60x40_sweep.npz", "mosfet_48x32_sweep.npz",
          "diode_24x16_defaults_sweep.npz",
          "trigate_24x18x14_vd005_sweep.npz", "trigate_24x18x14_vd07_sweep.npz",
          "slab_xgate_10x12x9_sweep.npz", "slab_zgate_12x9x10_sweep.npz"]
SNAPSHOTS = ["diode_14x9_fj.npz", "mosfet_16x12_fj.npz", "trigate_12x10x8_fj.npz",
             "slab_xgate_10x12x9_fj.npz", "slab_zgate_12x9x10_fj.npz"]


@pytest.mark.parametrize("name", SWEEPS)
def test_sweep_matches_reference(name):
    spec, d = load(name)
    keep = tuple(int(k.split("_")[1]) for k in d.files
                 if k.startswith("psi_") and k != "psi_eq")
    dev, its, cur, st = oracle_sweep(spec, keep=keep)
    assert dev.last_eq_iterations == int(d["eq_iterations"])
    assert np.array_equal(its, d["newton_iterations"])
    assert np.array_equal(cur, d["currents"])
    for p, (psi, n, pp) in st.items():
        assert np.array_equal(psi, d[f"psi_{p}"])
        assert np.array_equal(n, d[f"n_{p}"])
        assert np.array_equal(pp, d[f"p_{p}"])
    last = len(spec.sweep) - 1
    if spec.dim == 3 and f"Jnx_{last}" in d.files:
        # the physical edge currents the sweep leaves on the device
        for a, ax in enumerate("xyz"):
            assert np.array_equal(dev.Jn[a], d[f"Jn{ax}_{last}"])
            assert np.array_equal(dev.Jp[a], d[f"Jp{ax}_{last}"])


@pytest.mark.parametrize("name", SNAPSHOTS)
def test_residual_jacobian_matches_reference(name):
    spec, d = load(name)
    dev = oracle_device(spec)
    dev.solve_equilibrium()
    if "Feq" in d.files:
        # equilibrium Poisson F, J at the perturbed psi (gates at 0 V)
        Feq, Jeq = dev._residual_jacobian_poisson(d["in_psi"])
        Jr = sp.csr_matrix((d["Jeq_data"], d["Jeq_indices"], d["Jeq_indptr"]), shape=Jeq.shape)
        assert np.array_equal(Feq, d["Feq"])
        assert abs(Jeq - Jr).max() == 0.0
    volts = dict(zip([str(k) for k in d["volt_keys"]], d["volt_vals"]))
    dev._set_voltages(volts)
    if spec.models["surface_mobility"]:
        dev._update_surface_mobility(d["in_psi"])
    contacts = {c.name for c in spec.contacts}
    F, J, Jn, Jp, Fn, Fp = dev._residual_jacobian(
        d["in_psi"], d["in_n"], d["in_p"], {k: v for k, v in volts.items() if k in contacts})
    Jr = sp.csr_matrix((d["J_data"], d["J_indices"], d["J_indptr"]), shape=J.shape)
    assert np.array_equal(F, d["F"])
    assert abs(J - Jr).max() == 0.0
    assert np.array_equal(Fn, d["Fn"]) and np.array_equal(Fp, d["Fp"])
    for a, ax in enumerate("xyz"[:spec.dim]):
        assert np.array_equal(Jn[a], d[f"Jn{ax}"]) and np.array_equal(Jp[a], d[f"Jp{ax}"])
    assert np.array_equal(dev._dirichlet_rows, d["dirichlet_rows"])
