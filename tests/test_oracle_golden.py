"""The oracle against the golden fixtures generated from the reference
(tests/golden/make_golden.py).  The oracle reproduces the reference
bit for bit on these inputs, so the bar is np.array_equal."""
import numpy as np
import pytest
import scipy.sparse as sp

from golden.fixtures import load
from oracle_build import oracle_device, oracle_sweep

SWEEPS = ["diode_60x40_sweep.npz", "mosfet_48x32_sweep.npz",
          "diode_24x16_defaults_sweep.npz"]


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


@pytest.mark.parametrize("name", ["diode_14x9_fj.npz", "mosfet_16x12_fj.npz"])
def test_residual_jacobian_matches_reference(name):
    spec, d = load(name)
    dev = oracle_device(spec)
    dev.solve_equilibrium()
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
    assert np.array_equal(Jn[0], d["Jnx"]) and np.array_equal(Jn[1], d["Jny"])
    assert np.array_equal(Jp[0], d["Jpx"]) and np.array_equal(Jp[1], d["Jpy"])
    assert np.array_equal(dev._dirichlet_rows, d["dirichlet_rows"])
