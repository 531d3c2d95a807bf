"""Generate the small-signal AC golden fixtures by running the REFERENCE
(pytcad.ac2d.y_parameters on pytcad.device2d.Device2D) in the
development container.  Not run by the test suite; the fixtures it writes
(tests/golden/ac_*.npz) are committed and the reference never travels.

Only the reference pins the complex N-port Y(f) of its formulation
(ohmic forcing/observation through the raw current rows, gate
displacement terms, the storage matrix and the physical scaling).

Usage:  python tests/golden/make_golden_ac.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as mg                              # noqa: E402  (sets up the reference)

from pytcad import ac2d as rac                        # noqa: E402
from pytcad.device import NewtonOptions               # noqa: E402

from tcadhip import configs                           # noqa: E402


def run(spec, bias, freqs):
    dev = mg.ref_device(spec)
    opts = NewtonOptions(linsolve="direct")
    dev.solve_equilibrium(opts)
    if bias:
        dev.solve_bias(bias, opts)
    res = rac.y_parameters(dev, freqs)
    return {"freqs": np.asarray(freqs, dtype=float), "Y_re": res.Y.real, "Y_im": res.Y.imag,
            "port_names": np.array(res.port_names),
            "bias_keys": np.array(sorted(bias or {})),
            "bias_vals": np.array([bias[k] for k in sorted(bias or {})], dtype=float)}


def main():
    spec = configs.diode_2d(nx=24, ny=16, Na=3e18, Nd=2e15, xj_frac=0.4, biases=[0.0],
                            models=dict(doping_mobility=True, srh=True, auger=True,
                                        bgn=True, surface_mobility=False),
                            material={}, name="ac_diode_24x16")
    freqs = [1e3, 1e6, 1e9, 1e11]
    mg.save("ac_diode_fwd.npz", spec, run(spec, {"anode": 0.4}, freqs))
    mg.save("ac_diode_rev.npz", spec, run(spec, {"anode": -1.0}, freqs))
    spec = configs.mosfet_2d(nx=24, ny=16, vg_list=[0.8], Vd=0.05, name="ac_mosfet_24x16")
    mg.save("ac_mosfet.npz", spec, run(spec, {"gate": 0.8, "drain": 0.05},
                                       [1e6, 1e9, 1e11, 1e12]))


if __name__ == "__main__":
    main()
