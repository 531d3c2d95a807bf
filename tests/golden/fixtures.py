"""Load the committed golden fixtures (written by make_golden.py and
make_golden_3d.py from the reference) back into DeviceSpec form, so
tests build oracle and GPU devices from exactly the arrays the reference
was run on.  A 3D fixture carries a z axis, (k, j, i) node lists and the
gates' normal axes; a 2D one (j, i) node lists and y-normal gates."""
import os

import numpy as np

from tcadhip.configs import ContactSpec, DeviceSpec, GateSpec

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_KEYS = ("doping_mobility", "srh", "auger", "bgn", "surface_mobility")


def load(name):
    d = np.load(os.path.join(HERE, name), allow_pickle=False)
    three = "z" in d.files
    axes_idx = ("k", "j", "i") if three else ("j", "i")

    def idx(prefix):
        return tuple(d[f"{prefix}_{a}"] for a in axes_idx)

    contacts = [ContactSpec(str(nm), idx(f"contact_{nm}")) for nm in d["contact_names"]]
    normals = ([str(a) for a in d["gate_axes"]] if three
               else ["y"] * len(d["gate_names"]))
    gates = [GateSpec(str(nm), idx(f"gate_{nm}"), float(par[0]), float(par[1]), ax)
             for nm, par, ax in zip(d["gate_names"], d["gate_params"], normals)]
    keys = [str(k) for k in d["sweep_keys"]]
    sweep = [{k: float(v) for k, v in zip(keys, row) if not np.isnan(v)}
             for row in d["sweep_vals"]]
    ntot = d["Ntotal"] if d["Ntotal"].size else None
    coords = [d["x"], d["y"], d["z"]] if three else [d["x"], d["y"]]
    spec = DeviceSpec(coords, d["doping"], ntot,
                      dict(zip(MODEL_KEYS, (bool(v) for v in d["models"]))),
                      {str(k): float(v) for k, v in zip(d["material_keys"], d["material_vals"])},
                      contacts, gates, sweep, name)
    return spec, d
