"""Load the committed golden fixtures (written by make_golden.py from the
reference) back into DeviceSpec form, so tests build oracle and GPU
devices from exactly the arrays the reference was run on."""
import os

import numpy as np

from tcadhip.configs import ContactSpec, DeviceSpec, GateSpec

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_KEYS = ("doping_mobility", "srh", "auger", "bgn", "surface_mobility")


def load(name):
    d = np.load(os.path.join(HERE, name), allow_pickle=False)
    contacts = [ContactSpec(str(nm), (d[f"contact_{nm}_j"], d[f"contact_{nm}_i"]))
                for nm in d["contact_names"]]
    gates = [GateSpec(str(nm), (d[f"gate_{nm}_j"], d[f"gate_{nm}_i"]),
                      float(par[0]), float(par[1]), "y")
             for nm, par in zip(d["gate_names"], d["gate_params"])]
    keys = [str(k) for k in d["sweep_keys"]]
    sweep = [{k: float(v) for k, v in zip(keys, row) if not np.isnan(v)}
             for row in d["sweep_vals"]]
    ntot = d["Ntotal"] if d["Ntotal"].size else None
    spec = DeviceSpec([d["x"], d["y"]], d["doping"], ntot,
                      dict(zip(MODEL_KEYS, (bool(v) for v in d["models"]))),
                      {str(k): float(v) for k, v in zip(d["material_keys"], d["material_vals"])},
                      contacts, gates, sweep, name)
    return spec, d
