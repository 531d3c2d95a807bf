"""Build oracle devices from a DeviceSpec (test helper)."""
import numpy as np

from oracle import dd
from oracle import physics as oph


def oracle_device(spec):
    mat = oph.replace(oph.SILICON, **spec.material) if spec.material else oph.SILICON
    models = oph.Models(**spec.models)
    if spec.dim == 2:
        x, y = spec.coords
        dev = dd.Device2D(x, y, spec.doping, spec.Ntotal, material=mat, models=models)
        for c in spec.contacts:
            dev.add_contact(c.name, c.idx[1], c.idx[0], V=c.V)
        for g in spec.gates:
            dev.add_gate(g.name, g.idx[1], g.idx[0], g.tox_cm, g.Vfb, g.Vg)
    else:
        x, y, z = spec.coords
        dev = dd.Device3D(x, y, z, spec.doping, spec.Ntotal, material=mat, models=models)
        for c in spec.contacts:
            dev.add_contact(c.name, c.idx[2], c.idx[1], c.idx[0], V=c.V)
        for g in spec.gates:
            dev.add_gate(g.name, g.idx[2], g.idx[1], g.idx[0], g.tox_cm, g.Vfb, g.Vg,
                         normal_axis=g.normal_axis)
    return dev


def oracle_sweep(spec, opts=None, keep=()):
    """Equilibrium then the warm-started sweep; returns per-point Newton
    iteration counts, currents (contact order of the spec) and kept
    states."""
    dev = oracle_device(spec)
    dev.solve_equilibrium(opts)
    its, cur, states = [], [], {}
    for p, bias in enumerate(spec.sweep):
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
        if p in keep:
            states[p] = (dev.psi.copy(), dev.n.copy(), dev.p.copy())
    return dev, np.array(its), np.array(cur), states
