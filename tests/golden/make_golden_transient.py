"""Generate the transient golden fixtures by running the REFERENCE
(pytcad.transient2d.solve_transient on pytcad.device2d.Device2D) in the
development container.  Not run by the test suite; the fixtures it writes
(tests/golden/transient_*.npz) are committed and the reference never
travels anywhere.

Only the reference pins the time-step sequence (adaptive dt), the line-
search Newton iterates and iteration counts, and the per-step terminal
currents.  Each fixture also holds the state the transient starts from
(psi_init, n_init, p_init: equilibrium, then the pre-bias if any), so a
test can start from exactly the reference's initial state.  The reference keeps no iteration counter, so its
transient2d._newton_step is wrapped to record the count it returns.

Usage:  python tests/golden/make_golden_transient.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as mg                              # noqa: E402  (sets up the reference)

from pytcad import transient as rt                    # noqa: E402
from pytcad import transient2d as rt2                 # noqa: E402
from pytcad.device import NewtonOptions               # noqa: E402

from tcadhip import configs                           # noqa: E402

ITS = []
_newton_step = rt2._newton_step


def _count_newton_step(*a, **k):
    out = _newton_step(*a, **k)
    if out[3]:
        ITS.append(out[4])
    return out


rt2._newton_step = _count_newton_step


def run(spec, prebias, waveforms, wf_desc, t_end, dt0, theta, output_times=None):
    dev = mg.ref_device(spec)
    opts = NewtonOptions(linsolve="direct")
    dev.solve_equilibrium(opts)
    if prebias:
        dev.solve_bias(prebias, opts)
    init = {"psi_init": dev.psi.copy(), "n_init": dev.n.copy(), "p_init": dev.p.copy()}
    ITS.clear()
    res = rt2.solve_transient(dev, waveforms, t_end, dt0, theta=theta, opts=opts,
                              output_times=output_times)
    names = [c.name for c in spec.contacts]
    out = {"times": res.times, "dt_hist": res.dt_hist, "iterations": np.array(ITS),
           "currents": np.array([res.terminal_current[nm] for nm in names]).T,
           "psi_final": res.psi_hist[-1], "n_final": res.n_hist[-1], "p_final": res.p_hist[-1],
           "t_end": np.array(t_end), "dt0": np.array(dt0), "theta": np.array(theta),
           "output_times": np.array(output_times or [], dtype=float),
           "prebias_keys": np.array(sorted(prebias or {})),
           "prebias_vals": np.array([prebias[k] for k in sorted(prebias or {})], dtype=float),
           "wf_desc": np.array(wf_desc, dtype=str), **init}
    return out


def main():
    spec = configs.diode_2d(nx=24, ny=16, Na=3e18, Nd=2e15, xj_frac=0.4, biases=[0.0],
                            models=dict(doping_mobility=True, srh=True, auger=True,
                                        bgn=True, surface_mobility=False),
                            material={}, name="transient_diode_24x16")
    # backward Euler, forward voltage step on the anode
    mg.save("transient_diode_be.npz", spec,
            run(spec, None, {"anode": rt.StepWaveform(0.0, 0.5, 0.0)},
                ["anode", "step", "0.0", "0.5", "0.0"], 2e-10, 1e-12, 1.0))
    # Crank-Nicolson, ramp to reverse bias, with output times
    mg.save("transient_diode_cn.npz", spec,
            run(spec, None, {"anode": rt.RampWaveform(0.0, -1.0, 0.0, 5e-11)},
                ["anode", "ramp", "0.0", "-1.0", "0.0", "5e-11"], 1e-10, 2e-12, 0.5,
                output_times=[1e-11, 5e-11]))
    # MOSFET: gate at 0.8 V, drain step 0 -> 0.5 V (gate held; the lagged
    # surface mobility is not updated by the transient driver)
    spec = configs.mosfet_2d(nx=24, ny=16, vg_list=[0.8], Vd=0.0, name="transient_mosfet_24x16")
    mg.save("transient_mosfet_be.npz", spec,
            run(spec, {"gate": 0.8, "drain": 0.0}, {"drain": rt.StepWaveform(0.0, 0.5, 0.0)},
                ["drain", "step", "0.0", "0.5", "0.0"], 5e-11, 1e-13, 1.0))


if __name__ == "__main__":
    main()
