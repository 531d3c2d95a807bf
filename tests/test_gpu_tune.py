"""The autotuner (tcadhip.tune) on the reduced 24x18x14 tri-gate:

* every trial restarts from the same saved state, and the device's state
  and terminal voltages come back bit for bit after tuning;
* the chosen configuration was accepted (converged, currents within 1e-9
  of the untuned configuration's largest current, no extra Newton
  iteration on the tuning point);
* the sweep run with it meets the reference's own fixture: every current
  within 1e-6 of the largest (no floor: the tri-gate's contacts share
  nodes), Newton counts within one of the reference's.
"""
import numpy as np
import pytest

from golden.fixtures import load

pytestmark = pytest.mark.gpu


def test_tuner_restores_state_and_tuned_sweep_meets_reference():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import tune
    from tcadhip.device import NewtonOptions, device_from_spec
    tune.clear_cache()
    spec, d = load("trigate_24x18x14_vd07_sweep.npz")
    print("[tune] 24x18x14 tri-gate, Vd 0.7", flush=True)
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions(max_iter=400))
    dev.solve_bias(spec.sweep[0], NewtonOptions())
    before = [a.copy() for a in (dev.psi, dev.n, dev.p)]
    volts = {nm: (bc.V if hasattr(bc, "V") else bc.Vg) for nm, bc in dev.bcs.items()}
    opts, report = tune.tune(dev, spec.sweep[1], NewtonOptions(),
                             say=lambda m: print(" ", m, flush=True))
    after = [a.copy() for a in (dev.psi, dev.n, dev.p)]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert volts == {nm: (bc.V if hasattr(bc, "V") else bc.Vg) for nm, bc in dev.bcs.items()}
    rows = report["candidates"]
    assert rows[0]["accepted"]
    chosen = [r for r in rows if r["accepted"] and
              all(r.get(k) == v for k, v in report["chosen"].items())]
    assert chosen, report["chosen"]
    assert all(r["newton"] <= rows[0]["newton"] for r in rows if r["accepted"])
    # the sweep with the tuned options against the reference's fixture
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions(max_iter=400))
    its, cur = [], []
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
    its, cur = np.array(its), np.array(cur)
    ref_its, ref = d["newton_iterations"], d["currents"]
    print(f"  tuned {report['chosen']}: Newton {its.tolist()} reference {ref_its.tolist()}",
          flush=True)
    assert np.all(np.abs(its - ref_its) <= 1)
    for p in range(ref.shape[0]):
        big = np.abs(ref[p]).max()
        assert np.all(np.abs(cur[p] - ref[p]) <= 1e-6 * big), (p, cur[p], ref[p])
