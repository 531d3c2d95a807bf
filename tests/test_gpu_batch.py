"""GPU batching: masked solves, explicit node offsets, per-device outcome
codes and the batched continuation sweep.

Bars: a device's results in a batch are the same bits as alone (the
kernels never mix devices and every reduction runs in a fixed order per
device); against the CPU oracle, terminal currents within the current bar
of tests/current_bar.py (1e-6 of the largest current plus each side's own
conservation error, |sum_c I|, as the floor) and Newton counts within +-1.
"""
import warnings

import numpy as np
import pytest

from oracle_build import oracle_sweep

pytestmark = pytest.mark.gpu


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import configs, sweep, _lib
    from tcadhip.device import GPUBatch, NewtonOptions, device_from_spec, setup_from_spec
    return configs, sweep, _lib, GPUBatch, NewtonOptions, device_from_spec, setup_from_spec


def _say(msg):
    print(msg, flush=True)


def _small_diodes(count, seed=7):
    configs = _mods()[0]
    return configs.batch_diodes(count=count, seed=seed, nx_range=(16, 40), ny_range=(10, 24),
                                max_step=0.25)


def _alone(spec, opts):
    device_from_spec = _mods()[5]
    dev = device_from_spec(spec)
    dev.solve_equilibrium(opts)
    its, cur = [], []
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
    return np.array(its), np.array(cur), dev


DRIVERS = {"rounds": dict(queue=False), "queue": dict(queue=True),
           "queue2": dict(queue=True, queue_depth=2)}


@pytest.mark.parametrize("mode,driver", [("bicgstab", "rounds"), ("bicgstab", "queue"),
                                         ("bicgstab", "queue2"), ("bicgstab_mixed", "rounds"),
                                         ("bicgstab_mixed", "queue")])
def test_batched_sweep_equals_alone(mode, driver):
    """Rounds (one masked solve per point over every device) and the point
    queue (each device through its points without a barrier, whole list or
    two points per pass with compaction between) give every device the bits
    it gets alone."""
    configs, sweep, _lib, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = _small_diodes(6)
    _say(f"[batch-sweep] 6 diodes, {mode}, {driver}: batched continuation vs alone")
    opts = NewtonOptions(linsolve=mode)
    setups = [setup_from_spec(s) for s in specs]
    res = sweep.BatchedSweep(setups, [s.sweep for s in specs], opts, **DRIVERS[driver]).run()
    for d, s in enumerate(specs):
        R = res[d]
        assert not R.failed, R.failed
        assert R.backoffs == 0
        its, cur, _ = _alone(s, opts)
        got = np.array([[c[k.name] for k in s.contacts] for c in R.currents])
        assert got.shape == cur.shape
        assert np.array_equal(np.array(R.newton), its)
        assert np.array_equal(got, cur), (d, got, cur)


def test_batched_sweep_vs_oracle():
    configs, sweep, _lib, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = _small_diodes(3, seed=11)
    _say("[batch-sweep] 3 diodes vs the CPU oracle")
    opts = NewtonOptions()
    res = sweep.BatchedSweep([setup_from_spec(s) for s in specs],
                             [s.sweep for s in specs], opts).run()
    from current_bar import check, oracle_sweep
    for d, s in enumerate(specs):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            rits, rcur = oracle_sweep(s)
        got = np.array([[c[k.name] for k in s.contacts] for c in res[d].currents])
        check(got, rcur, s.name, say=_say)
        assert np.all(np.abs(np.array(res[d].newton) - rits) <= 1)


def test_mask_leaves_unselected_devices_untouched():
    configs, sweep, _lib, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = _small_diodes(3, seed=5)
    _say("[mask] solve devices 0 and 2 only")
    opts = NewtonOptions()
    setups = [setup_from_spec(s) for s in specs]
    B = GPUBatch(setups)
    B.solve_equilibrium(opts)
    before = [t.clone() for t in (B.psi, B.n, B.p)]
    for s, sp in zip(setups, specs):
        for nm, v in sp.sweep[0].items():
            s.bcs[nm].V = float(v)
    its, conv, kry, status = B.solve_bias(opts, active=[1, 0, 1])
    assert conv[0] and conv[2] and not conv[1]
    assert its[1] == 0 and kry[1] == 0 and status[1] == _lib.NEWTON_OK
    sl = B.device_slice(1)
    for a, b in zip((B.psi, B.n, B.p), before):
        assert bool((a[sl] == b[sl]).all())
    sl0 = B.device_slice(0)
    assert not bool((B.psi[sl0] == before[0][sl0]).all())


def test_explicit_offsets_match_packed():
    import torch
    configs, sweep, _lib, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = _small_diodes(2, seed=3)
    _say("[offsets] two diodes at explicit offsets with a gap vs packed")
    opts = NewtonOptions()
    packed = GPUBatch([setup_from_spec(s) for s in specs])
    packed.solve_equilibrium(opts)
    setups = [setup_from_spec(s) for s in specs]
    gap = 1000
    offs = [gap, gap + setups[0].N + 77]
    L = offs[1] + setups[1].N + 5
    planes = {k: torch.zeros(L, dtype=torch.float64, device="cuda")
              for k in GPUBatch.FIELD_KEYS + ("psi", "n", "p")}
    for s, o in zip(setups, offs):
        f = s.node_fields()
        for k in GPUBatch.FIELD_KEYS:
            planes[k][o:o + s.N] = torch.from_numpy(f[k]).to("cuda")
    B = GPUBatch(setups, offsets=offs, planes=planes)
    B.solve_equilibrium(opts)
    for d in range(2):
        a = packed.psi[packed.device_slice(d)]
        b = B.psi[B.device_slice(d)]
        assert bool((a == b).all())
    assert bool((planes["psi"][:gap] == 0).all())


def test_linear_failure_is_reported_per_device():
    configs, sweep, _lib, GPUBatch, NewtonOptions, device_from_spec, setup_from_spec = _mods()
    specs = _small_diodes(2, seed=9)
    _say("[linfail] 1 Krylov iteration per refinement sweep, no fallback: the device must stop, "
         "not step")
    opts = NewtonOptions(lin_maxiter=1, linear_fallback=False)
    B = GPUBatch([setup_from_spec(s) for s in specs])
    B.solve_equilibrium(NewtonOptions())
    psi0 = B.psi.clone()
    for s, sp in zip(B.setups, specs):
        for nm, v in sp.sweep[0].items():
            s.bcs[nm].V = float(v)
    its, conv, kry, status = B.solve_bias(opts)
    assert list(status) == [_lib.NEWTON_LINEAR_FAIL] * 2
    assert not conv.any()
    dev = device_from_spec(specs[0])
    dev.solve_equilibrium()
    with pytest.raises(_lib.LinearSolveError):
        dev.solve_bias(specs[0].sweep[0], opts)
    del psi0


def test_linear_fallback_per_iteration():
    """With BiCGStab cut to one iteration per refinement sweep every Newton
    solve fails, and each is solved again for that iteration by the
    fallback (GMRES + ILU(0) on the device's CSR system), as the
    reference's 3D Newton loop falls back to a direct solve per iteration
    (device3d.py:2154-2168): the sweep then converges with the counts and
    currents of the plain path (the fallback solves the same system to the
    same target), batched and alone alike."""
    configs, sweep, _lib, GPUBatch, NewtonOptions, device_from_spec, setup_from_spec = _mods()
    specs = _small_diodes(2, seed=9)
    _say("[linfallback] 2 diodes, every Newton solve through the fallback")
    weak = NewtonOptions(lin_maxiter=1)
    for s in specs:
        dev = device_from_spec(s)
        dev.solve_equilibrium()
        ref = device_from_spec(s)
        ref.solve_equilibrium()
        for bias in s.sweep[:3]:
            dev.solve_bias(bias, weak)
            ref.solve_bias(bias, NewtonOptions())
            assert dev.last_converged
            assert abs(dev.last_iterations - ref.last_iterations) <= 1
            got = np.array([dev.terminal_current(c.name) for c in s.contacts])
            want = np.array([ref.terminal_current(c.name) for c in s.contacts])
            big = np.abs(want).max()
            floor = abs(want.sum()) + abs(got.sum())
            assert np.all(np.abs(got - want) <= 1e-6 * big + floor), (got, want)
    B = GPUBatch([setup_from_spec(s) for s in specs])
    B.solve_equilibrium(NewtonOptions())
    for s, sp in zip(B.setups, specs):
        for nm, v in sp.sweep[0].items():
            s.bcs[nm].V = float(v)
    its, conv, kry, status = B.solve_bias(weak)
    assert list(status) == [_lib.NEWTON_OK] * 2 and conv.all()
