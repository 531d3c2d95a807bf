"""The secant predictor of sweeps (tcad_predict_state; tcadhip.sweep
device_sweep and BatchedSweep(predictor=True)).

The predictor changes only where Newton starts, so its bars are the plain
sweep's:
  * the kernel: psi_prev + (1 + f)(psi - psi_prev) extrapolation, n and p
    geometric inside the Newton clip [0.1x, 10x], against the same formula
    in NumPy (f = 1: the same operations, bit for bit; f = 0.5: pow, within
    1e-14 relative); devices with f <= 0 untouched;
  * sweeps against the reference's fixtures: terminal currents within the
    current bar (tests/current_bar.py), converged psi within 1e-9 and n, p
    within 1e-7 relative (densities above 1e-10), as the plain sweep is
    held; Newton counts at most the reference's + 1 at every point.  The
    +-1 match of the reference's counts is a property of its own start (the
    previous converged state), which the plain sweep keeps and is held to
    (test_gpu_parity.py); a predicted start is closer to the solution and
    may save more than one iteration where the bias steps are coarse (the
    reduced tri-gate fixtures step Vg by 0.2 V: 5 -> 3 at the last point);
  * a batch with the predictor gives every device the bits it gets alone
    with the predictor;
  * forcing terms for predicted starts only (NewtonOptions.
    forcing_predicted, TCAD_TUNE_FORCING_PREDICTED): in one solve of a batch
    whose starts were predicted for some devices only, the predicted
    devices get the bits of the same solve with forcing terms for all, the
    others the bits of the solve without forcing terms (Newton and Krylov
    counts, psi, n, p).
"""
import numpy as np
import pytest

from current_bar import check, contacts_overlap
from golden.fixtures import load

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import configs, sweep
    from tcadhip.device import GPUBatch, NewtonOptions, device_from_spec, setup_from_spec
    return torch, configs, sweep, GPUBatch, NewtonOptions, device_from_spec, setup_from_spec


def test_predict_kernel_formula():
    torch, configs, sweep, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = configs.batch_diodes(count=3, seed=5, nx_range=(16, 40), ny_range=(10, 24))
    B = GPUBatch([setup_from_spec(s) for s in specs])
    N = B.psi.numel()
    rng = np.random.default_rng(3)
    psi0, psi = rng.normal(0, 10, N), rng.normal(0, 10, N)
    n0, n = 10.0 ** rng.uniform(-20, 5, N), 10.0 ** rng.uniform(-20, 5, N)
    p0, p = 10.0 ** rng.uniform(-20, 5, N), 10.0 ** rng.uniform(-20, 5, N)
    prev = [torch.from_numpy(v).cuda() for v in (psi0, n0, p0)]
    for t, v in ((B.psi, psi), (B.n, n), (B.p, p)):
        t.copy_(torch.from_numpy(v))
    torch.cuda.synchronize()
    ratios = np.array([1.0, 0.5, 0.0])
    with B.on_stream():
        B.predict(prev, ratios)
    B.stream_sync()
    got = [t.cpu().numpy() for t in (B.psi, B.n, B.p)]
    for d, f in enumerate(ratios):
        sl = B.device_slice(d)
        if f <= 0.0:
            for g, v in zip(got, (psi, n, p)):
                assert np.array_equal(g[sl], v[sl])
            continue
        ep = psi[sl] + f * (psi[sl] - psi0[sl])
        rn = (n[sl] / n0[sl]) if f == 1.0 else (n[sl] / n0[sl]) ** f
        rp = (p[sl] / p0[sl]) if f == 1.0 else (p[sl] / p0[sl]) ** f
        en = n[sl] * np.clip(rn, 0.1, 10.0)
        epp = p[sl] * np.clip(rp, 0.1, 10.0)
        tol = 0.0 if f == 1.0 else 1e-14
        for g, e, nm in ((got[0][sl], ep, "psi"), (got[1][sl], en, "n"), (got[2][sl], epp, "p")):
            err = np.abs(g - e) / np.maximum(np.abs(e), 1e-300)
            _say(f"[predict] device {d} f={f} {nm}: max rel err {err.max():.2e}")
            assert err.max() <= tol, (d, nm, err.max())


SWEEPS = ["diode_60x40_sweep.npz", "mosfet_48x32_sweep.npz", "trigate_24x18x14_vd005_sweep.npz",
          "trigate_24x18x14_vd07_sweep.npz", "slab_xgate_10x12x9_sweep.npz"]


@pytest.mark.parametrize("name", SWEEPS)
def test_predicted_sweep_vs_reference(name):
    torch, configs, sweep, GPUBatch, NewtonOptions, device_from_spec, _ = _mods()
    _say(f"[predicted sweep] {name}")
    spec, d = load(name)
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions(max_iter=400) if spec.dim == 3 else NewtonOptions())
    cur, states = [], {}

    def rec(k, its, kry):
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
        states[k] = (dev.psi, dev.n, dev.p)

    out = sweep.device_sweep(dev, spec.sweep, NewtonOptions(), predictor=True, on_point=rec)
    its = np.array([o[0] for o in out])
    ref_its = d["newton_iterations"]
    _say(f"  Newton predicted {its.tolist()} reference {ref_its.tolist()}; Krylov "
         f"{sum(o[1] for o in out)}")
    assert np.all(its <= ref_its + 1), (its, ref_its)
    n, strict, fl = check(np.array(cur), d["currents"], name, contacts_overlap(spec), say=_say)
    _say(f"  {strict}/{n} points within 1e-6 with no floor, {fl} needed the floor")
    for k in d.files:
        if k.startswith("psi_") and k != "psi_eq":
            p = int(k.split("_")[1])
            psi, nn, pp = states[p]
            assert np.abs(psi - d[f"psi_{p}"]).max() <= 1e-9
            for got, key in ((nn, f"n_{p}"), (pp, f"p_{p}")):
                ref = d[key]
                m = ref > 1e-10
                if m.any():
                    assert (np.abs(got[m] - ref[m]) / ref[m]).max() <= 1e-7


def test_predicted_batch_equals_alone():
    torch, configs, sweep, GPUBatch, NewtonOptions, device_from_spec, setup_from_spec = _mods()
    specs = configs.batch_diodes(count=5, seed=13, nx_range=(16, 40), ny_range=(10, 24),
                                 max_step=0.1)
    _say("[predicted batch] 5 diodes, batched with the predictor vs each alone with it")
    opts = NewtonOptions()
    res = sweep.BatchedSweep([setup_from_spec(s) for s in specs], [s.sweep for s in specs], opts,
                             predictor=True).run()
    plain = sweep.BatchedSweep([setup_from_spec(s) for s in specs], [s.sweep for s in specs],
                               opts).run()
    for k, s in enumerate(specs):
        R = res[k]
        assert not R.failed, R.failed
        alone = sweep.BatchedSweep([setup_from_spec(s)], [s.sweep], opts, predictor=True).run()[0]
        got = np.array([[c[x.name] for x in s.contacts] for c in R.currents])
        one = np.array([[c[x.name] for x in s.contacts] for c in alone.currents])
        assert np.array_equal(np.array(R.newton), np.array(alone.newton))
        assert np.array_equal(got, one), (k, got, one)
        ref = np.array([[c[x.name] for x in s.contacts] for c in plain[k].currents])
        _say(f"  device {k}: Newton predicted {sum(R.newton)} plain {sum(plain[k].newton)}")
        check(got, ref, f"device {k} vs plain", say=_say)


def test_forcing_for_predicted_starts_only():
    torch, configs, sweep, GPUBatch, NewtonOptions, _, setup_from_spec = _mods()
    specs = configs.batch_diodes(count=3, seed=21, nx_range=(20, 40), ny_range=(12, 24))
    setups = [setup_from_spec(s) for s in specs]
    B = GPUBatch(setups)
    B.solve_equilibrium(NewtonOptions())
    volts = (0.15, 0.3, 0.45)                 # anode bias of the three points

    def bias(k):
        for s in setups:
            s.bcs["anode"].V = volts[k]

    def clone():
        return tuple(t.clone() for t in (B.psi, B.n, B.p))

    bias(0)
    its, conv, _, _ = B.solve_bias(NewtonOptions())
    assert conv.all()
    prev = clone()
    bias(1)
    its, conv, _, _ = B.solve_bias(NewtonOptions())
    assert conv.all()
    start = clone()
    bias(2)
    f = np.array([sweep.secant_ratio(np.array([volts[0]]), np.array([volts[1]]),
                                     np.array([volts[2]]))] * 3)
    assert np.allclose(f, 1.0), f
    f[1] = 0.0                                # device 1 starts unpredicted
    _say(f"[forcing predicted only] ratios {f.tolist()}")
    runs = {}
    for key, opts in (("pred_only", NewtonOptions(forcing=True, forcing_predicted=True)),
                      ("forcing", NewtonOptions(forcing=True)), ("exact", NewtonOptions())):
        for t, v in zip((B.psi, B.n, B.p), start):
            t.copy_(v)
        B.predict(prev, f)
        its, conv, kry, status = B.solve_bias(opts)
        B.stream_sync()
        assert conv.all() and (status == 0).all(), (key, conv, status)
        runs[key] = (its.copy(), kry.copy(), [t.cpu().numpy() for t in (B.psi, B.n, B.p)])
        _say(f"  {key}: Newton {its.tolist()} Krylov {kry.tolist()}")
    for d in range(3):
        ref = runs["exact" if f[d] == 0.0 else "forcing"]
        got = runs["pred_only"]
        sl = B.device_slice(d)
        assert got[0][d] == ref[0][d] and got[1][d] == ref[1][d], (d, got[:2], ref[:2])
        for a, b in zip(got[2], ref[2]):
            assert np.array_equal(a[sl], b[sl]), d
    # forcing terms change the predicted devices' Krylov work, so the
    # comparison above tells the two option sets apart
    assert any(runs["forcing"][1][d] != runs["exact"][1][d] for d in (0, 2)), runs
