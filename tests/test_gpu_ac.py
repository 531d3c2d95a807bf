"""Small-signal AC on the GPU (tcad_y_parameters through tcadhip.ac)
against the golden fixtures written from the reference's
ac2d.y_parameters, and against the oracle.

Bars:
* every entry of Y(f) within 1e-6 of the largest |Y| entry at that
  frequency, plus FLOOR times the entry's FP64 rounding floor
  sum|S_i| max|du_k| (tests/ac_floor.py), of the oracle's Y refined in
  extended precision (oracle/ac.py, refine=True; pinned against the
  fixtures and the row-sum identity in tests/test_oracle_ac.py).  At a
  heavily doped contact an ohmic entry cancels by up to 1e11: on the
  reverse-biased diode at 1 kHz the reference's own direct solve is off by
  4e-4 of |Y| there (6e-16 of the floor; measured GPU: 2e-16), so the
  fixture is checked with its own distance to the refined value added;
* a batch of devices gives every device the Y it gets alone, bit for bit
  (per-device reductions in a fixed order);
* at 1 Hz the ohmic block is the DC conductance: a central difference of
  two GPU terminal-current solves;
* a z-invariant 3D device matches the oracle's 3D Y.
The DC operating point is the GPU's; each complex solve is refined until
its true residual stops falling.
"""
import warnings

import numpy as np
import pytest

from golden.fixtures import load

pytestmark = pytest.mark.gpu

FIXTURES = ["ac_diode_fwd.npz", "ac_diode_rev.npz", "ac_mosfet.npz"]
FLOOR = 1e-14


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import ac
    from tcadhip.device import NewtonOptions, device_from_spec
    return ac, NewtonOptions, device_from_spec


def _operating_point(name):
    ac, NewtonOptions, device_from_spec = _mods()
    spec, d = load(name)
    dev = device_from_spec(spec)
    opts = NewtonOptions()
    dev.solve_equilibrium(opts)
    bias = {str(k): float(v) for k, v in zip(d["bias_keys"], d["bias_vals"])}
    if bias:
        dev.solve_bias(bias, opts)
    return ac, spec, d, dev


def _oracle_point(spec, d):
    from oracle import physics as oph
    from oracle_build import oracle_device
    ref = oracle_device(spec)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref.solve_equilibrium(oph.NewtonOptions())
        bias = {str(k): float(v) for k, v in zip(d["bias_keys"], d["bias_vals"])}
        if bias:
            ref.solve_bias(bias, oph.NewtonOptions())
    return ref


def _check_close(Y, Yref, floor, what, extra=None):
    """Per entry: |Y - Yref| <= 1e-6 max_f |Yref| + FLOOR * floor (+ extra);
    returns the worst error over tolerance and the worst error over floor."""
    worst, worst_floor = 0.0, 0.0
    for f in range(Yref.shape[0]):
        tol = 1e-6 * np.abs(Yref[f]).max() + FLOOR * floor[f]
        if extra is not None:
            tol = tol + extra[f]
        err = np.abs(Y[f] - Yref[f])
        worst = max(worst, (err / tol).max())
        worst_floor = max(worst_floor, (err / floor[f]).max())
        assert np.all(err <= tol), (what, f, Y[f], Yref[f], floor[f])
    return worst, worst_floor


@pytest.mark.parametrize("name", FIXTURES)
def test_ac_vs_reference(name):
    print(f"[ac] {name}", flush=True)
    from ac_floor import observation_floor
    from oracle import ac as oac
    ac, spec, d, dev = _operating_point(name)
    res = ac.y_parameters(dev, d["freqs"])
    assert list(res.port_names) == [str(p) for p in d["port_names"]]
    ref = _oracle_point(spec, d)
    Yr = oac.y_parameters(ref, d["freqs"], refine=True).Y
    floor = observation_floor(ref, d["freqs"])
    worst, wf = _check_close(res.Y, Yr, floor, name)
    Yfix = d["Y_re"] + 1j * d["Y_im"]
    _check_close(res.Y, Yfix, floor, name + " (fixture)", extra=np.abs(Yfix - Yr))
    print(f"  Krylov iterations {res.iterations}, worst error / tolerance {worst:.2e}, "
          f"worst error / floor {wf:.2e}", flush=True)


def test_ac_batch_matches_alone():
    print("[ac] batch of three devices vs alone", flush=True)
    from tcadhip.device import GPUBatch
    devs = [_operating_point(nm)[3] for nm in FIXTURES]
    freqs = np.array([1e4, 1e9, 1e11])
    alone = []
    for dv in devs:
        B = dv._gpu()
        Y, status, _ = B.y_parameters(freqs, 1e-16, 20000)
        assert status[0] == 0
        alone.append(Y[0])
    batch = GPUBatch([dv._s for dv in devs])
    for k, dv in enumerate(devs):
        sl = batch.device_slice(k)
        src = dv._gpu()
        batch.psi[sl].copy_(src.psi)
        batch.n[sl].copy_(src.n)
        batch.p[sl].copy_(src.p)
        # the lagged surface mobility (MOSFET) lives in the edge fields
        for key in GPUBatch.FIELD_KEYS:
            batch._t[key][sl].copy_(src._t[key])
    Yb, status, _ = batch.y_parameters(freqs, 1e-16, 20000)
    assert np.all(status == 0)
    for k, Ya in enumerate(alone):
        P = Ya.shape[-1]
        assert np.array_equal(Yb[k][:, :P, :P], Ya), k
        assert not np.any(Yb[k][:, P:, :]) and not np.any(Yb[k][:, :, P:])


def test_ac_active_mask_skips_unsolved_devices():
    """Only the selected devices are assembled and preconditioned: a batch
    whose other devices were never solved (their matrices still zero)
    gives the selected device's Y bit for bit as alone, and zeros for the
    others."""
    print("[ac] active mask over devices never assembled", flush=True)
    from tcadhip.device import GPUBatch
    devs = [_operating_point(nm)[3] for nm in FIXTURES]
    freqs = np.array([1e4, 1e9])
    alone, status, _ = devs[0]._gpu().y_parameters(freqs, 1e-16, 20000)
    assert status[0] == 0
    batch = GPUBatch([dv._s for dv in devs])
    sl = batch.device_slice(0)
    src = devs[0]._gpu()
    batch.psi[sl].copy_(src.psi)
    batch.n[sl].copy_(src.n)
    batch.p[sl].copy_(src.p)
    for key in GPUBatch.FIELD_KEYS:
        batch._t[key][sl].copy_(src._t[key])
    Yb, status, _ = batch.y_parameters(freqs, 1e-16, 20000, active=[1, 0, 0])
    assert status[0] == 0
    P = alone.shape[-1]
    assert np.array_equal(Yb[0][:, :P, :P], alone[0])
    assert not np.any(Yb[1:])


def test_ac_low_frequency_is_dc_conductance():
    print("[ac] 1 Hz conductance vs DC difference", flush=True)
    ac, NewtonOptions, _ = _mods()
    _, spec, d, dev = _operating_point("ac_diode_fwd.npz")
    G = ac.y_parameters(dev, [1.0]).Y[0, 0, 0]
    V = float(d["bias_vals"][list(d["bias_keys"]).index("anode")])
    dV = 1e-4
    I = []
    for v in (V - dV, V + dV):
        dev.solve_bias({"anode": v}, NewtonOptions())
        I.append(dev.terminal_current("anode"))
    G_fd = (I[1] - I[0]) / (2 * dV)
    assert abs(G.real - G_fd) <= 1e-4 * abs(G_fd), (G, G_fd)
    assert abs(G.imag) <= 1e-6 * abs(G.real)


def test_ac_3d_vs_oracle():
    print("[ac] z-invariant 3D diode vs oracle", flush=True)
    ac, NewtonOptions, _ = _mods()
    from oracle import ac as oac
    from oracle import dd
    from oracle import physics as oph
    import tcadhip
    x = np.linspace(0, 1e-4, 14)
    y = np.linspace(0, 0.3e-4, 6)
    z = np.linspace(0, 0.2e-4, 3)
    dop = np.tile(np.where(x < 0.45e-4, -2e17, 5e17), (3, 6, 1))
    J, K = np.meshgrid(np.arange(6), np.arange(3), indexing="ij")
    m = oph.Models(doping_mobility=True, srh=True, auger=True, bgn=True)
    ref = dd.Device3D(x, y, z, dop, models=m)
    gpu = tcadhip.Device3D(tcadhip.Mesh3D(x, y, z), dop,
                           models=tcadhip.Models(doping_mobility=True, srh=True, auger=True,
                                                 bgn=True))
    for dv in (ref, gpu):
        dv.add_contact("a", 0, J.ravel(), K.ravel())
        dv.add_contact("c", 13, J.ravel(), K.ravel())
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref.solve_equilibrium()
        ref.solve_bias({"a": 0.3})
    gpu.solve_equilibrium(NewtonOptions())
    gpu.solve_bias({"a": 0.3}, NewtonOptions())
    from ac_floor import observation_floor
    f = [1e6, 1e10]
    Yr = oac.y_parameters(ref, f, refine=True).Y
    Yg = ac.y_parameters(gpu, f).Y
    worst, wf = _check_close(Yg, Yr, observation_floor(ref, f), "3d")
    print(f"  worst error / tolerance {worst:.2e}, worst error / floor {wf:.2e}", flush=True)


def test_cutoff_frequency_matches_oracle():
    """f_T of the MOSFET (|h21| = 1 between gate and drain) from the GPU's
    Y and from the oracle's Y of the same device."""
    print("[ac] MOSFET cutoff frequency", flush=True)
    from ac_floor import observation_floor
    from oracle import ac as oac
    ac, spec, d, dev = _operating_point("ac_mosfet.npz")
    freqs = np.logspace(8, 13, 11)
    res = ac.y_parameters(dev, freqs)
    ref = _oracle_point(spec, d)
    ores = oac.y_parameters(ref, freqs, refine=True)
    _check_close(res.Y, ores.Y, observation_floor(ref, freqs), "mosfet sweep")
    fg = ac.cutoff_frequency(res, "gate", "drain")
    fo = oac.cutoff_frequency(ores, "gate", "drain")
    print(f"  f_T gpu {fg} oracle {fo}", flush=True)
    assert (fg is None) == (fo is None)
    if fg is not None:
        assert abs(fg - fo) <= 1e-5 * fo


def test_ac_3d_trigate_with_gates_vs_oracle():
    """A small tri-gate (top and side gates, 12x10x8): the 3D gate ports'
    displacement rows and the ohmic rows against the oracle's refined 3D Y
    at the same bias."""
    print("[ac] small 3D tri-gate with gate ports vs oracle", flush=True)
    ac, NewtonOptions, device_from_spec = _mods()
    from ac_floor import observation_floor
    from oracle import ac as oac
    from oracle import physics as oph
    from oracle_build import oracle_device
    from tcadhip import configs
    spec = configs.trigate_3d(nx=12, ny=10, nz=8, vg_list=[0.0, 0.25, 0.5], Vd=0.05,
                              name="ac_trigate_small")
    gpu = device_from_spec(spec)
    gpu.solve_equilibrium(NewtonOptions(max_iter=400))
    ref = oracle_device(spec)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref.solve_equilibrium(oph.NewtonOptions(max_iter=400))
        for bias in spec.sweep:
            ref.solve_bias(bias, oph.NewtonOptions())
    for bias in spec.sweep:
        gpu.solve_bias(bias, NewtonOptions())
    f = [1e9, 1e11]
    res = ac.y_parameters(gpu, f)
    ores = oac.y_parameters(ref, f, refine=True)
    assert list(res.port_names) == list(ores.port_names)
    assert sum(p.startswith("gate") for p in res.port_names) == 3
    worst, wf = _check_close(res.Y, ores.Y, observation_floor(ref, f), "3d tri-gate")
    print(f"  ports {list(res.port_names)}, worst error / tolerance {worst:.2e}", flush=True)
