"""3D tri-gate on the GPU beyond the reference fixtures of
tests/test_gpu_parity.py: ragged patch edges against the live CPU oracle,
and the full 64x48x48 mesh (442,368 unknowns, where the oracle's direct
solve is out of reach) through invariants.

Bars: Newton counts within +-1 of the oracle's; terminal currents within
1e-6 * max_c |I_ref| per point (the tri-gate's currents are not
cancellation-limited; its overlapping source/drain/body node lists make
sum_c I nonzero by construction, so no conservation floor is used);
psi within 1e-9 (scaled).  Full size: the device is mirror-symmetric in z
(symmetric mesh, doping and side gates), so the converged psi must be too,
to 1e-8 (scaled, i.e. 2.6e-10 V); the FP32-matrix path must give the FP64
path's drain current to 1e-6 relative.
"""
import numpy as np
import pytest

from oracle_build import oracle_sweep

pytestmark = pytest.mark.gpu


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import configs
    from tcadhip.device import NewtonOptions, device_from_spec
    return configs, NewtonOptions, device_from_spec


def _say(msg):
    print(msg, flush=True)


def _sweep(spec, opts):
    configs, NewtonOptions, device_from_spec = _mods()
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions(max_iter=400))
    its, cur = [], []
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        assert dev.last_converged
        its.append(dev.last_iterations)
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
    return dev, np.array(its), np.array(cur)


def test_trigate_full_size_invariants():
    configs, NewtonOptions, _ = _mods()
    _say("[3d] trigate 64x48x48 Vd=0.05 Vg 0,0.4: z-mirror symmetry, FP32 vs FP64")
    spec = configs.config3_trigate(0.05)
    spec.sweep = [spec.sweep[0], spec.sweep[8]]
    dev64, its64, cur64 = _sweep(spec, NewtonOptions(linsolve="bicgstab"))
    psi = dev64.psi
    assert psi.shape == (48, 48, 64)
    asym = np.abs(psi - psi[::-1]).max()
    _say(f"  z-mirror asymmetry of psi {asym:.2e}")
    assert asym <= 1e-8
    n = dev64.n
    assert (np.abs(n - n[::-1]) / n).max() <= 1e-7
    dev32, its32, cur32 = _sweep(spec, NewtonOptions(linsolve="bicgstab_mixed"))
    names = [c.name for c in spec.contacts]
    d = names.index("drain")
    rel = np.abs(cur32[:, d] - cur64[:, d]) / np.abs(cur64[:, d])
    _say(f"  FP32-matrix vs FP64 drain current {rel.max():.2e}; Newton {its64} {its32}")
    assert rel.max() <= 1e-6
    assert np.all(np.abs(its32 - its64) <= 1)


@pytest.mark.parametrize("shape", [("trigate", (19, 7, 9)), ("mosfet", (37, 11))])
def test_ragged_patches_tile_and_node_precond(shape):
    """Devices whose sizes leave the patches ragged (3D patches are
    16 x 1 x 8 nodes, 2D patches 32 x 4: 19 % 16, 9 % 8, 37 % 32 and 11 % 4
    are all nonzero), so the tile preconditioner's edge tiles
    hold nodes outside the device: the tile and node block-Jacobi
    preconditioners reach the same Newton counts and states, and both match
    the live oracle (Newton +-1, currents 1e-6 of the largest, psi 1e-9)."""
    configs, NewtonOptions, _ = _mods()
    kind, dims = shape
    _say(f"[3d] ragged {kind} {dims}: tile vs node block-Jacobi vs oracle")
    if kind == "trigate":
        spec = configs.trigate_3d(nx=dims[0], ny=dims[1], nz=dims[2], Vd=0.05,
                                  vg_list=[0.0, 0.8])
    else:
        spec = configs.mosfet_2d(nx=dims[0], ny=dims[1], vg_list=[0.0, 1.0])
    odev, oits, ocur, ost = oracle_sweep(spec, keep=(1,))
    dt, its_t, cur_t = _sweep(spec, NewtonOptions(precond="tile"))
    dn, its_n, cur_n = _sweep(spec, NewtonOptions(precond="node"))
    assert np.array_equal(its_t, its_n), (its_t, its_n)
    assert np.all(np.abs(its_t - oits) <= 1), (its_t, oits)
    for p in range(len(spec.sweep)):
        big = np.abs(ocur[p]).max()
        assert np.abs(cur_t[p] - cur_n[p]).max() <= 1e-9 * big, (p, cur_t[p], cur_n[p])
        assert np.abs(cur_t[p] - ocur[p]).max() <= 1e-6 * big, (p, cur_t[p], ocur[p])
    assert np.abs(dt.psi - dn.psi).max() <= 1e-9
    assert np.abs(dt.psi - ost[1][0]).max() <= 1e-9


def test_linear_fallback_3d_vs_reference():
    """Every Newton solve of the reduced tri-gate through the per-iteration
    fallback (BiCGStab cut to one iteration per refinement sweep, so each
    solve is redone by GMRES + ILU(0) on the device's CSR system, as
    device3d.py:2154-2168 redoes a failed iterative solve directly): the
    first two points of the reference's Vd = 0.05 V sweep, Newton counts
    within one, every current within 1e-6 of the largest (the contacts
    share nodes: no floor)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from golden.fixtures import load
    from tcadhip.device import NewtonOptions, device_from_spec
    spec, d = load("trigate_24x18x14_vd005_sweep.npz")
    print("[fallback-3d] 24x18x14 tri-gate, every solve through GMRES + ILU(0)", flush=True)
    dev = device_from_spec(spec)
    dev.solve_equilibrium(NewtonOptions(max_iter=400))
    weak = NewtonOptions(lin_maxiter=1)
    for p, bias in enumerate(spec.sweep[:2]):
        dev.solve_bias(bias, weak)
        cur = np.array([dev.terminal_current(c.name) for c in spec.contacts])
        ref = d["currents"][p]
        print(f"  point {p}: Newton {dev.last_iterations} reference "
              f"{int(d['newton_iterations'][p])}", flush=True)
        assert abs(dev.last_iterations - int(d["newton_iterations"][p])) <= 1
        assert np.all(np.abs(cur - ref) <= 1e-6 * np.abs(ref).max()), (cur, ref)
