"""The red-black tile order (NewtonOptions.tile_rb, TCAD_TUNE_TILE_RB)
against tile block-Jacobi on whole sweeps of the reference's fixtures.

The ordering changes only the preconditioner, so a sweep must converge to
the same states either way.  Bars, per fixture:
  * terminal currents of each ordering within the current bar
    (tests/current_bar.py) of the reference's fixture currents;
  * Newton counts of the two orderings within +-1 of each other at every
    point (the refinement decides each linear solve to the same target);
  * fewer BiCGStab iterations over the sweep with the red-black order
    than with block-Jacobi (the forward block Gauss-Seidel sweep over the
    tile checkerboard halves them on these systems).
"""
import numpy as np
import pytest

from current_bar import check, contacts_overlap
from golden.fixtures import load

pytestmark = pytest.mark.gpu

CASES = ["mosfet_48x32_sweep.npz", "diode_60x40_sweep.npz", "trigate_24x18x14_vd07_sweep.npz"]


def _say(msg):
    print(msg, flush=True)


def _sweep(spec, opts, eq_opts):
    from tcadhip.device import device_from_spec
    dev = device_from_spec(spec)
    dev.solve_equilibrium(eq_opts)
    its, cur, kry = [], [], 0
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        kry += dev.last_krylov_iterations
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
    return np.array(its), np.array(cur), kry


@pytest.mark.parametrize("name", CASES)
def test_red_black_against_block_jacobi(name):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip.device import NewtonOptions
    spec, d = load(name)
    out = {}
    for rb in (True, False):
        eq = NewtonOptions(max_iter=400, tile_rb=rb) if spec.dim == 3 else NewtonOptions(tile_rb=rb)
        out[rb] = _sweep(spec, NewtonOptions(tile_rb=rb), eq)
        its, cur, kry = out[rb]
        _say(f"[{name}] {'red-black' if rb else 'block-Jacobi'}: Newton {its.tolist()} "
             f"Krylov {kry}")
        check(cur, d["currents"], f"{name} tile_rb={rb}", contacts_overlap(spec), say=_say)
    assert np.all(np.abs(out[True][0] - out[False][0]) <= 1), (out[True][0], out[False][0])
    assert out[True][2] < out[False][2], (out[True][2], out[False][2])
