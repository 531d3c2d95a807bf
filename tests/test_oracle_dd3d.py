"""3D oracle invariants, beside the bit-exact reference fixtures of
tests/test_oracle_golden.py (made by tests/golden/make_golden_3d.py):

* a z-invariant 3D device equals the golden-pinned 2D oracle
  (the reference's own gate, pytcad/tests/test_validation_3d.py:68-109, 156-187):
  psi identical in every z-plane, I_3D = I_2D * Lz, same Newton counts;
* axis-permutation symmetry: the same device built along x and along z
  gives the permuted solution (pins the z-axis code on the x-axis code);
* finite-difference Jacobian (pytcad/tests/test_validation_3d.py:112-153);
* terminal-current conservation (pytcad/tests/test_validation_3d.py:190-217).
"""
import numpy as np

from oracle import dd
from oracle import physics as oph

MODELS = oph.Models(doping_mobility=True, srh=True, auger=True, bgn=True)


def diode2d(nx=16, ny=6):
    x = np.linspace(0, 1e-4, nx)
    y = np.linspace(0, 0.3e-4, ny)
    dop = np.tile(np.where(x < 0.45e-4, -2e17, 5e17), (ny, 1))
    d = dd.Device2D(x, y, dop, models=MODELS)
    d.add_contact("a", 0, np.arange(ny))
    d.add_contact("c", nx - 1, np.arange(ny))
    return d, x, y, dop


def test_z_invariant_3d_equals_2d():
    d2, x, y, dop = diode2d()
    nz = 4
    z = np.linspace(0, 0.2e-4, nz)
    d3 = dd.Device3D(x, y, z, np.tile(dop, (nz, 1, 1)), models=MODELS)
    J, K = np.meshgrid(np.arange(y.size), np.arange(nz), indexing="ij")
    d3.add_contact("a", 0, J.ravel(), K.ravel())
    d3.add_contact("c", x.size - 1, J.ravel(), K.ravel())
    for dev in (d2, d3):
        dev.solve_equilibrium()
    assert d2.last_eq_iterations == d3.last_eq_iterations
    for k in range(nz):
        assert np.abs(d3.psi[k] - d2.psi).max() <= 1e-9
    for V in (0.2, 0.45):
        d2.solve_bias({"a": V})
        d3.solve_bias({"a": V})
        assert d2.last_iterations == d3.last_iterations
        for k in range(nz):
            assert np.abs(d3.psi[k] - d2.psi).max() <= 1e-9
            assert (np.abs(d3.n[k] - d2.n) / d2.n).max() <= 1e-7
        I2 = d2.terminal_current("a")
        I3 = d3.terminal_current("a")
        # floor: the 2D oracle's own conservation error (rounding noise of
        # the contact residuals at low current)
        floor = 10 * abs(I2 + d2.terminal_current("c")) * z[-1]
        assert abs(I3 - I2 * z[-1]) <= 1e-7 * abs(I2 * z[-1]) + floor


def _gated_3d(perm):
    """A gated p-n slab; perm=False: built along x with a z-normal gate;
    perm=True: the same device with x and z exchanged."""
    nx, ny, nz = 9, 5, 4
    a = np.linspace(0, 0.2e-4, nx)
    b = np.linspace(0, 0.1e-4, ny)
    c = np.linspace(0, 0.05e-4, nz)
    dopx = np.where(a < 0.08e-4, -1e17, 3e17)
    dop = np.broadcast_to(dopx[None, None, :], (nz, ny, nx)).copy()
    if not perm:
        dev = dd.Device3D(a, b, c, dop, models=MODELS)
        J, K = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        dev.add_contact("s", 0, J.ravel(), K.ravel())
        dev.add_contact("d", nx - 1, J.ravel(), K.ravel())
        I, J = np.meshgrid(np.arange(2, 6), np.arange(ny), indexing="ij")
        dev.add_gate("g", I.ravel(), J.ravel(), 0, 2e-7, -0.9, normal_axis="z")
    else:
        dev = dd.Device3D(c, b, a, np.transpose(dop, (2, 1, 0)).copy(), models=MODELS)
        I, J = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
        dev.add_contact("s", I.ravel(), J.ravel(), 0)
        dev.add_contact("d", I.ravel(), J.ravel(), nx - 1)
        K, J = np.meshgrid(np.arange(2, 6), np.arange(ny), indexing="ij")
        dev.add_gate("g", 0, J.ravel(), K.ravel(), 2e-7, -0.9, normal_axis="x")
    return dev


def test_axis_permutation_symmetry():
    # forward-biased gated diode; the equilibrated direct solve keeps the
    # two orderings' Newton sequences identical (SuperLU's unscaled solve
    # can differ by one iteration between orderings through rounding)
    d0, d1 = _gated_3d(False), _gated_3d(True)
    opts = oph.NewtonOptions(linsolve="direct_equilibrated")
    for dev in (d0, d1):
        dev.solve_equilibrium()
        dev.solve_bias({"d": -0.3, "g": 0.6}, opts)
    assert d0.last_iterations == d1.last_iterations
    psi1 = np.transpose(d1.psi, (2, 1, 0))
    assert np.abs(d0.psi - psi1).max() <= 1e-10 * np.abs(d0.psi).max()
    n1 = np.transpose(d1.n, (2, 1, 0))
    assert (np.abs(d0.n - n1) / d0.n).max() <= 1e-9
    Is = d0.terminal_current("d"), d1.terminal_current("d")
    floor = 10 * abs(Is[0] + d0.terminal_current("s"))
    assert abs(Is[0] - Is[1]) <= 1e-6 * abs(Is[0]) + floor


def test_fd_jacobian_3d():
    dev = _gated_3d(False)
    dev.solve_equilibrium()
    dev._set_voltages({"g": 0.5, "d": 0.2})
    rng = np.random.default_rng(7)
    psi = dev.psi + 0.02 * rng.standard_normal(dev.shape)
    n = dev.n * (1 + 0.01 * rng.standard_normal(dev.shape))
    p = dev.p * (1 + 0.01 * rng.standard_normal(dev.shape))
    cur = {"s": 0.0, "d": 0.2}
    F, J, *_ = dev._residual_jacobian(psi, n, p, cur)
    Jc = J.tocsc()
    u = np.stack([psi.ravel(), n.ravel(), p.ravel()], 1).ravel()
    worst = 0.0
    for c in rng.choice(3 * dev.N, 30, replace=False):
        step = 1e-7 * max(abs(u[c]), 1.0)
        u2 = u.copy(); u2[c] += step
        F2, *_ = dev._residual_jacobian(u2[0::3].reshape(dev.shape), u2[1::3].reshape(dev.shape),
                                        u2[2::3].reshape(dev.shape), cur)
        an = np.asarray(Jc[:, c].todense()).ravel()
        worst = max(worst, np.abs((F2 - F) / step - an).max() / (np.abs(an).max() + 1e-30))
    assert worst < 1e-3, worst


def test_terminal_current_conservation_3d():
    dev = _gated_3d(False)
    dev.solve_equilibrium()
    dev.solve_bias({"d": -0.4, "g": 0.4})          # forward-biased junction
    Is, Id = dev.terminal_current("s"), dev.terminal_current("d")
    assert abs(Id) > 1e-12
    assert abs(Is + Id) <= 1e-6 * max(abs(Is), abs(Id))
