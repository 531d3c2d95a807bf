"""The GPU solve_linear (tcadhip.linsolve, tcad_csr_solve) against the
reference's own bars for its Krylov methods (pytcad/tests/test_m22_linsolve.py):

* G3: GMRES / BiCGStab agree with the direct solve (SciPy spsolve, the
  reference's "direct") within 1e-6 relative, on random diagonally
  dominant systems and on a device's Newton system
  (test_m22_linsolve.py:58-95), with every preconditioner: node
  block-Jacobi with 3-blocks and 1-blocks, and level-scheduled ILU(0);
* G4: a solve that cannot converge RAISES LinearSolveError
  (test_m22_linsolve.py:98-121);
* G5: non-finite input raises LinearSolveError (test_m22_linsolve.py:124);
* the true relative residual reported is <= rtol on success;
* the same solve twice gives the same bits (fixed reduction order).
"""
import warnings

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

pytestmark = pytest.mark.gpu


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import _lib, linsolve
    return linsolve, _lib


def _random_dd_system(n, seed, density=0.02):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr()
    A = A + A.T
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0)
    return A.tocsr(), rng.standard_normal(n)


PRECONDS = [("block_jacobi", 3), ("block_jacobi", 1), ("ilu0", None)]


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
@pytest.mark.parametrize("precond,bs", PRECONDS)
@pytest.mark.parametrize("n,seed", [(30, 10), (300, 11), (2001, 12)])
def test_agrees_with_direct(method, precond, bs, n, seed):
    linsolve, _ = _mods()
    A, b = _random_dd_system(n, seed)
    ref = spsolve(A.tocsc(), b)
    x, info = linsolve.solve_linear(A, b, method=method, rtol=1e-10, block_size=bs,
                                    precond=precond)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    print(f"[csr] {method} {precond}/{bs} n={n}: {info['iterations']} its, resid "
          f"{info['residual']:.2e}, rel err {rel:.2e}", flush=True)
    assert info["converged"] and info["residual"] <= 1e-10
    assert rel <= 1e-6


def _device_system(kind):
    """The eliminated Newton system of a small 2D device at its warm start
    (the oracle's, pinned bit for bit to the reference's Device2D)."""
    from oracle import dd
    from oracle.linsolve import eliminate_csr
    from tcadhip import configs
    from oracle_build import oracle_device
    if kind == "diode":
        spec = configs.diode_2d(nx=30, ny=12, biases=[0.4])
        bias = {"anode": 0.4}
    else:
        spec = configs.mosfet_2d(nx=24, ny=16, vg_list=[1.0])
        bias = {"drain": 0.05, "gate": 1.0}
    dev = oracle_device(spec)
    dev.solve_equilibrium()
    dev._set_voltages(bias)
    if spec.models.get("surface_mobility"):
        dev._update_surface_mobility(dev.psi)
    cur = {nm: b.V for nm, b in dev.bcs.items() if isinstance(b, dd.DirichletBC)}
    psi, n, p = dev.psi.copy(), dev.n.copy(), dev.p.copy()
    for bc in dev.bcs.values():
        if isinstance(bc, dd.DirichletBC):
            psi[bc.idx], n[bc.idx], p[bc.idx] = dev._bc_contact_values(bc, bc.V)
    F, J, *_ = dev._residual_jacobian(psi, n, p, cur)
    Jd, rhs = eliminate_csr(J, -F, dev._dirichlet_rows)
    # equilibrated as the Newton path solves it: columns scaled by
    # (1, n, p) per node, rows by their inverse largest magnitude
    cs = np.stack([np.ones(n.size), n.ravel(), p.ravel()], axis=1).ravel()
    Js = Jd @ sp.diags(cs)
    rs = 1.0 / abs(Js).max(axis=1).toarray().ravel()
    return (sp.diags(rs) @ Js).tocsr(), rs * rhs


@pytest.mark.parametrize("kind", ["diode", "mosfet"])
@pytest.mark.parametrize("method,precond,bs", [("bicgstab", "block_jacobi", 3),
                                               ("gmres", "block_jacobi", 3),
                                               ("gmres", "ilu0", None)])
def test_device_newton_system(kind, method, precond, bs):
    linsolve, _lib = _mods()
    Jd, rhs = _device_system(kind)
    ref = spsolve(Jd.tocsc(), rhs)
    x, info = linsolve.solve_linear(Jd, rhs, method=method, rtol=1e-10, block_size=bs,
                                    precond=precond, maxiter=2000)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    print(f"[csr] {kind} {method} {precond}: {info['iterations']} its, rel err {rel:.2e}",
          flush=True)
    assert np.linalg.norm(Jd @ x - rhs) / np.linalg.norm(rhs) <= 1e-6
    assert rel <= 1e-6


def test_non_convergence_raises():
    linsolve, _lib = _mods()
    rng = np.random.default_rng(99)
    n = 198
    A = (sp.random(n, n, density=0.05, random_state=rng,
                   data_rvs=lambda k: rng.standard_normal(k)) + sp.eye(n) * 0.5).tocsr()
    b = rng.standard_normal(n)
    for method in ("gmres", "bicgstab"):
        with pytest.raises(_lib.LinearSolveError):
            linsolve.solve_linear(A, b, method=method, rtol=1e-14, maxiter=1, block_size=3)


def test_bad_input():
    linsolve, _lib = _mods()
    A, b = _random_dd_system(30, 1)
    bad = b.copy()
    bad[3] = np.nan
    with pytest.raises(_lib.LinearSolveError):
        linsolve.solve_linear(A, bad, method="gmres", block_size=3)
    with pytest.raises(ValueError):
        linsolve.solve_linear(A, b, method="direct")
    with pytest.raises(ValueError):
        linsolve.solve_linear(sp.eye(31).tocsr(), np.ones(31), method="gmres", block_size=3)


def test_deterministic():
    linsolve, _ = _mods()
    A, b = _random_dd_system(3000, 5)
    x1, i1 = linsolve.solve_linear(A, b, method="bicgstab", block_size=3)
    x2, i2 = linsolve.solve_linear(A, b, method="bicgstab", block_size=3)
    assert np.array_equal(x1, x2) and i1["iterations"] == i2["iterations"]
    y1, _ = linsolve.solve_linear(A, b, method="gmres", precond="ilu0")
    y2, _ = linsolve.solve_linear(A, b, method="gmres", precond="ilu0")
    assert np.array_equal(y1, y2)
