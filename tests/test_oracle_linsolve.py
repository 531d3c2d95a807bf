"""Pins of the oracle's linear algebra (oracle/linsolve.py) by the
reference's own bars and by what the definitions fix:

* eliminate_csr against a dense transcription of dirichlet.py:161-191
  (rhs <- rhs - J g, constrained columns zeroed, unit diagonal), and the
  identity that the eliminated system has the row-stamped system's
  solution;
* solve_linear("direct") is SciPy's spsolve bit for bit
  (pytcad/tests/test_m22_linsolve.py:33-55);
* the node block-Jacobi preconditioner equals each node's dense block
  solve to 1e-10 (pytcad/tests/test_m22_linsolve.py:271-298), on non-symmetric
  blocks so a transposed block fails;
* GMRES and BiCGStab with it agree with the direct solve within 1e-6 on
  random diagonally dominant systems and on a real device Jacobian
  (pytcad/tests/test_m22_linsolve.py:58-95);
* a Krylov solve that cannot converge raises LinearSolveError, never
  returns a half-solved state (pytcad/tests/test_m22_linsolve.py:98-121).
"""
import warnings

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from oracle import dd
from oracle import linsolve as ols


def _random_dd_system(n, seed, density=0.02):
    """A random, strongly diagonally dominant sparse system (the shape of
    the reference's own parity systems)."""
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr()
    A = A + A.T
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0)
    return A.tocsr(), rng.standard_normal(n)


def test_eliminate_csr_matches_definition():
    rng = np.random.default_rng(5)
    n = 60
    A, b = _random_dd_system(n, 7, density=0.1)
    rows = np.array([0, 1, 2, 17, 30, 59])
    Ad = A.toarray()
    for r in rows:                       # contact rows stamped e_k
        Ad[r, :] = 0.0
        Ad[r, r] = 1.0
    b[rows] = rng.standard_normal(rows.size)
    J = sp.csr_matrix(Ad)
    Jd, rhs = ols.eliminate_csr(J, b, rows)
    g = np.zeros(n)
    g[rows] = b[rows]
    ref_rhs = b - Ad @ g
    ref_rhs[rows] = g[rows]
    ref = Ad.copy()
    ref[:, rows] = 0.0
    ref[rows, rows] = 1.0
    assert np.array_equal(Jd.toarray(), ref)
    assert np.allclose(rhs, ref_rhs, rtol=0, atol=1e-13)
    # same solution as the row-stamped system
    x1 = np.linalg.solve(Ad, b)
    x2 = np.linalg.solve(Jd.toarray(), rhs)
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x1).max()


@pytest.mark.parametrize("n,seed", [(5, 0), (51, 1), (402, 2)])
def test_direct_is_spsolve_bit_for_bit(n, seed):
    A, b = _random_dd_system(n, seed, density=0.2)
    x, info = ols.solve_linear(A, b, method="direct")
    assert np.array_equal(x, spsolve(A, b))
    assert info["method"] == "direct"


def test_block_jacobi_is_the_exact_node_block_inverse():
    rng = np.random.default_rng(3)
    n_nodes, bs = 200, 3
    n = n_nodes * bs
    A = sp.lil_matrix((n, n))
    for k in range(n_nodes):
        A[3 * k:3 * k + 3, 3 * k:3 * k + 3] = rng.standard_normal((bs, bs)) * 3 + np.eye(bs) * 10
    for _ in range(50):                   # off-block entries the preconditioner ignores
        i, j = rng.integers(0, n, size=2)
        if i // 3 != j // 3:
            A[i, j] += rng.standard_normal()
    A = A.tocsr()
    M = ols.block_jacobi(A, bs)
    x = rng.standard_normal(n)
    y = M.matvec(x)
    Ad = A.toarray()
    y_ref = np.concatenate([np.linalg.solve(Ad[3 * k:3 * k + 3, 3 * k:3 * k + 3],
                                            x[3 * k:3 * k + 3]) for k in range(n_nodes)])
    assert np.abs(y - y_ref).max() < 1e-10


@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
@pytest.mark.parametrize("n,seed", [(30, 10), (300, 11), (2001, 12)])
def test_krylov_agrees_with_direct(method, n, seed):
    A, b = _random_dd_system(n, seed)
    ref = spsolve(A.tocsc(), b)
    x, info = ols.solve_linear(A, b, method=method, rtol=1e-10)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    assert rel <= 1e-6, (method, n, rel)
    assert info["residual"] <= 1e-10


@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
def test_krylov_on_a_device_jacobian(method):
    """The eliminated Newton system of a small 2D diode at 0.3 V forward
    bias, from the fixture-pinned residual/Jacobian."""
    x = np.linspace(0, 1e-4, 14)
    y = np.linspace(0, 0.4e-4, 6)
    dop = np.tile(np.where(x < 0.5e-4, -1e17, 1e18), (6, 1))
    dev = dd.Device2D(x, y, dop)
    dev.add_contact("a", 0, np.arange(6))
    dev.add_contact("c", 13, np.arange(6))
    dev.solve_equilibrium()
    dev._set_voltages({"a": 0.3})
    F, J, *_ = dev._residual_jacobian(dev.psi, dev.n, dev.p, {"a": 0.3, "c": 0.0})
    Jd, rhs = ols.eliminate_csr(J, -F, dev._dirichlet_rows)
    ref = spsolve(Jd.tocsc(), rhs)
    xs, info = ols.solve_linear(Jd, rhs, method=method, rtol=1e-10)
    assert np.linalg.norm(xs - ref) / np.linalg.norm(ref) <= 1e-6


def test_non_convergence_raises():
    n = 200
    rng = np.random.default_rng(99)
    A = sp.random(n, n, density=0.05, random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr() + sp.eye(n) * 1e-3
    # a node block that is singular would make the preconditioner itself
    # fail; keep every 3x3 diagonal block invertible
    A = A + sp.eye(n) * 0.5
    A = A[:198, :198].tocsr()
    b = rng.standard_normal(198)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        for method in ("gmres", "bicgstab"):
            with pytest.raises(ols.LinearSolveError):
                ols.solve_linear(A, b, method=method, rtol=1e-14, maxiter=1)
