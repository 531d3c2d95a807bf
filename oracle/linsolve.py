"""Oracle linear algebra: symmetric Dirichlet elimination and the linear
solve contract.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning (tests/test_oracle_linsolve.py): `eliminate_csr` against a dense
transcription of the definition (rhs <- rhs - J g, constrained columns
zeroed, unit diagonal kept) and the identity that the eliminated system
has the same solution as the row-stamped one; `solve_linear("direct")`
is SciPy's spsolve (the reference's own call, linsolve.py:860); the
block-Jacobi Krylov path against direct within 1e-6 and block-Jacobi
exact on a block-diagonal matrix to 1e-10 (the reference's own bars,
pytcad/tests/test_m22_linsolve.py:58-95, 271-298).
"""
import warnings

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import LinearOperator, bicgstab, gmres, spsolve


class LinearSolveError(RuntimeError):
    """linsolve.py:344-352."""


def eliminate_csr(J, rhs, dirichlet):
    """dirichlet.py:161-191: symmetric Dirichlet elimination of an
    assembled matrix whose constrained rows are already e_k."""
    n = J.shape[0]
    rhs = np.array(rhs, dtype=float, copy=True)
    is_d = np.zeros(n, dtype=bool)
    is_d[dirichlet] = True
    if not is_d.any():
        return J.tocsr(), rhs
    g = np.zeros(n)
    g[is_d] = rhs[is_d]
    Jc = J.tocsr()
    rhs = rhs - Jc @ g
    rhs[is_d] = g[is_d]
    keep_cols = sp.diags((~is_d).astype(float), format="csr")
    Jout = Jc @ keep_cols + sp.diags(is_d.astype(float), format="csr")
    return Jout.tocsr(), rhs


def block_jacobi(A, block_size):
    """linsolve.py:363-427: invert each node's dense diagonal block."""
    n = A.shape[0]
    nblk = n // block_size
    Ac = A.tocsr()
    blocks = np.zeros((nblk, block_size, block_size))
    idx = np.arange(nblk) * block_size
    for r in range(block_size):
        for c in range(block_size):
            off = c - r
            d = Ac.diagonal(off)
            blocks[:, r, c] = d[idx + r] if off >= 0 else d[idx + c]
    inv = np.linalg.inv(blocks)

    def apply(x):
        return np.einsum("nij,nj->ni", inv, x.reshape(nblk, block_size)).ravel()

    return LinearOperator(A.shape, apply)


def solve_linear(A, b, method="direct", rtol=1e-10, maxiter=500,
                 restart=None, block_size=3):
    """linsolve.py:725-985 for methods direct / gmres / bicgstab.
    Iterative methods raise LinearSolveError unless the true relative
    residual is <= max(rtol, 1e-6) (linsolve.py:971-982)."""
    b = np.asarray(b, dtype=float)
    if not np.all(np.isfinite(b)) or not np.all(np.isfinite(A.data)):
        raise LinearSolveError("non-finite input")
    if method == "direct":
        with warnings.catch_warnings():
            warnings.simplefilter("error", sp.linalg.MatrixRankWarning)
            x = spsolve(A, b)
        if not np.all(np.isfinite(x)):
            raise LinearSolveError("direct solve returned non-finite result")
        return x, {"method": "direct", "iterations": 1}
    A = A.tocsr()
    M = block_jacobi(A, block_size)
    its = [0]

    def cb(_):
        its[0] += 1

    kw = dict(rtol=rtol, atol=0.0, maxiter=maxiter, M=M, callback=cb)
    if method == "gmres":
        kw["callback_type"] = "pr_norm"
        kw["restart"] = min(restart or 100, A.shape[0])
        x, code = gmres(A, b, **kw)
    elif method == "bicgstab":
        x, code = bicgstab(A, b, **kw)
    else:
        raise ValueError(method)
    resid = np.linalg.norm(A @ x - b) / max(np.linalg.norm(b), 1e-300)
    if code != 0 or not np.all(np.isfinite(x)) or resid > max(rtol, 1e-6):
        raise LinearSolveError(f"{method} did not converge (resid {resid:.3e})")
    return x, {"method": method, "iterations": its[0], "residual": resid}
