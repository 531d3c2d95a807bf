"""GPU drop-in for pytcad.linsolve.solve_linear's Krylov methods.

`solve_linear(A, b, *, method, rtol, atol, maxiter, x0, restart,
block_size, precond)` has the reference's signature and contract
(pytcad/pytcad/linsolve.py:725-985): "gmres" and "bicgstab" precondition
with node block-Jacobi when `block_size` is given (linsolve.py:363-427)
and otherwise with an incomplete factorisation (the reference's ILU
fallback, linsolve.py:429-585; here level-scheduled ILU(0)); the solution
is accepted when ||A x - b|| / ||b|| <= max(rtol, 1e-6); a solve that does
not get there RAISES LinearSolveError, never returns the unconverged
iterate; non-finite input raises LinearSolveError too (linsolve.py:
344-360).  The returned info holds "method", "iterations", "converged",
"residual" and "backend" ("hip").

The system is solved by libtcadhip's tcad_csr_solve on the GPU (FP64,
deterministic reductions).  No direct factorisation runs on the GPU:
method="direct" raises ValueError; the reference's direct path stays
SciPy's SuperLU on the host.
"""
import ctypes

import numpy as np

from . import _lib

_METHODS = ("gmres", "bicgstab")
_PRECOND = ("auto", "block_jacobi", "ilu0")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("tcadhip needs a HIP (ROCm) GPU; torch.cuda.is_available() is False")
    return torch


def _prepare(A, b):
    import scipy.sparse as sp
    A = sp.csr_matrix(A) if not sp.issparse(A) else A.tocsr()
    if not A.has_canonical_format:
        A = A.copy()
        A.sum_duplicates()                 # sorts the indices as well
    b = np.asarray(b, dtype=np.float64).ravel()
    n = A.shape[0]
    if A.shape != (n, n) or b.size != n:
        raise ValueError(f"A must be square and b of length {n}; got {A.shape}, {b.size}")
    if n >= 2 ** 31 or A.nnz >= 2 ** 31:
        raise ValueError("the GPU CSR solve takes int32 indices (n, nnz < 2^31)")
    if not np.all(np.isfinite(A.data)) or not np.all(np.isfinite(b)):
        raise _lib.LinearSolveError("solve_linear: non-finite entries in A or b")
    return A, b


def solve_linear(A, b, *, method="bicgstab", rtol=1e-10, atol=0.0, maxiter=500, x0=None,
                 restart=None, block_size=None, precond="auto", stream=None, noalloc=False):
    """Solve A x = b on the GPU; returns (x, info) with x a NumPy array.

    noalloc=True runs the allocation-free entry point
    (tcad_csr_solve_noalloc: block-Jacobi only, a fixed 64 MiB device
    workspace per GPU) instead of tcad_csr_solve."""
    if method == "direct":
        raise ValueError("method='direct' is a host factorisation (SciPy SuperLU); the GPU "
                         "solve_linear runs 'gmres' and 'bicgstab'")
    if method not in _METHODS:
        raise ValueError(f"unknown method {method!r}; choose from {_METHODS}")
    if precond not in _PRECOND:
        raise ValueError(f"unknown precond {precond!r}; choose from {_PRECOND}")
    torch = _torch()
    A, b = _prepare(A, b)
    n = A.shape[0]
    if precond == "ilu0" or (precond == "auto" and block_size is None):
        prec, bs = _lib.CSR_PREC_ILU0, 1
    else:
        prec, bs = _lib.CSR_PREC_BLOCK_JACOBI, int(block_size or 1)
    bnorm = float(np.linalg.norm(b))
    # atol folds into the relative target (the Krylov test is ||r|| <= max(rtol ||b||, atol))
    rt = max(rtol, atol / bnorm) if (atol and bnorm > 0) else rtol
    dev = torch.device("cuda")
    ptr = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    idx = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    val = torch.from_numpy(A.data.astype(np.float64)).to(dev)
    bd = torch.from_numpy(b).to(dev)
    x = (torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None else
         torch.from_numpy(np.asarray(x0, dtype=np.float64).ravel().copy()).to(dev))
    s = stream if stream is not None else torch.cuda.current_stream()
    torch.cuda.current_stream().synchronize()
    its = ctypes.c_int32()
    res = ctypes.c_double()
    fn = _lib.lib().tcad_csr_solve_noalloc if noalloc else _lib.lib().tcad_csr_solve
    meth = _lib.CSR_GMRES if method == "gmres" else _lib.CSR_BICGSTAB
    rc = fn(n, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), bd.data_ptr(), x.data_ptr(),
            meth, prec, bs, float(rt), int(maxiter), int(restart or 0),
            ctypes.c_void_p(s.cuda_stream), ctypes.byref(its), ctypes.byref(res))
    if rc == _lib.TCAD_E_NOT_CONVERGED:
        raise _lib.LinearSolveError(
            f"{method} did not converge within {maxiter} iterations (relative residual="
            f"{res.value:.3e}, target rtol={rtol:.3e}) -- refusing to return the unconverged "
            f"iterate")
    _lib.check(rc, "tcad_csr_solve_noalloc" if noalloc else "tcad_csr_solve")
    return x.cpu().numpy(), {"method": method, "iterations": int(its.value), "converged": True,
                             "residual": float(res.value), "backend": "hip"}
