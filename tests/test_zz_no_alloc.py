"""The allocation-free entry point tcad_csr_solve_noalloc against
tcad_csr_solve, through the C ABI: two calls in a row, the second on a
smaller system, each compared with the allocating entry point on the same
input.  Both run the same kernels in the same order, so the bar is bit
equality of x, iteration count and residual (stronger than the 1e-6 of
the Krylov solve itself).  Also: sizes beyond the fixed workspace and the
ILU(0) preconditioner are refused with a clear error."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _sys(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=min(1.0, 12.0 / n), random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr()
    A = A + A.T
    A = (A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0)).tocsr()
    A.sort_indices()
    return A, rng.standard_normal(n)


def _call(fn, A, b, method, bs=3, restart=0, rtol=1e-10, maxiter=500):
    import torch
    from tcadhip import _lib
    n = A.shape[0]
    dev = torch.device("cuda")
    ptr = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    idx = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    val = torch.from_numpy(A.data.astype(np.float64)).to(dev)
    bd = torch.from_numpy(np.asarray(b, dtype=np.float64)).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    its = ctypes.c_int32()
    res = ctypes.c_double()
    rc = fn(n, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), bd.data_ptr(), x.data_ptr(),
            method, _lib.CSR_PREC_BLOCK_JACOBI, bs, rtol, maxiter, restart,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.byref(its),
            ctypes.byref(res))
    return rc, x.cpu().numpy(), its.value, res.value


@pytest.mark.parametrize("method_name", ["bicgstab", "gmres"])
def test_noalloc_equals_allocating(method_name):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import _lib
    L = _lib.lib()
    method = _lib.CSR_BICGSTAB if method_name == "bicgstab" else _lib.CSR_GMRES
    for n, seed in ((30003, 1), (6000, 2)):          # second call smaller
        A, b = _sys(n, seed)
        print(f"[no-alloc] {method_name} n={n}", flush=True)
        rc0, x0, i0, r0 = _call(L.tcad_csr_solve, A, b, method, restart=40)
        rc1, x1, i1, r1 = _call(L.tcad_csr_solve_noalloc, A, b, method, restart=40)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1, L.tcad_last_error())
        assert np.array_equal(x0, x1), np.abs(x0 - x1).max()
        assert i0 == i1 and r0 == r1
        assert r1 <= 1e-10
        assert np.linalg.norm(A @ x1 - b) / np.linalg.norm(b) <= 1e-9


def test_noalloc_refusals():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import _lib
    L = _lib.lib()
    # GMRES(100) on 200k unknowns needs ~170 MB of workspace: refused
    A = sp.identity(200001, format="csr")
    b = np.ones(200001)
    rc, *_ = _call(L.tcad_csr_solve_noalloc, A, b, _lib.CSR_GMRES, restart=100)
    assert rc == _lib.TCAD_E_INVALID
    assert b"fixed scratch" in L.tcad_last_error()
    # ILU(0) is refused
    import torch as _t
    n = 30
    A, b = _sys(n, 3)
    dev = _t.device("cuda")
    ptr = _t.from_numpy(A.indptr.astype(np.int32)).to(dev)
    idx = _t.from_numpy(A.indices.astype(np.int32)).to(dev)
    val = _t.from_numpy(A.data).to(dev)
    bd = _t.from_numpy(b).to(dev)
    x = _t.zeros(n, dtype=_t.float64, device=dev)
    rc = L.tcad_csr_solve_noalloc(n, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(),
                                  bd.data_ptr(), x.data_ptr(), _lib.CSR_BICGSTAB,
                                  _lib.CSR_PREC_ILU0, 1, 1e-10, 100, 0, None, None, None)
    assert rc == _lib.TCAD_E_INVALID
    assert b"ILU" in L.tcad_last_error()
