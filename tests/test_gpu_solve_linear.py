"""tcad_solve_linear: the FP64-refined Krylov solve of the last assembled
Newton system (linsolve.solve_linear's iterative contract,
pytcad/pytcad/linsolve.py:725-985), with each preconditioner the library
offers: the reference's node block-Jacobi and the 16-node tile inverses,
applied in red-black order (the default) and as block-Jacobi
(TCAD_TUNE_TILE_RB 0), with the FP64 and the FP32-stored matrix.

The system is the reference's own: the Newton matrix at the fixtures'
perturbed states (tests/golden/make_golden*.py), assembled on the GPU
(its F and raw J equal the reference's to 1e-12, test_gpu_parity.py).
Checked per solve, stated tolerances:
  * scaled residual ||R (b - Jd du)|| / ||R b|| <= 1e-9 (the solve aims at
    1e-10 on the row-equilibrated system);
  * du against the oracle's direct solve of the same equilibrated system
    (oracle.dd.equilibrated_direct): max |du - du_ref| per unknown kind
    (psi, n/n, p/p scaled by the column scale) <= 1e-6 of that kind's max.
"""
import ctypes

import numpy as np
import pytest

from golden.fixtures import load
from oracle import dd

pytestmark = pytest.mark.gpu

CASES = ["mosfet_16x12_fj.npz", "trigate_12x10x8_fj.npz", "slab_zgate_12x9x10_fj.npz"]
MODES = [("node", False), ("tile", False), ("tile-jacobi", False), ("node", True), ("tile", True),
         ("tile-jacobi", True)]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", CASES)
def test_solve_linear_preconditioners(name):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import _lib
    from tcadhip.device import device_from_spec, stencil_to_csr
    spec, d = load(name)
    dev = device_from_spec(spec)
    volts = dict(zip([str(k) for k in d["volt_keys"]], d["volt_vals"]))
    psi, n, p = d["in_psi"], d["in_n"], d["in_p"]
    dev._set_voltages(volts)
    if spec.models["surface_mobility"]:
        dev._update_surface_mobility(psi)
    B = dev._gpu()
    B.set_bias()
    N = dev.N
    L = _lib.lib()
    # the state stays in these buffers: tcad_solve_linear unscales du with
    # the n and p of the last assembly
    st = [torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).ravel()).to("cuda")
          for v in (psi, n, p)]
    Fd = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    _lib.check(L.tcad_assemble(B.handle, st[0].data_ptr(), st[1].data_ptr(), st[2].data_ptr(),
                               Fd.data_ptr(), None), "tcad_assemble")
    ne = ctypes.c_int32()
    _lib.check(L.tcad_export_jacobian(B.handle, None, None, None, ctypes.byref(ne)),
               "tcad_export_jacobian")
    blocks = np.zeros(ne.value * N)
    bs = np.zeros(3 * N)
    rs = np.zeros(3 * N)
    _lib.check(L.tcad_export_jacobian(B.handle, _lib.dptr(blocks), _lib.dptr(rs), _lib.dptr(bs),
                                      None), "tcad_export_jacobian")
    Jd = stencil_to_csr(blocks.reshape(ne.value, N), rs, np.ravel(n), np.ravel(p), dev._s)
    # planes (c * N + k) -> interleaved unknowns (3 k + c)
    R = rs.reshape(3, N).T.ravel()
    rhs = (bs.reshape(3, N).T.ravel()) / R
    du_ref = dd.equilibrated_direct(Jd, rhs, n, p)
    cs = np.stack([np.ones(N), np.ravel(n), np.ravel(p)], 1).ravel()
    y_ref = (du_ref / cs).reshape(N, 3)
    for kind, mixed in MODES:
        mode = _lib.LIN_BICGSTAB_MIXED if mixed else _lib.LIN_BICGSTAB_F64
        mode |= _lib.PRECOND_NODE if kind == "node" else 0
        _lib.check(L.tcad_set_tuning(B.handle, _lib.TUNE_TILE_RB,
                                     0.0 if kind == "tile-jacobi" else 1.0), "tcad_set_tuning")
        du = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        its = (ctypes.c_int32 * 1)()
        res = (ctypes.c_double * 1)()
        _lib.check(L.tcad_solve_linear(B.handle, mode, 1e-10, 5000, du.data_ptr(), its,
                                       res), "tcad_solve_linear")
        got = du.cpu().numpy().reshape(3, N).T.ravel()
        r = R * (rhs - Jd @ got)
        rel = np.linalg.norm(r) / np.linalg.norm(R * rhs)
        y = (got / cs).reshape(N, 3)
        err = [np.abs(y[:, c] - y_ref[:, c]).max() / max(np.abs(y_ref[:, c]).max(), 1e-300)
               for c in range(3)]
        _say(f"[solve_linear] {name} {kind}{' fp32' if mixed else ''}: {its[0]} Krylov, "
             f"residual {rel:.2e} (reported {res[0]:.2e}), du vs direct {max(err):.2e}")
        assert rel <= 1e-9
        assert max(err) <= 1e-6
