"""ctypes binding of libtcadhip.so (the C ABI declared in include/tcad_hip.h).

The library is loaded from this package directory.  There is no CPU
fallback: if the shared library is missing, or no HIP device is present,
every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TCADHIP_LIB lets a developer point at another build of the same library
# (e.g. a launch-shape variant); the default is the in-tree build.
LIB_PATH = os.environ.get("TCADHIP_LIB") or os.path.join(HERE, "libtcadhip.so")

TCAD_OK = 0
TCAD_E_INVALID = -1
TCAD_E_INDEX = -2
TCAD_E_NONFINITE = -3
TCAD_E_NOT_CONVERGED = -4
TCAD_E_HIP = -5
TCAD_E_NO_DEVICE = -6

LIN_BICGSTAB_F64 = 0
LIN_BICGSTAB_MIXED = 1
LIN_FORCING = 0x100
PRECOND_NODE = 0x200

TUNE_TILE_ROWS = 1
TUNE_FORCING_SLOPE = 2
TUNE_FORCING_MAX = 3
TUNE_FORCING_FP32 = 4
TUNE_GRAPH = 5
TUNE_IR_CONTINUE = 6
TUNE_FP32_INNER = 7
TUNE_FORCING_POW = 8
TUNE_TILE_LAG = 9
TUNE_LINEAR_FALLBACK = 10
TUNE_FORCING_PREDICTED = 11
TUNE_TILE_RB = 12

CSR_BICGSTAB = 0
CSR_GMRES = 1
CSR_PREC_BLOCK_JACOBI = 0
CSR_PREC_ILU0 = 1

NEWTON_OK = 0
NEWTON_MAX_ITER = 1
NEWTON_LINEAR_FAIL = 2
NEWTON_NONFINITE = 3

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)


class LinearSolveError(RuntimeError):
    """Raised when a linear solve fails (pytcad/pytcad/linsolve.py:344-352)."""


class DeviceDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("h_hx", c_double_p), ("h_hy", c_double_p), ("h_hz", c_double_p),
        ("h_dvx", c_double_p), ("h_dvy", c_double_p), ("h_dvz", c_double_p),
        ("VT", ctypes.c_double), ("Ns", ctypes.c_double), ("LD", ctypes.c_double),
        ("J0", ctypes.c_double), ("R0", ctypes.c_double),
        ("Cn_auger", ctypes.c_double), ("Cp_auger", ctypes.c_double),
        ("srh", ctypes.c_int32), ("auger", ctypes.c_int32),
        ("surface_mobility", ctypes.c_int32), ("T", ctypes.c_double),
        ("h_mun_row0", c_double_p), ("h_mun_row1", c_double_p),
        ("h_mup_row0", c_double_p), ("h_mup_row1", c_double_p),
        ("ncontacts", ctypes.c_int32), ("h_contact_ptr", c_int32_p),
        ("h_contact_nodes", c_int32_p), ("h_contact_lnr", c_double_p),
        ("h_contact_n0", c_double_p), ("h_contact_p0", c_double_p),
        ("ngates", ctypes.c_int32), ("h_gate_ptr", c_int32_p), ("h_gate_nodes", c_int32_p),
        ("h_gate_kw", c_double_p), ("h_gate_psib", c_double_p),
        ("node_offset", ctypes.c_int64),
    ]


class StaticFields(ctypes.Structure):
    _fields_ = [
        ("d_C", ctypes.c_void_p), ("d_nie_s", ctypes.c_void_p), ("d_nie", ctypes.c_void_p),
        ("d_tau_n", ctypes.c_void_p), ("d_tau_p", ctypes.c_void_p), ("d_dV", ctypes.c_void_p),
        ("d_an", ctypes.c_void_p * 3), ("d_ap", ctypes.c_void_p * 3),
        ("d_dlnnie", ctypes.c_void_p * 3),
    ]


_lib = None


def lib():
    """Load libtcadhip.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python tcad-dev_amd/build.py` "
            "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sigs = {
        "tcad_batch_create": ([ctypes.POINTER(DeviceDesc), ctypes.c_int32,
                               ctypes.POINTER(StaticFields), vp, ctypes.POINTER(vp)], ctypes.c_int),
        "tcad_batch_destroy": ([vp], ctypes.c_int),
        "tcad_batch_nodes": ([vp], ctypes.c_int64),
        "tcad_launch_count": ([vp], ctypes.c_int64),
        "tcad_set_bias": ([vp, c_double_p, c_double_p, c_double_p], ctypes.c_int),
        "tcad_assemble": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "tcad_predict_state": ([vp, vp, vp, vp, vp, vp, vp, c_double_p], ctypes.c_int),
        "tcad_export_jacobian": ([vp, c_double_p, c_double_p, c_double_p, c_int32_p],
                                 ctypes.c_int),
        "tcad_residual_jacobian": ([vp, vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "tcad_solve_linear": ([vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, vp,
                               c_int32_p, c_double_p], ctypes.c_int),
        "tcad_solve_equilibrium": ([vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int32, c_int32_p, vp, vp, vp, c_int32_p,
                                    c_int32_p, c_int32_p], ctypes.c_int),
        "tcad_solve_bias": ([vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int32, ctypes.c_double, ctypes.c_int32, c_int32_p,
                             vp, vp, vp, c_int32_p, c_int32_p, c_int32_p, c_int32_p],
                            ctypes.c_int),
        "tcad_solve_queue": ([vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                              c_int32_p, c_double_p, c_double_p, vp, vp, vp, c_int32_p,
                              c_int32_p, c_int32_p, c_int32_p, c_double_p], ctypes.c_int),
        "tcad_transient_step": ([vp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int32, ctypes.c_double, ctypes.c_int32, c_int32_p,
                                 c_double_p, ctypes.c_double, vp, vp, vp, vp, vp, vp, vp,
                                 c_int32_p, c_int32_p, c_int32_p], ctypes.c_int),
        "tcad_continuity_residuals": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "tcad_y_parameters": ([vp, c_int32_p, ctypes.c_int32, c_double_p, ctypes.c_double,
                               ctypes.c_int32, vp, vp, vp, c_double_p, c_int32_p,
                               ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "tcad_terminal_currents": ([vp, vp, vp, vp, c_double_p], ctypes.c_int),
        "tcad_edge_currents": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "tcad_surface_mobility": ([vp, vp], ctypes.c_int),
        "tcad_time_kernel": ([vp, ctypes.c_int32, ctypes.c_int32, c_double_p], ctypes.c_int),
        "tcad_probe_kernel": ([vp, ctypes.c_int32], ctypes.c_int),
        "tcad_set_tuning": ([vp, ctypes.c_int32, ctypes.c_double], ctypes.c_int),
        "tcad_csr_solve": ([ctypes.c_int32, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, vp,
                            c_int32_p, c_double_p], ctypes.c_int),
        "tcad_csr_solve_noalloc": ([ctypes.c_int32, vp, vp, vp, vp, vp, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_int32, ctypes.c_int32, vp, c_int32_p, c_double_p],
                                   ctypes.c_int),
        "tcad_probe_read": ([vp, ctypes.c_double, ctypes.POINTER(ctypes.c_int64), c_double_p],
                            ctypes.c_int),
        "tcad_last_error": ([], ctypes.c_char_p),
        "tcad_version": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc, what):
    """Map a return code to the reference's exception classes
    (pytcad/core/bindings/module.cpp:109-140)."""
    if rc == TCAD_OK:
        return
    msg = f"{what}: {lib().tcad_last_error().decode()}"
    if rc == TCAD_E_INVALID:
        raise ValueError(msg)
    if rc == TCAD_E_INDEX:
        raise IndexError(msg)
    if rc in (TCAD_E_NONFINITE, TCAD_E_NOT_CONVERGED):
        raise LinearSolveError(msg)
    raise RuntimeError(msg)


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int32_p)


def exported_symbols():
    """Names declared in include/tcad_hip.h that the library must export."""
    return ["tcad_batch_create", "tcad_batch_destroy", "tcad_batch_nodes", "tcad_set_bias",
            "tcad_assemble", "tcad_predict_state", "tcad_export_jacobian", "tcad_residual_jacobian",
            "tcad_solve_linear",
            "tcad_solve_equilibrium", "tcad_solve_bias", "tcad_terminal_currents",
            "tcad_edge_currents", "tcad_surface_mobility", "tcad_time_kernel",
            "tcad_transient_step", "tcad_continuity_residuals", "tcad_y_parameters",
            "tcad_solve_queue",
            "tcad_probe_kernel", "tcad_probe_read", "tcad_set_tuning", "tcad_csr_solve", "tcad_csr_solve_noalloc",
            "tcad_launch_count", "tcad_last_error", "tcad_version"]
