"""tcadhip: MI355X (gfx950) drift-diffusion Newton solve behind the
reference's Device2D / Device3D operator API.

All compute runs in libtcadhip.so (hand-written HIP kernels, C ABI in
include/tcad_hip.h); there is no CPU fallback.
"""
from . import configs, mesh, physics
from ._lib import LinearSolveError
from .device import Device2D, Device3D, DeviceSetup, GPUBatch, NewtonOptions, device_from_spec
from .mesh import Mesh2D, Mesh3D
from .physics import Models, SILICON, Semiconductor

__all__ = ["Device2D", "Device3D", "DeviceSetup", "GPUBatch", "NewtonOptions", "Models",
           "SILICON", "Semiconductor", "Mesh2D", "Mesh3D", "LinearSolveError",
           "device_from_spec", "configs", "mesh", "physics"]
