"""Workload definitions: the diode, MOSFET, tri-gate, batched-diode and
serving workloads, as plain data (`DeviceSpec`) that both the GPU
device classes and the CPU oracle are built from, so both sides see
identical input arrays.

Structure builders follow the reference's closed-form doping and
contact/gate placement:
  mosfet_doping  pytcad/pytcad/mosfet.py:26-54 (Gaussian depth x
                 erfc lateral source/drain on a -N_A background)
  build_mosfet   pytcad/pytcad/mosfet.py:57-87 (source/drain contacts
                 on row 0, body on the last row, gate over the channel)
  build_finfet3d pytcad/pytcad/finfet3d.py:61-152 (top gate y-normal
                 over all k, side gates z-normal excluding j = 0, body
                 on j = Ny-1, source/drain over the full cross-section)
  flatband_voltage pytcad/pytcad/moscap.py:63-90
  id_vg_sweep    pytcad/pytcad/mosfet.py:90-103 (warm-started sweep)

Choices the workload statements leave open are fixed here and used
identically by the oracle and the GPU path:
  * diode: anode = left edge (p side) carries the forward bias, cathode
    = right edge at 0 V; constant mobility (mu_max), SRH only, no BGN,
    tau_n = tau_p = 1e-7 s via tau_Nref = 1e300 (lifetime_scharfetter
    then returns tau0 exactly);
  * MOSFET: L_sd = 50 nm, depth 100 nm, x refined at the two gate
    edges, y geometric from a 0.2 nm first cell, n+poly gate;
  * tri-gate: L_g = 20 nm, L_sd = 15 nm, sigma_y = 100 nm (S/D nearly
    uniform through the 30 nm fin height), n+poly gate, Models()
    defaults;
  * batched diodes: 1.0 x 0.5 um domain, seed 20261203, default
    Models(); each target bias is reached by continuation steps of at
    most 0.1 V from equilibrium.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
from scipy.special import erfc

from .mesh import Mesh2D, Mesh3D, focused_mesh, geometric_mesh, uniform_mesh
from .physics import SILICON, flatband_voltage

NM = 1e-7        # cm


@dataclass
class ContactSpec:
    name: str
    idx: Tuple[np.ndarray, ...]      # array-order index arrays ((j,i) or (k,j,i))
    V: float = 0.0


@dataclass
class GateSpec:
    name: str
    idx: Tuple[np.ndarray, ...]
    tox_cm: float
    Vfb: float
    normal_axis: str = "y"
    Vg: float = 0.0


@dataclass
class DeviceSpec:
    coords: List[np.ndarray]                  # [x, y] or [x, y, z] in cm
    doping: np.ndarray                        # array order (Ny,Nx)/(Nz,Ny,Nx)
    Ntotal: Optional[np.ndarray]
    models: Dict[str, bool]
    material: Dict[str, float]
    contacts: List[ContactSpec]
    gates: List[GateSpec]
    sweep: List[Dict[str, float]] = field(default_factory=list)
    name: str = ""
    T: float = 300.0

    @property
    def dim(self):
        return len(self.coords)

    @property
    def shape(self):
        return tuple(c.size for c in self.coords[::-1])

    @property
    def N(self):
        return int(np.prod(self.shape))


def mosfet_doping(x, y, Lsd, Lg, Na, Nsd_peak, sigma_y, sigma_lat):
    """pytcad/pytcad/mosfet.py:26-54: returns (doping, Ntotal), (Ny, Nx)."""
    def sd(x_edge, high_side):
        vertical = np.exp(-(y[:, None]) ** 2 / (2.0 * sigma_y ** 2))
        s = (x_edge - x) if high_side == "left" else (x - x_edge)
        lateral = 0.5 * erfc(-s[None, :] / (np.sqrt(2.0) * sigma_lat))
        return Nsd_peak * vertical * lateral
    channel = -Na * np.ones((y.size, x.size))
    source = sd(Lsd, "left")
    drain = sd(Lsd + Lg, "right")
    return channel + source + drain, np.abs(channel) + source + drain


def diode_2d(nx=60, ny=40, Lx=1.0e-4, Ly=0.5e-4, xj_frac=0.5, Na=1e17, Nd=1e18,
             biases=None, models=None, material=None, name="diode"):
    """2D p-n diode, p side on the left, uniform mesh, ohmic left/right."""
    x = uniform_mesh(Lx, nx)
    y = uniform_mesh(Ly, ny)
    dop1 = np.where(x < xj_frac * Lx, -Na, Nd)
    doping = np.tile(dop1, (ny, 1))
    j = np.arange(ny)
    contacts = [ContactSpec("anode", (j, np.zeros(ny, dtype=int))),
                ContactSpec("cathode", (j, np.full(ny, nx - 1)))]
    if biases is None:
        biases = [round(0.05 * k, 10) for k in range(13)]
    return DeviceSpec([x, y], doping, None,
                      models or dict(doping_mobility=False, srh=True, auger=False,
                                     bgn=False, surface_mobility=False),
                      material if material is not None else
                      dict(tau_n0=1e-7, tau_p0=1e-7, tau_Nref=1e300),
                      contacts, [], [{"anode": float(v)} for v in biases], name)


def config1_diode():
    """60x40 diode, forward bias 0 -> 0.6 V in 0.05 V steps."""
    return diode_2d(name="diode_60x40")


def mosfet_2d(nx=200, ny=120, Lg=100 * NM, Lsd=50 * NM, depth=100 * NM, tox=2 * NM,
              Na=1e17, Npeak=1e20, sigma=10 * NM, h_if=0.2 * NM, Vd=0.05,
              vg_list=None, surface_mobility=True, name="mosfet"):
    """2D n-MOSFET (pytcad/pytcad/mosfet.py:57-87 layout)."""
    L = 2 * Lsd + Lg
    hx = L / (nx - 1)
    x = focused_mesh(L, [Lsd, Lsd + Lg], nx, h_min=0.5 * hx, h_max=1.5 * hx, ratio=1.15)
    y = geometric_mesh(depth, ny, h_if)
    doping, ntot = mosfet_doping(x, y, Lsd, Lg, Na, Npeak, sigma, sigma)
    i_s = np.where(x <= Lsd)[0]
    i_g = np.where((x > Lsd) & (x < Lsd + Lg))[0]
    i_d = np.where(x >= Lsd + Lg)[0]
    contacts = [ContactSpec("source", (np.zeros_like(i_s), i_s)),
                ContactSpec("drain", (np.zeros_like(i_d), i_d)),
                ContactSpec("body", (np.full(nx, ny - 1), np.arange(nx)))]
    Vfb = flatband_voltage(-Na, tox, "n+poly", 300.0, SILICON)
    gates = [GateSpec("gate", (np.zeros_like(i_g), i_g), tox, Vfb, "y")]
    if vg_list is None:
        vg_list = np.linspace(0.0, 1.0, 21)
    sweep = [{"drain": float(Vd), "gate": float(v)} for v in vg_list]
    return DeviceSpec([x, y], doping, ntot,
                      dict(doping_mobility=True, srh=True, auger=True, bgn=True,
                           surface_mobility=surface_mobility),
                      {}, contacts, gates, sweep, name)


def config2_mosfet():
    """200x120 MOSFET, Vd = 0.05 V, Vg 0 -> 1.0 V in 21 points."""
    return mosfet_2d(name="mosfet_200x120")


def trigate_3d(nx=64, ny=48, nz=48, Lg=20 * NM, Lsd=15 * NM, H=30 * NM, W=10 * NM,
               tox=1 * NM, Na=1e16, Npeak=1e20, sigma_lat=5 * NM, sigma_y=100 * NM,
               Vd=0.05, vg_list=None, name="trigate"):
    """3D tri-gate (pytcad/pytcad/finfet3d.py:61-152 layout)."""
    L = 2 * Lsd + Lg
    hx = L / (nx - 1)
    x = focused_mesh(L, [Lsd, Lsd + Lg], nx, h_min=0.5 * hx, h_max=1.5 * hx, ratio=1.15)
    hy = H / (ny - 1)
    y = focused_mesh(H, [0.0], ny, h_min=0.4 * hy, h_max=1.5 * hy, ratio=1.15)
    hz = W / (nz - 1)
    z = focused_mesh(W, [0.0, W], nz, h_min=0.4 * hz, h_max=1.5 * hz, ratio=1.15)
    d2, t2 = mosfet_doping(x, y, Lsd, Lg, Na, Npeak, sigma_y, sigma_lat)
    doping = np.tile(d2, (nz, 1, 1))
    ntot = np.tile(t2, (nz, 1, 1))
    i_s = np.where(x <= Lsd)[0]
    i_d = np.where(x >= Lsd + Lg)[0]
    i_g = np.where((x > Lsd) & (x < Lsd + Lg))[0]

    def grid(ii, jj, kk):
        I, J, K = np.meshgrid(ii, jj, kk, indexing="ij")
        return (K.ravel(), J.ravel(), I.ravel())

    contacts = [ContactSpec("source", grid(i_s, np.arange(ny), np.arange(nz))),
                ContactSpec("drain", grid(i_d, np.arange(ny), np.arange(nz))),
                ContactSpec("body", grid(np.arange(nx), [ny - 1], np.arange(nz)))]
    Vfb = flatband_voltage(-Na, tox, "n+poly", 300.0, SILICON)
    j_side = np.arange(1, ny)
    gates = [GateSpec("gate_top", grid(i_g, [0], np.arange(nz)), tox, Vfb, "y"),
             GateSpec("gate_left", grid(i_g, j_side, [0]), tox, Vfb, "z"),
             GateSpec("gate_right", grid(i_g, j_side, [nz - 1]), tox, Vfb, "z")]
    if vg_list is None:
        vg_list = np.linspace(0.0, 0.8, 17)
    sweep = [{"drain": float(Vd), "gate_top": float(v), "gate_left": float(v),
              "gate_right": float(v)} for v in vg_list]
    return DeviceSpec([x, y, z], doping, ntot,
                      dict(doping_mobility=True, srh=True, auger=True, bgn=True,
                           surface_mobility=False),
                      {}, contacts, gates, sweep, name)


def gated_slab_3d(axis, nx, ny, nz, L=60 * NM, depth=30 * NM, width=20 * NM, tox=2 * NM,
                  Na=1e17, Nsd=1e19, Vd=0.1, vg_list=(0.0, 0.5, 1.0), name=None):
    """3D MOS slab whose single gate face is normal to `axis` ('x' or
    'z'): the face at index 0 of that axis, over the middle half of the
    channel.  The channel runs along y for an x-normal gate and along x
    for a z-normal gate; abrupt n+ source/drain over the first and last
    quarter of the channel length on a p body.  Contacts: source and
    drain over the two end faces of the channel axis, body over the
    face opposite the gate.  The mesh is geometric away from the gate
    face (first cell depth/40) and uniform along the other two axes."""
    ax = {"x": 0, "y": 1, "z": 2}
    g = ax[axis]
    c = 1 if axis == "x" else 0                     # channel axis
    w = 3 - g - c                                   # the remaining axis
    n = (nx, ny, nz)
    lengths = [0.0, 0.0, 0.0]
    lengths[g], lengths[c], lengths[w] = depth, L, width
    coords = [None, None, None]
    coords[g] = geometric_mesh(depth, n[g], depth / 40)
    coords[c] = uniform_mesh(L, n[c])
    coords[w] = uniform_mesh(width, n[w])
    shape = (nz, ny, nx)
    cc = coords[c]
    sd = (cc < 0.25 * L) | (cc > 0.75 * L)
    prof = np.where(sd, Nsd, -Na)
    bshape = [1, 1, 1]
    bshape[2 - c] = n[c]
    doping = np.broadcast_to(prof.reshape(bshape), shape).copy()

    def face(axis_idx, value, sel_axis=None, sel=None):
        ranges = [np.arange(nx), np.arange(ny), np.arange(nz)]
        ranges[axis_idx] = np.array([value])
        if sel_axis is not None:
            ranges[sel_axis] = sel
        I, J, K = np.meshgrid(*ranges, indexing="ij")
        return (K.ravel(), J.ravel(), I.ravel())

    chan = np.where((cc > 0.25 * L) & (cc < 0.75 * L))[0]
    contacts = [ContactSpec("source", face(c, 0)),
                ContactSpec("drain", face(c, n[c] - 1)),
                ContactSpec("body", face(g, n[g] - 1))]
    Vfb = flatband_voltage(-Na, tox, "n+poly", 300.0, SILICON)
    gates = [GateSpec("gate", face(g, 0, c, chan), tox, Vfb, axis)]
    sweep = [{"drain": float(Vd), "gate": float(v)} for v in vg_list]
    return DeviceSpec(coords, doping, None,
                      dict(doping_mobility=True, srh=True, auger=True, bgn=True,
                           surface_mobility=False),
                      {}, contacts, gates, sweep,
                      name or f"slab_{axis}gate_{nx}x{ny}x{nz}")


def config3_trigate(Vd=0.05):
    """64x48x48 tri-gate, Vg 0 -> 0.8 V in 17 points at the given Vd."""
    return trigate_3d(Vd=Vd, name=f"trigate_64x48x48_vd{Vd}")


BATCH_SEED = 20261203


def batch_diodes(count=256, seed=BATCH_SEED, nx_range=(40, 160), ny_range=(30, 100),
                 max_step=0.1):
    """Independent 2D diodes with seeded random size, doping, junction
    position and target bias.  Each spec's sweep is the continuation
    path from 0 V to its target in steps of at most `max_step`; only the
    last point is the unit of work."""
    rng = np.random.default_rng(seed)
    specs = []
    for d in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        Nd = float(10 ** rng.uniform(15, 19))
        Na = float(10 ** rng.uniform(15, 19))
        xj = float(rng.uniform(0.3, 0.7))
        vt = float(rng.uniform(-2.0, 0.8))
        steps = max(1, int(np.ceil(abs(vt) / max_step - 1e-12)))
        path = [vt * (s + 1) / steps for s in range(steps)]
        sp = diode_2d(nx, ny, xj_frac=xj, Na=Na, Nd=Nd, biases=path,
                      models=dict(doping_mobility=True, srh=True, auger=True,
                                  bgn=True, surface_mobility=False),
                      material={}, name=f"diode{d}_{nx}x{ny}")
        specs.append(sp)
    return specs
