"""Wire formats of the solver boundary: the DeviceSpec job JSON that
crosses into a solver process (pytcad/gui/services/device_spec.py) and
the npz result grammar a solver writes back (the key grammar documented
in pytcad/gui/services/solver_backend.py:17-80, as
solver_runner.run_job / extract_result / run_sweep / run_transient emit
it, solver_runner.py:279-589, 709-812, 1140-1246).

The field names, JSON keys, validation conditions and npz keys are the
reference's, so a job file written for PyTCAD runs here unchanged and a
result written here opens in PyTCAD's result store.  The solves behind
them are this package's GPU devices: structured 2D/3D meshes, one
silicon region, Boltzmann statistics, ohmic contacts and Robin gates,
with the physics flags tcadhip.physics.Models implements; a job asking
for anything else (1D, heterostructures, the other model flags) is
refused with a ValueError naming what is missing, never approximated.

Entry points: DeviceSpec.from_json / to_json, build_device(spec),
solve_job(spec) -> result dict, write_result(result, path) (atomic),
run_job(job_path, out_path), validate_result(path_or_mapping),
ResultFile(path) (reader); for the serving loop, request_from_job(spec)
-> a tcadhip.serve.Request and served_result(spec, served) -> the same
result dict from what the server handed back.
"""
import json
import math
import os
import warnings
from dataclasses import asdict, dataclass, field
from datetime import datetime, timezone

import numpy as np

MAX_SWEEP_POINTS = 100_000
CURRENT_SPEC_VERSION = 1
RESULT_SCHEMA_VERSION = 3
KNOWN_RESULT_SCHEMA_VERSIONS = frozenset({1, 2, 3})
GEOM_STRUCTURED = "structured_rectilinear"
GEOM_POINT_CLOUD = "point_cloud"
WAVEFORM_KINDS = ("step", "ramp", "pulse", "constant")
ENGINE_CHOICES = ("auto", "direct", "gpu_direct", "amg", "mpi_schwarz")
BACKEND = "tcadhip"

# model flags of the reference's wire format (device_spec.py:_default_models)
# and the subset the GPU devices implement
DEFAULT_MODELS = {"doping_mobility": True, "field_mobility": False, "srh": True,
                  "auger": True, "bgn": True, "tat": False, "fd": False,
                  "incomplete_ion": False, "impact": False, "btbt": False,
                  "btbt_nonlocal": False, "impact_nonlocal": False,
                  "surface_mobility": False, "dg": False}
GPU_MODELS = ("doping_mobility", "srh", "auger", "bgn", "surface_mobility")


class UnsupportedSpecVersionError(Exception):
    """A job file stamped with a spec_version this build does not read."""


class ResultSchemaError(ValueError):
    """A result file violates the npz result grammar."""


def _finite(v):
    return isinstance(v, (int, float)) and not isinstance(v, bool) and math.isfinite(v)


def _ramp_points(start, stop, step):
    """Point count of an inclusive ramp; an endpoint within 1e-9 steps of
    the grid counts (device_spec.py:55-61)."""
    k = (stop - start) / step
    r = round(k)
    return int(r) + 1 if abs(k - r) < 1e-9 else int(math.floor(k)) + 1


def _parse(cls, d, what, build):
    if not isinstance(d, dict):
        raise ValueError(f"{what} configuration must be an object, got {type(d).__name__}")
    try:
        obj = build(d)
    except KeyError as exc:
        raise ValueError(f"{what} configuration is missing field {exc}") from None
    except (TypeError, ValueError) as exc:
        raise ValueError(f"invalid {what} configuration: {exc}") from None
    return obj


def _check_contact(what, name, names):
    names = list(names)
    if not isinstance(name, str) or not name or name not in names:
        raise ValueError(f"{what} contact '{name}' is not a registered contact "
                         f"(have: {', '.join(names) or 'none'})")


# ---------------------------------------------------------------------------
# the job (DeviceSpec) wire format
# ---------------------------------------------------------------------------
@dataclass
class SweepSpec:
    """One contact (or gate) ramped start -> stop by step, every other
    contact at its DeviceSpec bias (device_spec.py:36-133)."""
    contact: str
    start: float
    stop: float
    step: float

    def voltages(self):
        return [self.start + i * self.step
                for i in range(_ramp_points(self.start, self.stop, self.step))]

    def n_points(self):
        return len(self.voltages())

    def to_dict(self):
        return asdict(self)

    @classmethod
    def from_dict(cls, d):
        spec = _parse(cls, d, "sweep", lambda d: cls(
            contact=d["contact"], start=float(d["start"]), stop=float(d["stop"]),
            step=float(d["step"])))
        spec.validate_values()
        return spec

    def validate_values(self):
        for label in ("start", "stop", "step"):
            v = getattr(self, label)
            if not _finite(v):
                raise ValueError(f"sweep {label} must be finite, got {v!r}")
        if self.step == 0:
            raise ValueError("sweep step must be nonzero")
        if self.start == self.stop:
            raise ValueError(f"a sweep needs at least 2 points; start == stop ({self.start}) "
                             "gives 1")
        if (self.stop - self.start) * self.step < 0:
            raise ValueError(f"sweep step {self.step} does not move from start {self.start} "
                             f"toward stop {self.stop}")
        n = _ramp_points(self.start, self.stop, self.step)
        if n > MAX_SWEEP_POINTS:
            raise ValueError(f"sweep would need {n} points; the limit is {MAX_SWEEP_POINTS} "
                             "(increase the step)")

    def validate(self, contact_names):
        _check_contact("sweep", self.contact, contact_names)
        self.validate_values()


@dataclass
class WaveformSpec:
    """Bias versus time, one shape for the four kinds
    (device_spec.py:136-189): step (v0 before t0, v1 from t0), ramp (v0 ->
    v1 over [t0, t1]), pulse (v1 on [t0, t0 + t1), v0 elsewhere; t1 is the
    width), constant (v0)."""
    kind: str
    v0: float = 0.0
    v1: float = 0.0
    t0: float = 0.0
    t1: float = 0.0

    def to_dict(self):
        return asdict(self)

    @classmethod
    def from_dict(cls, d):
        spec = _parse(cls, d, "waveform", lambda d: cls(
            kind=d["kind"], v0=float(d.get("v0", 0.0)), v1=float(d.get("v1", 0.0)),
            t0=float(d.get("t0", 0.0)), t1=float(d.get("t1", 0.0))))
        spec.validate()
        return spec

    def validate(self):
        if self.kind not in WAVEFORM_KINDS:
            raise ValueError(f"waveform kind '{self.kind}' must be one of {WAVEFORM_KINDS}")
        for label in ("v0", "v1", "t0", "t1"):
            v = getattr(self, label)
            if not _finite(v):
                raise ValueError(f"waveform {label} must be finite, got {v!r}")
        if self.kind == "ramp" and self.t1 <= self.t0:
            raise ValueError(f"ramp waveform needs t1 ({self.t1}) > t0 ({self.t0})")
        if self.kind == "pulse" and self.t1 <= 0.0:
            raise ValueError(f"pulse waveform needs width t1 > 0, got {self.t1}")

    def waveform(self):
        """The tcadhip.transient waveform object (solver_runner.py:483-495)."""
        from . import transient as tr
        if self.kind == "step":
            return tr.StepWaveform(self.v0, self.v1, t_step=self.t0)
        if self.kind == "ramp":
            return tr.RampWaveform(self.v0, self.v1, self.t0, self.t1)
        if self.kind == "pulse":
            return tr.PulseWaveform(self.v0, self.v1, self.t0, self.t1)
        return tr.ConstantWaveform(self.v0)


@dataclass
class TransientSpec:
    """One contact driven by a waveform in time, every other contact at
    its DeviceSpec bias (device_spec.py:192-250)."""
    contact: str
    waveform: WaveformSpec
    t_end: float
    dt0: float
    theta: float = 1.0

    def to_dict(self):
        return asdict(self)

    @classmethod
    def from_dict(cls, d):
        spec = _parse(cls, d, "transient", lambda d: cls(
            contact=d["contact"], waveform=WaveformSpec.from_dict(d["waveform"]),
            t_end=float(d["t_end"]), dt0=float(d["dt0"]), theta=float(d.get("theta", 1.0))))
        spec.validate_values()
        return spec

    def validate_values(self):
        self.waveform.validate()
        for label in ("t_end", "dt0", "theta"):
            v = getattr(self, label)
            if not _finite(v):
                raise ValueError(f"transient {label} must be finite, got {v!r}")
        if self.t_end <= 0.0:
            raise ValueError(f"transient t_end must be > 0, got {self.t_end}")
        if self.dt0 <= 0.0:
            raise ValueError(f"transient dt0 must be > 0, got {self.dt0}")
        if not 0.0 <= self.theta <= 1.0:
            raise ValueError(f"transient theta must be in [0, 1], got {self.theta}")

    def validate(self, contact_names):
        _check_contact("transient", self.contact, contact_names)
        self.validate_values()


@dataclass
class ACSpec:
    """Small-signal sweep of one driven port at the converged bias point,
    n_points log-spaced frequencies (device_spec.py:253-308)."""
    contact: str
    f_start: float
    f_stop: float
    n_points: int = 40

    def to_dict(self):
        return asdict(self)

    def freqs(self):
        return np.logspace(np.log10(self.f_start), np.log10(self.f_stop), self.n_points)

    @classmethod
    def from_dict(cls, d):
        spec = _parse(cls, d, "AC", lambda d: cls(
            contact=d["contact"], f_start=float(d["f_start"]), f_stop=float(d["f_stop"]),
            n_points=int(d.get("n_points", 40))))
        spec.validate_values()
        return spec

    def validate_values(self):
        for label in ("f_start", "f_stop"):
            v = getattr(self, label)
            if not _finite(v):
                raise ValueError(f"AC {label} must be finite, got {v!r}")
        if self.f_start <= 0.0:
            raise ValueError(f"AC f_start must be > 0, got {self.f_start}")
        if self.f_stop <= self.f_start:
            raise ValueError(f"AC f_stop must be > f_start, got f_stop={self.f_stop}, "
                             f"f_start={self.f_start}")
        if not isinstance(self.n_points, int) or self.n_points < 2:
            raise ValueError(f"AC n_points must be an integer >= 2, got {self.n_points!r}")

    def validate(self, contact_names):
        _check_contact("AC", self.contact, contact_names)
        self.validate_values()


@dataclass
class MeshSpec:
    """Axis node positions [cm] only (device_spec.py:311-334)."""
    dimensionality: int
    axes: dict

    def shape(self):
        """(Nx,) / (Ny, Nx) / (Nz, Ny, Nx)."""
        if self.dimensionality not in (1, 2, 3):
            raise ValueError(f"dimensionality must be 1, 2 or 3, got {self.dimensionality}")
        return tuple(len(self.axes[a]) for a in reversed("xyz"[:self.dimensionality]))


@dataclass
class DopingSpec:
    """Net doping N_D - N_A [cm^-3] as nested lists of MeshSpec.shape();
    ntotal the total ionised concentration (None: |values|)."""
    kind: str
    values: list
    ntotal: list = None


@dataclass
class ContactSpec:
    """A terminal: kind "ohmic" or "gate"; nodes {"i": [...], "j": [...]
    [, "k": [...]]}; tox_cm / Vfb / normal_axis for gates
    (device_spec.py:349-364)."""
    name: str
    kind: str
    nodes: dict
    V: float = 0.0
    tox_cm: float = None
    Vfb: float = None
    normal_axis: str = "z"


def _check_boxes(entries, label, key):
    if not isinstance(entries, list):
        raise ValueError(f"{label} must be a list")
    for i, e in enumerate(entries):
        if not isinstance(e, dict) or not isinstance(e.get(key), str) or not e[key]:
            raise ValueError(f"{label}[{i}] needs a '{key}' string")
        box = e.get("box")
        if not isinstance(box, (list, tuple)) or len(box) not in (2, 4, 6) \
                or not all(isinstance(v, (int, float)) for v in box):
            raise ValueError(f"{label}[{i}].box must be [x0,x1], [x0,x1,y0,y1] or "
                             "[x0,x1,y0,y1,z0,z1] in cm")


def default_models():
    return dict(DEFAULT_MODELS)


@dataclass
class DeviceSpec:
    """The job: mesh, doping, material, temperature, model flags,
    contacts, an optional bias and at most one of sweep / transient, and
    an optional AC sweep at the bias point (device_spec.py:407-557)."""
    mesh: MeshSpec
    doping: DopingSpec
    material: str = "SILICON"
    T: float = 300.0
    models: dict = field(default_factory=default_models)
    contacts: list = field(default_factory=list)
    bias: dict = None
    sweep: SweepSpec = None
    transient: TransientSpec = None
    region_materials: list = None
    backend: str = "pytcad"
    engine: str = "auto"
    structure_regions: list = None
    ac: ACSpec = None
    spec_version: int = CURRENT_SPEC_VERSION

    def to_dict(self):
        d = asdict(self)
        if self.spec_version == 1:
            del d["spec_version"]          # version 1 files carry no stamp
        return d

    @classmethod
    def from_dict(cls, d):
        version = d.get("spec_version", 1)
        if not isinstance(version, int) or isinstance(version, bool) or version < 1 \
                or version > CURRENT_SPEC_VERSION:
            raise UnsupportedSpecVersionError(
                f"device spec version {version!r} is not supported (this build supports "
                f"up to {CURRENT_SPEC_VERSION})")
        for key, label, name in (("region_materials", "region_materials", "material"),
                                 ("structure_regions", "structure_regions", "name")):
            if d.get(key) is not None:
                _check_boxes(d[key], label, name)
        sweep, transient, ac = d.get("sweep"), d.get("transient"), d.get("ac")
        return cls(
            mesh=MeshSpec(**d["mesh"]), doping=DopingSpec(**d["doping"]),
            material=d.get("material", "SILICON"), region_materials=d.get("region_materials"),
            T=d.get("T", 300.0), models=d.get("models") or default_models(),
            contacts=[ContactSpec(**c) for c in d.get("contacts", [])], bias=d.get("bias"),
            sweep=SweepSpec.from_dict(sweep) if sweep else None,
            transient=TransientSpec.from_dict(transient) if transient else None,
            ac=ACSpec.from_dict(ac) if ac else None, backend=d.get("backend", "pytcad"),
            engine=d.get("engine", "auto"), structure_regions=d.get("structure_regions"),
            spec_version=version)

    def to_json(self, path):
        with open(path, "w") as fh:
            json.dump(self.to_dict(), fh)

    @classmethod
    def from_json(cls, path):
        with open(path) as fh:
            return cls.from_dict(json.load(fh))

    def contact_names(self):
        return [c.name for c in self.contacts]


# ---------------------------------------------------------------------------
# job -> device
# ---------------------------------------------------------------------------
def check_supported(spec):
    """Refuse, with the reason, a job the GPU devices cannot run."""
    dim = spec.mesh.dimensionality
    if dim not in (2, 3):
        raise ValueError(f"tcadhip solves structured 2D and 3D devices; this job is {dim}D")
    if str(spec.material).upper() not in ("SILICON", "SI"):
        raise ValueError(f"tcadhip solves silicon devices only; this job's material is "
                         f"'{spec.material}'")
    if spec.region_materials:
        raise ValueError("tcadhip solves single-material devices; this job has "
                         "region_materials (a heterostructure)")
    models = dict(default_models())
    models.update(spec.models or {})
    unknown = sorted(k for k in models if k not in DEFAULT_MODELS)
    if unknown:
        raise ValueError(f"unknown model flags {unknown}")
    extra = sorted(k for k, v in models.items() if v and k not in GPU_MODELS)
    if extra:
        raise ValueError(f"model flags {extra} are not implemented by tcadhip (it implements "
                         f"{list(GPU_MODELS)})")
    if dim == 3 and models["surface_mobility"]:
        raise ValueError("the surface-mobility model is implemented for 2D devices only")
    if spec.engine not in ENGINE_CHOICES:
        raise ValueError(f"unknown engine {spec.engine!r}; expected one of {ENGINE_CHOICES}")
    if spec.doping.kind != "array":
        raise ValueError(f"unsupported doping kind '{spec.doping.kind}' ('array' only)")
    for c in spec.contacts:
        if c.kind not in ("ohmic", "gate"):
            raise ValueError(f"unknown contact kind '{c.kind}'")
        if c.kind == "gate" and (c.tox_cm is None or c.Vfb is None):
            raise ValueError(f"gate '{c.name}' needs tox_cm and Vfb")
    names = spec.contact_names()
    if spec.sweep is not None:
        spec.sweep.validate(names)
    if spec.transient is not None:
        spec.transient.validate(names)
    if spec.ac is not None:
        spec.ac.validate(names)
    return {k: bool(models[k]) for k in GPU_MODELS}


def build_device(spec):
    """A tcadhip Device2D/Device3D with the job's contacts and gates
    registered at their ContactSpec voltages (solver_runner.py:103-217)."""
    from . import physics as ph
    from .device import Device2D, Device3D
    from .mesh import Mesh2D, Mesh3D
    flags = check_supported(spec)
    shape = spec.mesh.shape()
    doping = np.asarray(spec.doping.values, dtype=float).reshape(shape)
    ntot = None if spec.doping.ntotal is None else \
        np.asarray(spec.doping.ntotal, dtype=float).reshape(shape)
    axes = [np.asarray(spec.mesh.axes[a], dtype=float) for a in "xyz"[:spec.mesh.dimensionality]]
    models = ph.Models(**flags)
    if spec.mesh.dimensionality == 2:
        dev = Device2D(Mesh2D(*axes), doping, ntot, T=float(spec.T), models=models)
    else:
        dev = Device3D(Mesh3D(*axes), doping, ntot, T=float(spec.T), models=models)
    for c in spec.contacts:
        idx = [np.asarray(c.nodes[k], dtype=int) for k in "ijk"[:spec.mesh.dimensionality]]
        if c.kind == "ohmic":
            dev.add_contact(c.name, *idx, V=float(c.V))
        elif spec.mesh.dimensionality == 3:
            dev.add_gate(c.name, *idx, tox_cm=float(c.tox_cm), Vfb=float(c.Vfb),
                         Vg=float(c.V), normal_axis=c.normal_axis)
        else:
            dev.add_gate(c.name, *idx, tox_cm=float(c.tox_cm), Vfb=float(c.Vfb), Vg=float(c.V))
    return dev


def merge_bias(spec, override=None):
    """ContactSpec.V, overridden by DeviceSpec.bias, overridden by a
    per-point value (solver_runner.py:220-231)."""
    bias = {c.name: float(c.V) for c in spec.contacts}
    bias.update({k: float(v) for k, v in (spec.bias or {}).items()})
    bias.update({k: float(v) for k, v in (override or {}).items()})
    return bias


# ---------------------------------------------------------------------------
# device -> result arrays
# ---------------------------------------------------------------------------
def _edges_to_nodes(a, axis):
    """Edge values averaged onto nodes along `axis`: the mean of the one
    or two edges touching each node (solver_runner.py:251-276)."""
    a = np.asarray(a, dtype=float)
    n = a.shape[axis] + 1
    pad_lo = [(0, 0)] * a.ndim
    pad_hi = [(0, 0)] * a.ndim
    pad_lo[axis] = (1, 0)
    pad_hi[axis] = (0, 1)
    total = np.pad(a, pad_hi) + np.pad(a, pad_lo)
    count = np.full(n, 2.0)
    count[0] = count[-1] = 1.0
    shp = [1] * a.ndim
    shp[axis] = n
    return total / count.reshape(shp)


def _ohmic(spec):
    return [c.name for c in spec.contacts if c.kind == "ohmic"]


def extract_fields(dev, spec, solved_bias):
    """The scalar fields, node-averaged current density and terminal
    currents of the device's current state (solver_runner.py:279-357)."""
    dim = spec.mesh.dimensionality
    out = {"dimensionality": np.array(dim), "solved_bias": np.array(bool(solved_bias))}
    for a in "xyz"[:dim]:
        out[f"axis_{a}"] = np.asarray(spec.mesh.axes[a], dtype=float)
    for name, arr, unit in (("potential", dev.psi_V, "V"),
                            ("electron_density", dev.n_cm3, "cm^-3"),
                            ("hole_density", dev.p_cm3, "cm^-3"),
                            ("doping", dev.doping, "cm^-3")):
        out[f"field__{name}"] = np.asarray(arr, dtype=float).reshape(spec.mesh.shape())
        out[f"unit__{name}"] = np.array(unit)
    if solved_bias:
        edges = dev.current_density_edges()
        for a, name in enumerate("xyz"[:dim]):
            Jn, Jp = edges[name]
            out[f"vector__current_density__{name}"] = _edges_to_nodes(Jn + Jp, dim - 1 - a)
        out["unit__current_density"] = np.array("A/cm^2")
        unit = "A/cm" if dim == 2 else "A"
        currents = dev.terminal_currents()
        for nm in _ohmic(spec):
            out[f"terminal__{nm}__value"] = np.array(float(currents[nm]))
            out[f"terminal__{nm}__unit"] = np.array(unit)
    out["band__available"] = np.array(False)
    return out


def _converged_solve(dev, voltages, opts):
    """solve_bias; converged unless Newton warned or a linear solve
    failed (solver_runner.py:413-431)."""
    from ._lib import LinearSolveError
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        try:
            dev.solve_bias(voltages, opts)
        except LinearSolveError:
            return False
    return not any("did not converge" in str(w.message) for w in caught)


def run_sweep(dev, spec, opts, fallback):
    """Warm-started ramp of spec.sweep.contact; fields of the last
    converged point (else `fallback`) and the sweep block
    (solver_runner.py:360-480)."""
    sw = spec.sweep
    dim = spec.mesh.dimensionality
    channels = _ohmic(spec)
    currents = {nm: [] for nm in channels}
    flags, fields = [], None
    snap_v, snaps = [], {"potential": [], "electron_density": [], "hole_density": []}
    voltages = sw.voltages()
    for V in voltages:
        ok = _converged_solve(dev, merge_bias(spec, {sw.contact: V}), opts)
        flags.append(ok)
        I = dev.terminal_currents()
        for nm in channels:
            currents[nm].append(float(I[nm]))
        if ok:
            fields = extract_fields(dev, spec, True)
            if dim == 3:
                snap_v.append(float(V))
                snaps["potential"].append(np.asarray(dev.psi_V, dtype=float))
                snaps["electron_density"].append(np.asarray(dev.n_cm3, dtype=float))
                snaps["hole_density"].append(np.asarray(dev.p_cm3, dtype=float))
    series = {
        "sweep__voltage": np.asarray(voltages, dtype=float),
        "sweep__converged": np.asarray(flags, dtype=bool),
        "unit__sweep_current": np.array({2: "A/cm", 3: "A"}[dim]),
        "sweep__meta": np.array(json.dumps({"contact": sw.contact, "start": sw.start,
                                            "stop": sw.stop, "step": sw.step,
                                            "dimensionality": dim})),
    }
    for nm, vals in currents.items():
        series[f"sweep__current__{nm}"] = np.asarray(vals, dtype=float)
    if dim == 3 and snap_v:
        series["sweep__snapshot__voltages"] = np.array(json.dumps(snap_v))
        for name, arrs in snaps.items():
            for i, arr in enumerate(arrs):
                series[f"sweep__snapshot__field__{name}__{i}"] = arr
    return (fields if fields is not None else fallback), series


def run_transient(dev, spec, opts):
    """DC solve with the stimulus at its waveform's v0, then the
    transient (solver_runner.py:498-587)."""
    from .transient import solve_transient
    tr = spec.transient
    dim = spec.mesh.dimensionality
    dev.solve_bias(merge_bias(spec, {tr.contact: tr.waveform.v0}), opts)
    res = solve_transient(dev, {tr.contact: tr.waveform.waveform()}, tr.t_end, tr.dt0,
                          theta=tr.theta, opts=opts, keep_history=False)
    fields = extract_fields(dev, spec, True)
    series = {
        "transient__times": np.asarray(res.times, dtype=float),
        "unit__transient_current": np.array({2: "A/cm", 3: "A"}[dim]),
        "transient__meta": np.array(json.dumps({
            "contact": tr.contact, "waveform": tr.waveform.to_dict(), "t_end": tr.t_end,
            "dt0": tr.dt0, "theta": tr.theta, "dimensionality": dim})),
    }
    for nm, vals in res.terminal_current.items():
        series[f"transient__current__{nm}"] = np.asarray(vals, dtype=float)
    return fields, series


def run_ac(dev, spec):
    """C(f) and G(f) of the driven port's own admittance at the current
    bias point (solver_runner.py:761-802)."""
    from .ac import y_parameters
    freqs = spec.ac.freqs()
    yres = y_parameters(dev, freqs)
    k = yres.port_names.index(spec.ac.contact)
    Ykk = yres.Y[:, k, k]
    three = spec.mesh.dimensionality == 3
    return {"ac__freqs": freqs, "ac__C": Ykk.imag / (2 * np.pi * freqs), "ac__G": Ykk.real,
            "ac__port": np.array(spec.ac.contact),
            "unit__ac_capacitance": np.array("F" if three else "F/cm^2"),
            "unit__ac_conductance": np.array("S" if three else "S/cm^2")}


def node_coords(mesh):
    """Flat node coordinates [cm], (N, dim), x fastest."""
    axes = [np.asarray(mesh.axes[a], dtype=float) for a in "xyz"[:mesh.dimensionality]]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel(order="F") for g in grids], axis=1)


def solve_job(spec, opts=None):
    """Equilibrium, then the job's transient, sweep or bias (+ AC), as
    solver_runner._solve_all orders them (solver_runner.py:709-812);
    returns the result arrays with the grammar's stamps and run record
    (solver_runner.py:1140-1226)."""
    from .device import NewtonOptions
    opts = opts or NewtonOptions()
    dev = build_device(spec)
    dev.solve_equilibrium(opts)
    if spec.transient is not None:
        result, series = run_transient(dev, spec, opts)
        result.update(series)
    elif spec.sweep is not None:
        eq_fields = extract_fields(dev, spec, False)
        result, series = run_sweep(dev, spec, opts, eq_fields)
        result.update(series)
    else:
        solved = spec.bias is not None
        if solved:
            dev.solve_bias(merge_bias(spec), opts)
        result = extract_fields(dev, spec, solved)
        if spec.ac is not None:
            result.update(run_ac(dev, spec))
    return _stamp(result, spec, opts)


def _stamp(result, spec, opts):
    """The grammar's stamps, geometry keys and run record
    (solver_runner.py:1140-1226)."""
    result["result__schema"] = np.array(RESULT_SCHEMA_VERSION)
    result["geom__kind"] = np.array(GEOM_STRUCTURED)
    result["mesh__shape"] = np.array(spec.mesh.shape())
    coords = node_coords(spec.mesh)
    result["nodes__count"] = np.array(int(coords.shape[0]))
    result["nodes__coords"] = coords
    for key, meta in (("region_materials__meta", spec.region_materials),
                      ("structure_regions__meta", spec.structure_regions)):
        if meta:
            result[key] = np.array(json.dumps(meta))
    sweep_meta = transient_meta = None
    if spec.transient is not None:
        transient_meta = json.loads(str(result["transient__meta"]))
    if spec.sweep is not None:
        sweep_meta = json.loads(str(result["sweep__meta"]))
        result["continuation__records"] = np.array(json.dumps([
            {"index": i, "parameter": float(v), "accepted": bool(c)}
            for i, (v, c) in enumerate(zip(result["sweep__voltage"],
                                           result["sweep__converged"]))]))
    numerics = asdict(opts)
    numerics["linsolve_requested"] = opts.linsolve
    numerics["engine"] = "tcadhip (" + ("FP32-stored matrix" if opts.linsolve in (
        "bicgstab_mixed", "mixed") else "FP64") + " BiCGStab on the GPU)"
    result["record__meta"] = np.array(json.dumps({
        "schema_version": RESULT_SCHEMA_VERSION, "backend": BACKEND,
        "created_utc": datetime.now(timezone.utc).isoformat(timespec="seconds"),
        "dimensionality": spec.mesh.dimensionality, "material": spec.material,
        "T": spec.T, "models": spec.models, "numerics": numerics,
        "sweep": sweep_meta, "transient": transient_meta}))
    return result


def config_spec(spec, points, name="job"):
    """The job as a tcadhip.configs.DeviceSpec with the given full bias
    points, the input of the batched sweep and serving drivers."""
    from . import configs
    flags = check_supported(spec)
    dim = spec.mesh.dimensionality
    shape = spec.mesh.shape()
    coords = [np.asarray(spec.mesh.axes[a], dtype=float) for a in "xyz"[:dim]]
    doping = np.asarray(spec.doping.values, dtype=float).reshape(shape)
    ntot = None if spec.doping.ntotal is None else \
        np.asarray(spec.doping.ntotal, dtype=float).reshape(shape)

    def idx(c):                                   # array order (j, i) / (k, j, i)
        return tuple(np.asarray(c.nodes[k], dtype=int) for k in "kji"[3 - dim:])

    contacts = [configs.ContactSpec(c.name, idx(c), float(c.V))
                for c in spec.contacts if c.kind == "ohmic"]
    gates = [configs.GateSpec(c.name, idx(c), float(c.tox_cm), float(c.Vfb),
                              "y" if dim == 2 else c.normal_axis, float(c.V))
             for c in spec.contacts if c.kind == "gate"]
    return configs.DeviceSpec(coords, doping, ntot, flags, {}, contacts, gates,
                              sweep=[dict(p) for p in points], name=name, T=float(spec.T))


def request_from_job(spec, rid, arrival=0.0):
    """A serving request for a sweep job (every point a full bias) or a
    single-bias job; the server hands back the final state with it."""
    from .serve import Request
    if spec.transient is not None or spec.ac is not None:
        raise ValueError("the serving loop runs sweep and bias jobs; transient and AC jobs "
                         "run through solve_job")
    if spec.sweep is not None:
        points = [merge_bias(spec, {spec.sweep.contact: V}) for V in spec.sweep.voltages()]
    elif spec.bias is not None:
        points = [merge_bias(spec)]
    else:
        raise ValueError("an equilibrium-only job has no bias point to serve")
    return Request(rid, config_spec(spec, points, name=f"job{rid}"), float(arrival), "job",
                   keep_state=True)


def served_result(spec, served, opts=None):
    """The result dict of a job the server solved (request_from_job):
    fields at the last converged point (the lagged surface mobility
    re-evaluated at that state), the terminal currents, and for a sweep
    job the sweep block -- points past a failure are reported not
    converged, with NaN currents."""
    from .device import NewtonOptions
    if served.state is None:
        raise ValueError("served_result needs a request made by request_from_job")
    res = served.result
    dev = build_device(spec)
    dev.set_state(*served.state)
    reached = len(res.currents)
    if reached:
        dev._set_voltages(res.points[-1])
        if dev.models.surface_mobility:
            dev._update_surface_mobility(served.state[0])
    fields = extract_fields(dev, spec, reached > 0)
    if spec.sweep is not None:
        sw = spec.sweep
        voltages = sw.voltages()
        conv = np.arange(len(voltages)) < reached
        fields.update({
            "sweep__voltage": np.asarray(voltages, dtype=float),
            "sweep__converged": conv,
            "unit__sweep_current": np.array({2: "A/cm", 3: "A"}[spec.mesh.dimensionality]),
            "sweep__meta": np.array(json.dumps({"contact": sw.contact, "start": sw.start,
                                                "stop": sw.stop, "step": sw.step,
                                                "dimensionality": spec.mesh.dimensionality}))})
        for nm in _ohmic(spec):
            vals = np.full(len(voltages), np.nan)
            vals[:reached] = [float(c[nm]) for c in res.currents]
            fields[f"sweep__current__{nm}"] = vals
    return _stamp(fields, spec, opts or NewtonOptions())


def write_result(result, out_path):
    """Atomic write: <out>.tmp.npz, then rename over out_path, so a killed
    writer never leaves a partial file at the canonical path."""
    tmp = out_path + ".tmp.npz"
    np.savez(tmp, **result)
    os.replace(tmp, out_path)
    return out_path


def run_job(job_path, out_path, opts=None):
    """The solver-process contract: DeviceSpec JSON in, validated npz out."""
    spec = DeviceSpec.from_json(job_path)
    check_supported(spec)
    result = solve_job(spec, opts)
    validate_result(result)
    return write_result(result, out_path)


# ---------------------------------------------------------------------------
# the result grammar: validation and reading
# ---------------------------------------------------------------------------
class _Mapping:
    """A dict of arrays with the .files / [key] interface of an NpzFile."""

    def __init__(self, d, name):
        self._d, self.filename = d, name

    @property
    def files(self):
        return list(self._d.keys())

    def __getitem__(self, k):
        return self._d[k]

    def __contains__(self, k):
        return k in self._d


def _scalar_int(d, key, path):
    if key not in d.files:
        raise ResultSchemaError(f"{path}: missing required key '{key}'")
    try:
        return int(np.asarray(d[key]).reshape(()))
    except Exception:
        raise ResultSchemaError(f"{path}: '{key}' must be a scalar integer") from None


def _need(d, key, why, path):
    if key not in d.files:
        raise ResultSchemaError(f"{path}: missing required key '{key}' ({why})")


def _json_of(d, key, path):
    try:
        return json.loads(str(np.asarray(d[key]).reshape(())))
    except Exception as exc:
        raise ResultSchemaError(f"{path}: {key} is not valid JSON ({exc})") from exc


def _series_block(d, files, prefix, required, axis_key, channel_prefix, path, what):
    keys = [k for k in files if k.startswith(prefix)]
    if not keys:
        return
    for r in required:
        _need(d, r, f"an incomplete {what} block is invalid", path)
    axis = np.asarray(d[axis_key], dtype=float)
    if channel_prefix is None:
        return axis
    chans = [k for k in keys if k.startswith(channel_prefix)]
    if not chans:
        raise ResultSchemaError(f"{path}: {what} block has no {channel_prefix}<channel> series")
    for ch in chans:
        v = np.asarray(d[ch], dtype=float)
        if v.shape != axis.shape:
            raise ResultSchemaError(f"{path}: {ch} length {v.size} does not match {axis_key} "
                                    f"length {axis.size}")
    return axis


def validate_result(npz):
    """Structural validation of a result (path, NpzFile or dict) against
    the grammar (solver_backend.py:197-461): the schema stamp is known;
    the axes match the dimensionality; every field and vector has a unit
    and the mesh shape; geometry keys agree; record/trace are JSON;
    terminals come in value/unit pairs; sweep, transient and AC blocks are
    complete with matching lengths.  Returns the schema version (1 for an
    unstamped file); raises ResultSchemaError."""
    close = None
    if isinstance(npz, str):
        if not npz.endswith(".npz") or not os.path.isfile(npz):
            raise ResultSchemaError(f"{npz}: result file not found or not an .npz path")
        try:
            d = close = np.load(npz)
        except Exception as exc:
            raise ResultSchemaError(f"{npz}: not a readable npz archive ({exc})") from exc
        path = npz
    elif isinstance(npz, dict):
        d, path = _Mapping(npz, "<result>"), "<result>"
    else:
        d, path = npz, getattr(npz, "filename", None) or "<npz>"
    try:
        return _validate(d, path)
    finally:
        if close is not None:
            close.close()


def _validate(d, path):
    files = set(getattr(d, "files", None) or ())
    if not files:
        raise ResultSchemaError(f"{path}: not an npz result archive")
    version = 1
    if "result__schema" in files:
        version = _scalar_int(d, "result__schema", path)
        if version not in KNOWN_RESULT_SCHEMA_VERSIONS:
            raise ResultSchemaError(f"{path}: result schema version {version} unsupported "
                                    f"(this build reads {sorted(KNOWN_RESULT_SCHEMA_VERSIONS)})")
    _need(d, "solved_bias", "always required", path)
    dim = _scalar_int(d, "dimensionality", path)
    if dim not in (1, 2, 3):
        raise ResultSchemaError(f"{path}: dimensionality must be 1, 2 or 3, got {dim}")
    axes = {}
    for a in "xyz"[:dim]:
        key = f"axis_{a}"
        _need(d, key, f"a {dim}D result needs it", path)
        arr = np.asarray(d[key])
        if arr.ndim != 1 or arr.size == 0:
            raise ResultSchemaError(f"{path}: '{key}' must be a non-empty 1D array of node "
                                    "positions")
        axes[a] = arr
    shape = tuple(axes[a].size for a in reversed("xyz"[:dim]))
    for key in sorted(files):
        if key.startswith("field__"):
            name = key[len("field__"):]
            _need(d, f"unit__{name}", f"field__{name} has no declared unit", path)
            if np.asarray(d[key]).shape != shape:
                raise ResultSchemaError(f"{path}: field__{name} shape "
                                        f"{np.asarray(d[key]).shape} does not match the {dim}D "
                                        f"mesh axes {shape}")
    groups = {}
    for key in sorted(files):
        if key.startswith("vector__"):
            rest = key[len("vector__"):]
            if "__" not in rest:
                raise ResultSchemaError(f"{path}: malformed vector key '{key}'")
            name, comp = rest.rsplit("__", 1)
            groups.setdefault(name, []).append(comp)
    for name, comps in groups.items():
        _need(d, f"unit__{name}", f"vector__{name}__* has no unit", path)
        if sorted(comps) != sorted(axes):
            raise ResultSchemaError(f"{path}: vector__{name}__ components {sorted(comps)} do "
                                    f"not match the {dim}D axes {list(axes)}")
        for comp in comps:
            if np.asarray(d[f"vector__{name}__{comp}"]).shape != shape:
                raise ResultSchemaError(f"{path}: vector__{name}__{comp} shape does not match "
                                        "the mesh axes")
    if "geom__kind" in files:
        kind = str(np.asarray(d["geom__kind"]).reshape(()))
        if kind not in (GEOM_STRUCTURED, GEOM_POINT_CLOUD):
            raise ResultSchemaError(f"{path}: unknown geom__kind '{kind}'")
        if kind == GEOM_POINT_CLOUD:
            raise ResultSchemaError(f"{path}: geom__kind 'point_cloud' is reserved and not "
                                    "readable; structured results only")
    if "geom__kind" in files or files & {"mesh__shape", "nodes__count", "nodes__coords"}:
        sizes = [a.size for a in axes.values()]
        count = None
        if "mesh__shape" in files:
            ms = [int(x) for x in np.asarray(d["mesh__shape"]).ravel()]
            if sorted(ms) != sorted(sizes):
                raise ResultSchemaError(f"{path}: mesh__shape {ms} does not match the axes "
                                        f"{sizes}")
            count = int(np.prod(ms))
        if "nodes__count" in files:
            c = _scalar_int(d, "nodes__count", path)
            if c != (count or int(np.prod(sizes))):
                raise ResultSchemaError(f"{path}: nodes__count {c} disagrees with the {dim}D "
                                        "mesh")
            count = c
        if "nodes__coords" in files:
            xy = np.asarray(d["nodes__coords"])
            if xy.ndim != 2 or xy.shape[1] != dim:
                raise ResultSchemaError(f"{path}: nodes__coords shape {xy.shape} must be "
                                        f"(N, {dim})")
            if count is not None and xy.shape[0] != count:
                raise ResultSchemaError(f"{path}: nodes__coords has {xy.shape[0]} rows but the "
                                        f"mesh declares {count} nodes")
    for key, kind in (("record__meta", dict), ("converge__trace", list)):
        if key in files and not isinstance(_json_of(d, key, path), kind):
            raise ResultSchemaError(f"{path}: {key} must be a JSON "
                                    f"{'object' if kind is dict else 'list'}")
    for key in sorted(files):
        if key.startswith("terminal__") and key.endswith("__value"):
            nm = key[len("terminal__"):-len("__value")]
            _need(d, f"terminal__{nm}__unit", f"terminal__{nm}__value has no unit", path)
        elif key.startswith("terminal__") and key.endswith("__unit"):
            nm = key[len("terminal__"):-len("__unit")]
            _need(d, f"terminal__{nm}__value", f"orphan terminal__{nm}__unit has no value",
                  path)
    v = _series_block(d, files, "sweep__", ("sweep__voltage", "sweep__converged",
                                            "unit__sweep_current", "sweep__meta"),
                      "sweep__voltage", "sweep__current__", path, "sweep")
    if v is not None:
        conv = np.asarray(d["sweep__converged"])
        if conv.shape != v.shape or conv.ndim != 1:
            raise ResultSchemaError(f"{path}: sweep__converged {conv.shape} must be a 1D bool "
                                    f"array matching sweep__voltage {v.shape}")
        meta = _json_of(d, "sweep__meta", path)
        if not isinstance(meta, dict) or not isinstance(meta.get("dimensionality"), int):
            raise ResultSchemaError(f"{path}: sweep__meta must be a JSON object with an "
                                    "integer 'dimensionality'")
    t = _series_block(d, files, "transient__", ("transient__times", "unit__transient_current",
                                                "transient__meta"),
                      "transient__times", "transient__current__", path, "transient")
    if t is not None:
        if t.ndim != 1 or t.size == 0:
            raise ResultSchemaError(f"{path}: transient__times must be a non-empty 1D array")
        meta = _json_of(d, "transient__meta", path)
        if not isinstance(meta, dict) or not isinstance(meta.get("dimensionality"), int):
            raise ResultSchemaError(f"{path}: transient__meta must be a JSON object with an "
                                    "integer 'dimensionality'")
    f = _series_block(d, files, "ac__", ("ac__freqs", "ac__C", "ac__G", "ac__port",
                                         "unit__ac_capacitance", "unit__ac_conductance"),
                      "ac__freqs", None, path, "ac")
    if f is not None:
        if f.ndim != 1 or f.size == 0:
            raise ResultSchemaError(f"{path}: ac__freqs must be a non-empty 1D array")
        for key in ("ac__C", "ac__G"):
            if np.asarray(d[key], dtype=float).shape != f.shape:
                raise ResultSchemaError(f"{path}: {key} length does not match ac__freqs")
    return version


class ResultFile:
    """A validated result read fully into memory (no file handle kept):
    fields, terminals and the sweep / transient / AC blocks, with the
    conventions of the reference's store (result_store.py:350-560) --
    non-converged sweep points read as NaN, the converged flags kept."""

    def __init__(self, path):
        with np.load(path) as z:
            self._d = {k: z[k] for k in z.files}
        self.path = path
        validate_result(self._d)

    def keys(self):
        return sorted(self._d)

    def __getitem__(self, k):
        return self._d[k]

    @property
    def dimensionality(self):
        return int(self._d["dimensionality"])

    def axes(self):
        return {a: self._d[f"axis_{a}"] for a in "xyz"[:self.dimensionality]}

    def scalar_names(self):
        return sorted(k[len("field__"):] for k in self._d if k.startswith("field__"))

    def scalar_field(self, name):
        """(values, unit)."""
        if f"field__{name}" not in self._d:
            raise KeyError(f"no scalar field '{name}' in {self.path}")
        return self._d[f"field__{name}"], str(self._d[f"unit__{name}"])

    def terminal_names(self):
        return sorted(k[len("terminal__"):-len("__value")] for k in self._d
                      if k.startswith("terminal__") and k.endswith("__value"))

    def terminal_current(self, name):
        """(value, unit)."""
        if f"terminal__{name}__value" not in self._d:
            raise KeyError(f"no terminal current '{name}' in {self.path}")
        return float(self._d[f"terminal__{name}__value"]), \
            str(self._d[f"terminal__{name}__unit"])

    def has_sweep(self):
        return "sweep__voltage" in self._d

    def sweep(self):
        """(meta, voltages, converged, {channel: currents with NaN where not
        converged}, unit)."""
        if not self.has_sweep():
            raise KeyError(f"no sweep series in {self.path}")
        conv = np.asarray(self._d["sweep__converged"], dtype=bool)
        chans = {}
        for k, v in self._d.items():
            if k.startswith("sweep__current__"):
                vals = np.asarray(v, dtype=float).copy()
                vals[~conv] = np.nan
                chans[k[len("sweep__current__"):]] = vals
        return (json.loads(str(self._d["sweep__meta"])),
                np.asarray(self._d["sweep__voltage"], dtype=float), conv, chans,
                str(self._d["unit__sweep_current"]))

    def has_transient(self):
        return "transient__times" in self._d

    def transient(self):
        """(meta, times, {channel: currents}, unit)."""
        if not self.has_transient():
            raise KeyError(f"no transient series in {self.path}")
        chans = {k[len("transient__current__"):]: np.asarray(v, dtype=float)
                 for k, v in self._d.items() if k.startswith("transient__current__")}
        return (json.loads(str(self._d["transient__meta"])),
                np.asarray(self._d["transient__times"], dtype=float), chans,
                str(self._d["unit__transient_current"]))

    def has_ac(self):
        return "ac__freqs" in self._d

    def ac(self):
        """(port, freqs, C, G, unit_C, unit_G)."""
        if not self.has_ac():
            raise KeyError(f"no AC sweep in {self.path}")
        return (str(self._d["ac__port"]), np.asarray(self._d["ac__freqs"], dtype=float),
                np.asarray(self._d["ac__C"], dtype=float),
                np.asarray(self._d["ac__G"], dtype=float),
                str(self._d["unit__ac_capacitance"]), str(self._d["unit__ac_conductance"]))

    def run_record(self):
        if "record__meta" not in self._d:
            return None
        return json.loads(str(self._d["record__meta"]))
