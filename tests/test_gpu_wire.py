"""DeviceSpec jobs solved on the GPU (tcadhip.wire.solve_job) against the
result files the reference's solver process wrote for the same job files
(tests/golden/make_golden_wire.py): the same keys, units, shapes and
metadata, and the same values within the solver bars used elsewhere:

* doping, axes, node coordinates: identical;
* potential within 1e-6 V; carrier densities within 1e-6 relative to
  max(value, 1e-10 Ns) (the update test floors densities at 1e-10 Ns);
* node-averaged current density within 1e-6 of its largest magnitude over
  all components (a laterally uniform device's y-component is rounding);
* terminal and sweep currents within 1e-6 of the largest current at that
  point plus both sides' own current-conservation errors |sum I| there
  (tests/current_bar.py);
* transient: the same time grid, currents per step as the transient
  tests (1e-6 of the step's largest plus 1e-9 of the run's);
* AC: the driven port's admittance G + j 2 pi f C within 1e-6 of |Y|.
A 3D job (the reference's job runner was not run for 3D; its Device3D
physics is pinned by tests/golden/make_golden_3d.py) must pass the
grammar and agree with the same device solved directly.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
JOBS = ["wire_resistor_2d_sweep", "wire_diode_2d_ac", "wire_diode_2d_transient",
        "wire_mosfet_2d_gate_sweep"]


def _wire():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import wire
    return wire


def _close_currents(got, ref, what):
    """tests/current_bar.py's bar: 1e-6 of the largest current plus the two
    sides' own current-conservation errors."""
    from current_bar import check
    check(np.atleast_2d(got), np.atleast_2d(ref), what, say=lambda m: print(m, flush=True))


@pytest.mark.parametrize("name", JOBS)
def test_job_matches_reference(name, tmp_path):
    print(f"[wire] {name}", flush=True)
    wire = _wire()
    spec = wire.DeviceSpec.from_json(os.path.join(GOLDEN, f"{name}.json"))
    out = wire.write_result(wire.solve_job(spec), str(tmp_path / "r.npz"))
    assert wire.validate_result(out) == 3
    got = dict(np.load(out))
    ref = dict(np.load(os.path.join(GOLDEN, f"{name}_ref.npz")))
    assert sorted(got) == sorted(ref)
    for k, v in ref.items():
        if v.dtype.kind == "U" and k not in ("record__meta",):
            if k.endswith("__meta") or k == "continuation__records":
                assert json.loads(str(got[k])) == json.loads(str(v)), k
            else:
                assert str(got[k]) == str(v), k
        else:
            assert got[k].shape == v.shape and got[k].dtype.kind == v.dtype.kind, k
    for k in ("field__doping", "axis_x", "axis_y", "nodes__coords", "mesh__shape",
              "nodes__count", "dimensionality", "solved_bias", "result__schema",
              "band__available"):
        assert np.array_equal(got[k], ref[k]), k
    rec, rrec = json.loads(str(got["record__meta"])), json.loads(str(ref["record__meta"]))
    assert set(rec) == set(rrec) and rec["backend"] == "tcadhip"
    for k in ("schema_version", "dimensionality", "material", "T", "models", "sweep",
              "transient"):
        assert rec[k] == rrec[k], k
    assert np.abs(got["field__potential"] - ref["field__potential"]).max() <= 1e-6
    Ns = max(float(np.abs(ref["field__doping"]).max()), 1.0)
    for k in ("field__electron_density", "field__hole_density"):
        rel = np.abs(got[k] - ref[k]) / np.maximum(ref[k], 1e-10 * Ns)
        assert rel.max() <= 1e-6, (k, rel.max())
    jmax = max(np.abs(ref[f"vector__current_density__{a}"]).max() for a in "xy")
    for a in "xy":
        k = f"vector__current_density__{a}"
        assert np.abs(got[k] - ref[k]).max() <= 1e-6 * jmax, k
    terms = sorted(k for k in ref if k.startswith("terminal__") and k.endswith("__value"))
    _close_currents(np.array([float(got[k]) for k in terms]),
                    np.array([float(ref[k]) for k in terms]), "terminals")
    if "sweep__voltage" in ref:
        assert np.array_equal(got["sweep__converged"], ref["sweep__converged"])
        chans = sorted(k for k in ref if k.startswith("sweep__current__"))
        G = np.array([got[k] for k in chans]).T
        R = np.array([ref[k] for k in chans]).T
        for s in range(R.shape[0]):
            _close_currents(G[s], R[s], f"sweep point {s}")
    if "transient__times" in ref:
        assert np.array_equal(got["transient__times"], ref["transient__times"])
        chans = sorted(k for k in ref if k.startswith("transient__current__"))
        G = np.array([got[k] for k in chans]).T
        R = np.array([ref[k] for k in chans]).T
        big = np.abs(R).max()
        for s in range(R.shape[0]):
            tol = 1e-6 * np.abs(R[s]).max() + 1e-9 * big
            assert np.abs(G[s] - R[s]).max() <= tol, (s, G[s], R[s])
    if "ac__freqs" in ref:
        w = 2 * np.pi * ref["ac__freqs"]
        Yg = got["ac__G"] + 1j * w * got["ac__C"]
        Yr = ref["ac__G"] + 1j * w * ref["ac__C"]
        assert np.all(np.abs(Yg - Yr) <= 1e-6 * np.abs(Yr)), (Yg, Yr)


def test_3d_job_grammar_and_values(tmp_path):
    print("[wire] 3D resistor job", flush=True)
    wire = _wire()
    from tcadhip.device import NewtonOptions
    x = np.linspace(0.0, 2e-4, 8)
    y = np.linspace(0.0, 1e-4, 4)
    z = np.linspace(0.0, 1e-4, 5)
    jj, kk = np.meshgrid(np.arange(y.size), np.arange(z.size))
    jj, kk = jj.ravel().tolist(), kk.ravel().tolist()
    job = {"mesh": {"dimensionality": 3,
                    "axes": {"x": x.tolist(), "y": y.tolist(), "z": z.tolist()}},
           "doping": {"kind": "array", "values": np.full((5, 4, 8), 1e17).tolist()},
           "contacts": [{"name": "left", "kind": "ohmic",
                         "nodes": {"i": [0] * len(jj), "j": jj, "k": kk}},
                        {"name": "right", "kind": "ohmic",
                         "nodes": {"i": [7] * len(jj), "j": jj, "k": kk}}],
           "bias": {"left": 0.05, "right": 0.0},
           "sweep": {"contact": "left", "start": 0.0, "stop": 0.1, "step": 0.05}}
    jpath = str(tmp_path / "job.json")
    with open(jpath, "w") as fh:
        json.dump(job, fh)
    out = wire.run_job(jpath, str(tmp_path / "r.npz"))
    r = wire.ResultFile(out)
    assert r.dimensionality == 3 and r.terminal_current("left")[1] == "A"
    assert r["vector__current_density__z"].shape == (5, 4, 8)
    meta, v, conv, chans, unit = r.sweep()
    assert unit == "A" and conv.all() and v.tolist() == [0.0, 0.05, 0.1]
    assert r["sweep__snapshot__field__potential__2"].shape == (5, 4, 8)
    dev = wire.build_device(wire.DeviceSpec.from_json(jpath))
    dev.solve_equilibrium(NewtonOptions())
    for V in (0.0, 0.05, 0.1):
        dev.solve_bias({"left": V, "right": 0.0}, NewtonOptions())
    I = dev.terminal_currents()
    assert chans["left"][-1] == I["left"] and r.terminal_current("right")[0] == I["right"]
    # an ohmic resistor conducts: I = V / R with R > 0, and |sum I| ~ 0
    assert chans["left"][2] > chans["left"][1] > 0.0
    assert abs(I["left"] + I["right"]) <= 1e-6 * abs(I["left"])


def test_served_jobs_emit_the_same_results():
    """Sweep jobs through the continuous-batching server (request_from_job /
    served_result): valid results whose sweep currents are the bits the
    same job gives through solve_job (the server's per-device reductions
    do not depend on what else is resident), fields at the same state."""
    print("[wire] served jobs", flush=True)
    wire = _wire()
    from tcadhip import serve
    names = ["wire_resistor_2d_sweep", "wire_mosfet_2d_gate_sweep"]
    specs = [wire.DeviceSpec.from_json(os.path.join(GOLDEN, f"{n}.json")) for n in names]
    reqs = [wire.request_from_job(s, rid, arrival=0.0) for rid, s in enumerate(specs)]
    S = serve.Server(20_000)
    served = S.run(reqs, realtime=False)
    for spec, sv in zip(specs, served):
        got = wire.served_result(spec, sv)
        assert wire.validate_result(got) == 3
        ref = wire.solve_job(spec)
        assert sorted(got) == sorted(ref)
        for k in ref:
            if k.startswith("sweep__current__") or k in ("sweep__converged", "sweep__voltage",
                                                         "field__potential",
                                                         "field__electron_density"):
                assert np.array_equal(got[k], ref[k]), k
        terms = sorted(k for k in ref if k.startswith("terminal__") and k.endswith("__value"))
        _close_currents(np.array([float(got[k]) for k in terms]),
                        np.array([float(ref[k]) for k in terms]), "served terminals")
