"""The solver boundary's wire formats on the CPU (tcadhip.wire): the
DeviceSpec job JSON and the npz result grammar.

* every result file the reference's solver process wrote for the golden
  jobs (tests/golden/make_golden_wire.py) passes validate_result, and
  reads back through ResultFile with the reference store's conventions;
* validate_result rejects the structural faults the reference's own
  backend tests plant (pytcad/gui/tests/test_solver_backend.py:230-360),
  and accepts an unstamped legacy file;
* DeviceSpec round-trips the golden job files and refuses a newer
  spec_version; the sweep / waveform / transient / AC validators enforce
  the reference's conditions (device_spec.py);
* check_supported refuses what the GPU devices do not implement.
"""
import json
import os

import numpy as np
import pytest

from tcadhip import wire

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
JOBS = ["wire_resistor_2d_sweep", "wire_diode_2d_ac", "wire_diode_2d_transient",
        "wire_mosfet_2d_gate_sweep"]


@pytest.mark.parametrize("name", JOBS)
def test_reference_results_validate_and_read(name):
    path = os.path.join(GOLDEN, f"{name}_ref.npz")
    assert wire.validate_result(path) == 3
    r = wire.ResultFile(path)
    assert r.dimensionality == 2
    pot, unit = r.scalar_field("potential")
    assert unit == "V" and pot.shape == tuple(a.size for a in reversed(list(r.axes().values())))
    assert r.run_record()["backend"] == "pytcad"
    for t in r.terminal_names():
        assert r.terminal_current(t)[1] == "A/cm"
    if r.has_sweep():
        meta, v, conv, chans, unit = r.sweep()
        assert unit == "A/cm" and meta["dimensionality"] == 2
        assert set(chans) == set(r.terminal_names())
        for c in chans.values():
            assert np.all(np.isnan(c[~conv])) and np.all(np.isfinite(c[conv]))
    if r.has_transient():
        meta, t, chans, unit = r.transient()
        assert meta["contact"] == "anode" and t.size == chans["anode"].size
    if r.has_ac():
        port, f, C, G, uc, ug = r.ac()
        assert port == "anode" and (uc, ug) == ("F/cm^2", "S/cm^2") and C.shape == f.shape


@pytest.mark.parametrize("name", JOBS)
def test_job_files_round_trip(name, tmp_path):
    with open(os.path.join(GOLDEN, f"{name}.json")) as fh:
        raw = json.load(fh)
    spec = wire.DeviceSpec.from_dict(raw)
    d = spec.to_dict()
    assert "spec_version" not in d

    def covers(got, want, where):
        # every value the job file states survives; defaults may be added
        if isinstance(want, dict):
            for k, v in want.items():
                covers(got[k], v, f"{where}.{k}")
        elif isinstance(want, list) and want and isinstance(want[0], dict):
            assert len(got) == len(want), where
            for i, (g, w) in enumerate(zip(got, want)):
                covers(g, w, f"{where}[{i}]")
        else:
            assert got == want, where

    covers(d, raw, "job")
    p = str(tmp_path / "job.json")
    spec.to_json(p)
    assert wire.DeviceSpec.from_json(p).to_dict() == d
    wire.check_supported(spec)


def test_spec_version_gate():
    raw = json.load(open(os.path.join(GOLDEN, "wire_resistor_2d_sweep.json")))
    raw["spec_version"] = 2
    with pytest.raises(wire.UnsupportedSpecVersionError):
        wire.DeviceSpec.from_dict(raw)
    raw["spec_version"] = 1
    assert "spec_version" not in wire.DeviceSpec.from_dict(raw).to_dict()


def test_sweep_points_and_validation():
    # 0.3 / 0.1 is 2.9999...: the endpoint still counts (device_spec.py:55-61)
    assert wire.SweepSpec("a", 0.0, 0.3, 0.1).n_points() == 4
    assert wire.SweepSpec("a", 0.0, 0.25, 0.1).n_points() == 3
    assert wire.SweepSpec("a", 1.0, 0.0, -0.5).voltages() == [1.0, 0.5, 0.0]
    for bad in ({"contact": "a", "start": 0, "stop": 1, "step": 0},
                {"contact": "a", "start": 1, "stop": 1, "step": 0.1},
                {"contact": "a", "start": 0, "stop": 1, "step": -0.1},
                {"contact": "a", "start": 0, "stop": 1e6, "step": 1e-2},
                {"contact": "a", "start": float("nan"), "stop": 1, "step": 0.1},
                {"contact": "a", "start": 0, "stop": 1},
                ["not", "an", "object"]):
        with pytest.raises(ValueError):
            wire.SweepSpec.from_dict(bad)
    with pytest.raises(ValueError, match="not a registered contact"):
        wire.SweepSpec("gate", 0, 1, 0.5).validate(["source", "drain"])


def test_waveform_transient_ac_validation():
    for bad in ({"kind": "square"}, {"kind": "ramp", "t0": 1.0, "t1": 1.0},
                {"kind": "pulse", "t1": 0.0}, {"kind": "step", "v0": float("inf")}):
        with pytest.raises(ValueError):
            wire.WaveformSpec.from_dict(bad)
    wf = {"kind": "step", "v1": 1.0}
    for bad in ({"contact": "a", "waveform": wf, "t_end": 0.0, "dt0": 1e-12},
                {"contact": "a", "waveform": wf, "t_end": 1e-9, "dt0": -1.0},
                {"contact": "a", "waveform": wf, "t_end": 1e-9, "dt0": 1e-12, "theta": 1.5}):
        with pytest.raises(ValueError):
            wire.TransientSpec.from_dict(bad)
    for bad in ({"contact": "a", "f_start": 0.0, "f_stop": 1e6},
                {"contact": "a", "f_start": 1e6, "f_stop": 1e3},
                {"contact": "a", "f_start": 1e3, "f_stop": 1e6, "n_points": 1}):
        with pytest.raises(ValueError):
            wire.ACSpec.from_dict(bad)
    f = wire.ACSpec("a", 1e3, 1e9, 7).freqs()
    assert f.size == 7 and f[0] == pytest.approx(1e3) and f[-1] == pytest.approx(1e9)
    # waveform kinds map to the transient waveforms with the reference's
    # parameter meaning (solver_runner.py:483-495)
    assert wire.WaveformSpec("pulse", 0.0, 1.0, 1e-9, 2e-9).waveform().value(2e-9) == 1.0
    assert wire.WaveformSpec("pulse", 0.0, 1.0, 1e-9, 2e-9).waveform().value(3.5e-9) == 0.0
    assert wire.WaveformSpec("ramp", 0.0, 1.0, 0.0, 2.0).waveform().value(0.5) == 0.25
    assert wire.WaveformSpec("step", 0.0, 1.0, 1.0).waveform().value(0.999) == 0.0
    assert wire.WaveformSpec("constant", 0.7).waveform().value(5.0) == 0.7


def test_unsupported_jobs_are_refused():
    raw = json.load(open(os.path.join(GOLDEN, "wire_resistor_2d_sweep.json")))
    for patch, msg in (({"material": "GaAs"}, "silicon"),
                       ({"region_materials": [{"material": "GaAs", "box": [0, 1e-4]}]},
                        "heterostructure"),
                       ({"models": dict(wire.DEFAULT_MODELS, impact=True)}, "impact"),
                       ({"engine": "quantum"}, "engine")):
        with pytest.raises(ValueError, match=msg):
            wire.check_supported(wire.DeviceSpec.from_dict(dict(raw, **patch)))
    one_d = dict(raw, mesh={"dimensionality": 1, "axes": {"x": [0.0, 1e-4]}})
    with pytest.raises(ValueError, match="2D and 3D"):
        wire.check_supported(wire.DeviceSpec.from_dict(one_d))


def test_edges_to_nodes():
    a = np.array([[1.0, 3.0, 5.0]])
    assert np.array_equal(wire._edges_to_nodes(a, 1), [[1.0, 2.0, 4.0, 5.0]])
    b = np.arange(6.0).reshape(2, 3)
    assert np.array_equal(wire._edges_to_nodes(b, 0), [[0, 1, 2], [1.5, 2.5, 3.5], [3, 4, 5]])


def _minimal(path, **extra):
    d = {"dimensionality": np.array(1), "axis_x": np.array([0.0, 1e-4]),
         "field__potential": np.array([0.0, 1.0]), "unit__potential": np.array("V"),
         "solved_bias": np.array(False)}
    d.update(extra)
    np.savez(str(path), **d)
    return str(path)


def test_validator_faults(tmp_path):
    assert wire.validate_result(_minimal(tmp_path / "ok.npz")) == 1      # unstamped legacy
    faults = [
        ({"result__schema": np.array(99)}, "unsupported"),
        ({"unit__potential": None}, "no declared unit"),
        ({"field__potential": np.array([0.0, 1.0, 2.0])}, "does not match"),
        ({"solved_bias": None}, "solved_bias"),
        ({"terminal__a__value": np.array(1.0)}, "has no unit"),
        ({"terminal__a__unit": np.array("A/cm")}, "orphan"),
        ({"sweep__voltage": np.array([0.0, 0.1])}, "incomplete sweep"),
        ({"sweep__voltage": np.array([0.0, 0.1]), "sweep__converged": np.array([True, True]),
          "unit__sweep_current": np.array("A"),
          "sweep__meta": np.array(json.dumps({"dimensionality": 1})),
          "sweep__current__x": np.array([1.0])}, "does not match"),
        ({"transient__times": np.array([0.0])}, "incomplete transient"),
        ({"ac__freqs": np.array([1.0])}, "incomplete ac"),
        ({"geom__kind": np.array("point_cloud")}, "reserved"),
        ({"nodes__count": np.array(5)}, "disagrees"),
        ({"record__meta": np.array("{not json")}, "not valid JSON"),
        ({"vector__current_density__x": np.array([0.0, 1.0])}, "has no unit"),
    ]
    for i, (patch, msg) in enumerate(faults):
        d = {"dimensionality": np.array(1), "axis_x": np.array([0.0, 1e-4]),
             "field__potential": np.array([0.0, 1.0]), "unit__potential": np.array("V"),
             "solved_bias": np.array(False)}
        for k, v in patch.items():
            if v is None:
                d.pop(k)
            else:
                d[k] = v
        p = str(tmp_path / f"bad{i}.npz")
        np.savez(p, **d)
        with pytest.raises(wire.ResultSchemaError, match=msg):
            wire.validate_result(p)
    d = {"dimensionality": np.array(2), "axis_x": np.array([0.0, 1e-4]),
         "solved_bias": np.array(False)}
    np.savez(str(tmp_path / "noaxis.npz"), **d)
    with pytest.raises(wire.ResultSchemaError, match="axis_y"):
        wire.validate_result(str(tmp_path / "noaxis.npz"))
    with pytest.raises(wire.ResultSchemaError, match="not found"):
        wire.validate_result(str(tmp_path / "missing.npz"))
    (tmp_path / "junk.npz").write_text("not an archive")
    with pytest.raises(wire.ResultSchemaError, match="not a readable"):
        wire.validate_result(str(tmp_path / "junk.npz"))


def test_write_result_is_atomic(tmp_path):
    out = str(tmp_path / "r.npz")
    d = {"dimensionality": np.array(1), "axis_x": np.array([0.0, 1e-4]),
         "solved_bias": np.array(False)}
    wire.write_result(d, out)
    assert os.path.exists(out) and not os.path.exists(out + ".tmp.npz")
    assert wire.validate_result(out) == 1
