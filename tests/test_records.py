"""Reference-format records (metrics.jsonl, trajectory.jsonl, fitting-project
state and settings): field names and order of the serde types
(src/vbd.rs:18-25,61-83, src/ogc/world.rs:14-22,33-44, src/model.rs:116-121),
the members the reference's motion summary reads (scripts/analyze-motion.py:
38-68), and the reference's checks and messages (src/fitting.rs:117-238)."""
import io
import json
import math

import numpy as np
import pytest

from clothcraft_amd import abi, records
from clothcraft_amd.scenes import State
from clothcraft_amd.vbd import VbdSettings


def _metrics(**kw):
    m = abi.cv_metrics()
    vals = dict(time_s=0.05, dt_s=1 / 60, step_ms=3.5, detection_ms=1.25, solve_ms=2.0,
                gpu_compute_ms=float("nan"), refit_ms=0.1, candidates_ms=0.2, narrow_ms=0.3,
                candidate_pairs=1234, aabb_rejections=0, iterations=20, contacts=77, contact_terms=80,
                detection_updates=2, broad_phase_rebuilds=2, truncated_vertices=3,
                minimum_distance_m=0.0015, max_seam_gap_m=0.0002, extension_p95_percent=1.5,
                extension_max_percent=3.0, residual_rms_n=0.01, colors=7)
    vals.update(kw)
    for k, v in vals.items():
        setattr(m, k, v)
    return m.as_dict()


def test_metrics_line_is_the_serde_vbd_metrics_object():
    line = records.metrics_line(_metrics())
    rec = json.loads(line)
    assert list(rec) == ["time_s", "dt_s", "step_ms", "detection_ms", "contact_profile", "solve_ms",
                         "gpu_compute_ms", "iterations", "contacts", "contact_terms", "detection_updates",
                         "broad_phase_rebuilds", "truncated_vertices", "minimum_distance_m", "max_seam_gap_m",
                         "extension_p95_percent", "extension_max_percent", "residual_rms_n"]
    assert list(rec["contact_profile"]) == ["refit_ms", "candidates_ms", "narrow_ms", "candidate_pairs",
                                            "aabb_rejections"]
    assert rec["gpu_compute_ms"] is None  # Option<f64>::None
    assert rec["residual_rms_n"] == 0.01 and rec["iterations"] == 20 and isinstance(rec["contacts"], int)
    assert rec["contact_profile"]["candidate_pairs"] == 1234
    assert "colors" not in rec  # extensions stay out of the reference's type
    assert "\n" not in line
    none = json.loads(records.metrics_line(_metrics(residual_rms_n=float("nan"))))
    assert none["residual_rms_n"] is None


def test_logs_carry_what_the_motion_summary_reads():
    out = io.StringIO()
    w = records.JsonlWriter(out)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(10, 3))
    states = []
    for k in range(4):
        st = State(x + 0.01 * k, np.full((10, 3), 0.5 * k), 0.05 * k)
        states.append(st)
        w.state(st)
    lines = out.getvalue().splitlines()
    got = records.read_jsonl(lines)
    for st, rec in zip(states, got):
        assert len(rec["x"]) == 10 and len(rec["x"][0]) == 3
        assert np.array_equal(np.array(rec["x"]), st.x) and rec["time"] == st.time  # exact f64 round trip
        back = records.state_from_record(rec, 10)
        assert np.array_equal(back.v, st.v)
    out = io.StringIO()
    w = records.JsonlWriter(out)
    for k in range(3):
        w.metrics(_metrics(step_ms=1.0 + k))
    ms = records.read_jsonl(out.getvalue().splitlines())
    for name in ("iterations", "dt_s", "step_ms", "detection_ms", "solve_ms", "max_seam_gap_m",
                 "extension_p95_percent", "minimum_distance_m"):
        assert all(isinstance(m[name], (int, float)) for m in ms), name
    assert [m["step_ms"] for m in ms] == [1.0, 2.0, 3.0]


def _project(**kw):
    doc = {"version": 1, "template": {"panels": []}, "avatar_path": "avatar.obj", "contact_stiffness": 1e5,
           "contact_clearance": 0.0, "prearrange": False, "quasistatic": False, "solver": {"dt": 0.01},
           "simulation": {"dt": 1 / 240, "iterations": 16, "damping": 1.5, "gravity": [0.0, -9.81, 0.0],
                          "contact": {"radius": 0.003, "stiffness": 1e5, "relaxation": 0.45,
                                      "refresh_fraction": 0.01}},
           "state": None, "mesh_signature": "fnv1a:0123456789abcdef"}
    doc.update(kw)
    return doc


def test_project_state_and_settings_round_trip_and_keep_other_members():
    doc = _project()
    assert records.project_state(doc) is None
    s = records.project_settings(doc)
    assert s == VbdSettings()  # the reference defaults (src/vbd.rs:26-36, src/ogc/world.rs:45-54)
    assert records.settings_record(s) == doc["simulation"]
    st = State(np.arange(12.0).reshape(4, 3), np.ones((4, 3)), 1.25)
    new = json.loads(json.dumps(records.with_project_state(doc, st, 4)))
    assert {k: v for k, v in new.items() if k != "state"} == {k: v for k, v in doc.items() if k != "state"}
    back = records.project_state(new, 4)
    assert np.array_equal(back.x, st.x) and np.array_equal(back.v, st.v) and back.time == 1.25
    assert records.project_settings(_project(simulation=None)) is None


def test_project_checks_use_the_reference_messages():
    st = State(np.zeros((4, 3)), np.zeros((4, 3)), 0.0)
    with pytest.raises(ValueError, match="^Unsupported fitting project version$"):
        records.project_state(_project(version=2))
    with pytest.raises(ValueError, match="^Invalid cached fitting state$"):
        records.project_state(_project(state=records.state_record(st)), 5)
    bad = records.state_record(st)
    bad["time"] = math.inf
    with pytest.raises(ValueError, match="^Invalid cached fitting state$"):
        records.state_from_record(json.loads(json.dumps(bad)))
    with pytest.raises(ValueError, match="^Invalid fitting state; existing document was preserved$"):
        records.with_project_state(_project(), State(np.full((4, 3), np.nan), np.zeros((4, 3)), 0.0))
    with pytest.raises(ValueError, match="mesh signature"):
        records.with_project_state(_project(mesh_signature=None), st)
    sim = _project()["simulation"]
    with pytest.raises(ValueError, match="^Invalid VBD settings$"):
        records.settings_from_record(dict(sim, dt=0.03))
    with pytest.raises(ValueError, match="^Invalid OGC radius, stiffness or trust-region parameters$"):
        records.settings_from_record(dict(sim, contact=dict(sim["contact"], relaxation=0.5)))
