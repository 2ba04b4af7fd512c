"""Reference-format records around the step: the per-step metrics lines of
`metrics.jsonl`, the state lines of `trajectory.jsonl`, and the `state` and
`simulation` members of a fitting-project JSON document.

* metrics.jsonl — one serde-JSON `VbdMetrics` object per step, written by
  the fitting CLI after every step (src/bin/cloth-fit.rs:127-140); fields and
  their order as declared in src/vbd.rs:61-83, the detection profile as a
  nested `contact_profile` object (src/ogc/world.rs:14-22), and `null` for
  the optional `gpu_compute_ms` and `residual_rms_n`.
* trajectory.jsonl / project state — serde-JSON `State` (src/model.rs:116-121):
  `{"x": [[x, y, z], ...], "v": [[...], ...], "time": t}` (glam's DVec3
  serialises as a 3-array).
* fitting project — `FittingProject` (src/fitting.rs:18-42): this module
  reads and replaces its `state` (may be null) and reads its optional
  `simulation` (`VbdSettings`, src/vbd.rs:18-25, with `contact` the
  `OgcSettings` of src/ogc/world.rs:33-44); every other member (template,
  avatar path, solver settings) is carried through untouched.  The
  document's `mesh_signature` fingerprints the compiled template and mesh,
  not the state (src/fitting.rs:213-232, 244-270), so a state can replace
  the state of a document that already has a signature and the document
  still loads in the reference; computing a signature needs the template
  compiler, which is not part of this library.

The readers check what they read with the reference's rules and messages
(version 1, one finite position and velocity per vertex, finite time).
"""
from __future__ import annotations

import json
import math
from typing import IO, Iterable

import numpy as np

from .scenes import State
from .vbd import OgcSettings, VbdSettings

# src/vbd.rs:61-83, in declaration order (serde keeps it).
METRICS_FIELDS = ("time_s", "dt_s", "step_ms", "detection_ms", "contact_profile", "solve_ms", "gpu_compute_ms",
                  "iterations", "contacts", "contact_terms", "detection_updates", "broad_phase_rebuilds",
                  "truncated_vertices", "minimum_distance_m", "max_seam_gap_m", "extension_p95_percent",
                  "extension_max_percent", "residual_rms_n")
# src/ogc/world.rs:14-22.
PROFILE_FIELDS = ("refit_ms", "candidates_ms", "narrow_ms", "candidate_pairs", "aabb_rejections")
_OPTIONAL = ("gpu_compute_ms", "residual_rms_n")
_INTEGER = ("iterations", "contacts", "contact_terms", "detection_updates", "broad_phase_rebuilds",
            "truncated_vertices", "candidate_pairs", "aabb_rejections")


def _number(name: str, value):
    if name in _INTEGER:
        return int(value)
    value = float(value)
    if name in _OPTIONAL and math.isnan(value):
        return None
    if not math.isfinite(value):
        # serde_json writes non-finite floats as null as well
        return None
    return value


def metrics_record(m: dict) -> dict:
    """A step's metrics (the dict GpuVbd.step returns) as the reference's
    serde object: its fields only, in its order, the optional ones null when
    absent (NaN in the C struct)."""
    out = {}
    for name in METRICS_FIELDS:
        if name == "contact_profile":
            out[name] = {p: _number(p, m[p]) for p in PROFILE_FIELDS}
        else:
            out[name] = _number(name, m[name])
    return out


def metrics_line(m: dict) -> str:
    """One metrics.jsonl line (no trailing newline)."""
    return json.dumps(metrics_record(m), separators=(",", ":"))


def state_record(state: State) -> dict:
    """serde-JSON State: positions and velocities as lists of 3-arrays."""
    return {"x": np.asarray(state.x, np.float64).tolist(), "v": np.asarray(state.v, np.float64).tolist(),
            "time": float(state.time)}


def state_from_record(rec: dict, n: int | None = None) -> State:
    """State from its serde-JSON object; n (when given) is the vertex count
    the garment expects."""
    try:
        x = np.ascontiguousarray(np.asarray(rec["x"], np.float64).reshape(-1, 3))
        v = np.ascontiguousarray(np.asarray(rec["v"], np.float64).reshape(-1, 3))
        t = float(rec["time"])
    except (KeyError, TypeError, ValueError) as e:
        raise ValueError(f"invalid state record: {e}") from None
    if (len(x) != len(v) or (n is not None and len(x) != n)
            or not (np.isfinite(x).all() and np.isfinite(v).all() and math.isfinite(t))):
        raise ValueError("Invalid cached fitting state")  # src/fitting.rs:197-204
    return State(x, v, t)


class JsonlWriter:
    """Appends one JSON object per line, as the fitting CLI's logs do."""

    def __init__(self, stream: IO[str]):
        self.stream = stream

    def metrics(self, m: dict) -> None:
        self.stream.write(metrics_line(m) + "\n")

    def state(self, state: State) -> None:
        self.stream.write(json.dumps(state_record(state), separators=(",", ":")) + "\n")


def read_jsonl(lines: Iterable[str]) -> list:
    return [json.loads(line) for line in lines if line.strip()]


def settings_from_record(rec: dict) -> VbdSettings:
    """VbdSettings from its serde-JSON object (src/vbd.rs:18-25), validated by
    the reference's rules (src/vbd.rs:37-51, src/ogc/world.rs:56-71)."""
    try:
        c = rec["contact"]
        s = VbdSettings(float(rec["dt"]), int(rec["iterations"]), float(rec["damping"]),
                        tuple(float(g) for g in rec["gravity"]),
                        OgcSettings(float(c["radius"]), float(c["stiffness"]), float(c["relaxation"]),
                                    float(c["refresh_fraction"])))
    except (KeyError, TypeError, ValueError) as e:
        raise ValueError(f"invalid VBD settings record: {e}") from None
    c = s.contact
    if not (math.isfinite(c.radius) and c.radius > 0 and math.isfinite(c.stiffness) and c.stiffness > 0
            and math.isfinite(c.relaxation) and 0 < c.relaxation < 0.5
            and math.isfinite(c.refresh_fraction) and 0 <= c.refresh_fraction <= 1):
        raise ValueError("Invalid OGC radius, stiffness or trust-region parameters")
    if not (math.isfinite(s.dt) and 1e-6 <= s.dt <= 0.025 and 1 <= s.iterations <= 200
            and math.isfinite(s.damping) and s.damping >= 0 and len(s.gravity) == 3
            and all(math.isfinite(g) for g in s.gravity)):
        raise ValueError("Invalid VBD settings")
    return s


def settings_record(s: VbdSettings) -> dict:
    """serde-JSON VbdSettings (the library's extension fields are not part of
    the reference type and are left out)."""
    c = s.contact
    return {"dt": s.dt, "iterations": s.iterations, "damping": s.damping, "gravity": list(s.gravity),
            "contact": {"radius": c.radius, "stiffness": c.stiffness, "relaxation": c.relaxation,
                        "refresh_fraction": c.refresh_fraction}}


def _check_version(doc: dict) -> None:
    if doc.get("version") != 1:
        raise ValueError("Unsupported fitting project version")  # src/fitting.rs:130-132


def project_state(doc: dict, n: int | None = None) -> State | None:
    """The saved state of a fitting-project document (None when it has
    none); n is the garment's vertex count."""
    _check_version(doc)
    rec = doc.get("state")
    return None if rec is None else state_from_record(rec, n)


def project_settings(doc: dict) -> VbdSettings | None:
    """The interactive-backend settings of a fitting-project document (None
    in documents that predate them)."""
    _check_version(doc)
    rec = doc.get("simulation")
    return None if rec is None else settings_from_record(rec)


def with_project_state(doc: dict, state: State, n: int | None = None) -> dict:
    """A copy of the document with its state replaced; every other member is
    kept as it was.  The document must already carry a mesh signature (see
    the module notes); n is the garment's vertex count."""
    _check_version(doc)
    if not doc.get("mesh_signature"):
        raise ValueError("fitting project has no mesh signature: save it from the reference with a state first")
    x, v = np.asarray(state.x, np.float64), np.asarray(state.v, np.float64)
    if (x.shape != v.shape or x.ndim != 2 or x.shape[1] != 3 or (n is not None and len(x) != n)
            or not (np.isfinite(x).all() and np.isfinite(v).all() and math.isfinite(state.time))):
        raise ValueError("Invalid fitting state; existing document was preserved")  # src/fitting.rs:224-238
    out = dict(doc)
    out["state"] = state_record(state)
    return out
