"""Reference-format garment templates compiled into the solver's garment.

`compile_template` restates `GarmentTemplate::compile`
(src/garment/template.rs:214-451) for templates whose panels are grid
patterns (`Pattern::build`, restated by scenes.pattern_build): fabrics bound
to each panel's grain, panel placement, corner pins, and the seams of
src/garment/seam.rs — rest-arclength boundary curves (:80-236), the common
partition of both sides with notch correspondences, two-point Gauss stitch
samples and gap probes (:321-407).  The result is the `CompiledGarment` the
solver consumes: one `Garment` with per-panel materials, the stitch terms
`VbdSolver::new` builds from the seam samples (src/vbd.rs:128-144:
coefficient = seam stiffness x sample weight) and the seam probes of
`max_seam_gap`.  Polygon panels (template version 2) need the reference's
constrained-Delaunay mesher and are rejected with a clear error.

Error messages are the reference's `Err(String)` texts.
"""
from __future__ import annotations

import bisect
import json
import math
import re
from dataclasses import dataclass, field

import numpy as np

from .scenes import Garment, Material, State, euler_xyz, pattern_build

FABRIC_FIELDS = ("warp", "weft", "shear", "compression_ratio", "compression_transition", "bend_warp",
                 "bend_weft", "density", "thickness")


class TemplateError(ValueError):
    pass


# ------------------------------------------------------------ boundary curves
@dataclass
class BoundaryCurve:
    """src/garment/seam.rs:80-236: a directed boundary polyline parametrised
    by normalised rest arclength, trimmed to `range`, or a chain of them."""
    chain: list
    vertices: list
    knots: list
    range: tuple
    reverse: bool
    length: float

    @staticmethod
    def new(uv, vertices, rng, reverse):
        if (len(vertices) < 2 or any(i >= len(uv) for i in vertices)
                or not all(math.isfinite(v) for v in rng) or rng[0] < 0.0 or rng[1] > 1.0
                or rng[1] - rng[0] < 1e-9):
            raise TemplateError("Boundary requires valid vertices and 0 <= start < end <= 1")
        knots = [0.0]
        for a, b in zip(vertices[:-1], vertices[1:]):
            length = math.hypot(uv[a][0] - uv[b][0], uv[a][1] - uv[b][1])
            if not math.isfinite(length) or length <= 1e-12:
                raise TemplateError("Zero-length boundary segment")
            knots.append(knots[-1] + length)
        total = knots[-1]
        knots = [k / total for k in knots]
        return BoundaryCurve([], list(vertices), knots, (rng[0], rng[1]), bool(reverse),
                             total * (rng[1] - rng[0]))

    def breakpoints(self):
        if self.chain:
            offset, out = 0.0, []
            for c in self.chain:
                out += [(offset + t * c.length) / self.length for t in c.breakpoints()]
                offset += c.length
            return _dedup(sorted(out))
        out = [0.0, 1.0]
        for k in self.knots:
            if self.range[0] < k < self.range[1]:
                t = (k - self.range[0]) / (self.range[1] - self.range[0])
                out.append(1.0 - t if self.reverse else t)
        return sorted(out)

    def point(self, t, left=False):
        """Two (vertex, weight) pairs of the boundary point at parameter t;
        at a joint of a chain, `left` takes the end of the earlier segment."""
        if self.chain:
            distance = min(max(t, 0.0), 1.0) * self.length
            offset = 0.0
            for i, c in enumerate(self.chain):
                end = offset + c.length
                if distance < end - 1e-12 or (left and distance <= end + 1e-12) or i + 1 == len(self.chain):
                    return c.point((distance - offset) / c.length, left)
                offset = end
        t = 1.0 - t if self.reverse else t
        s = self.range[0] + min(max(t, 0.0), 1.0) * (self.range[1] - self.range[0])
        i = min(max(bisect.bisect_right(self.knots, s) - 1, 0), len(self.vertices) - 2)
        w = (s - self.knots[i]) / (self.knots[i + 1] - self.knots[i])
        return [(self.vertices[i], 1.0 - w), (self.vertices[i + 1], w)]

    @staticmethod
    def concatenate(chain, reverse):
        flat = []
        for c in chain:
            flat += c.chain if c.chain else [c]
        if not flat or len(flat) > 128:
            raise TemplateError("Sewing side needs 1..128 segments")
        if reverse:
            flat = flat[::-1]
            for c in flat:
                c.reverse = not c.reverse
        if len(flat) == 1:
            return flat[0]
        return BoundaryCurve(flat, [], [], (0.0, 1.0), False, sum(c.length for c in flat))

    def joints(self):
        out, acc = [], 0.0
        for c in self.chain[:-1]:
            acc += c.length
            out.append(acc / self.length)
        return out


def _dedup(values, tol=1e-12):
    """Vec::dedup_by with |a - b| < tol: drops a value close to the last kept one."""
    out = []
    for v in values:
        if not out or not abs(v - out[-1]) < tol:
            out.append(v)
    return out


def _interpolate(knots, u, frm):
    """src/garment/seam.rs:285-292."""
    col = [k[frm] for k in knots]
    i = min(max(bisect.bisect_right(col, u) - 1, 0), len(knots) - 2)
    t = (u - knots[i][frm]) / (knots[i + 1][frm] - knots[i][frm])
    return knots[i][1 - frm] * (1.0 - t) + knots[i + 1][1 - frm] * t


def _stitch_point(a, b, s, t, left=False):
    """StitchPoint::limit (src/garment/seam.rs:216-230): a's point minus b's,
    merged per vertex in id order, near-zero coefficients dropped."""
    terms: dict = {}
    for v, w in a.point(s, left):
        terms[v] = terms.get(v, 0.0) + w
    for v, w in b.point(t, left):
        terms[v] = terms.get(v, 0.0) - w
    return [(v, c) for v, c in sorted(terms.items()) if abs(c) > 1e-15]


@dataclass
class CompiledSeam:
    id: str
    samples: list          # (terms, weight_m)
    probes: list           # terms
    stiffness: float
    lengths: tuple
    notches: list


def compile_seam(spec, a, b) -> CompiledSeam:
    """CompiledSeam::compile (src/garment/seam.rs:321-407)."""
    k = spec["stiffness"]
    if not math.isfinite(k) or k <= 0.0:
        raise TemplateError("Seam stiffness must be positive and finite (N/m²)")
    ease = spec.get("allow_ease", False)
    if not ease and abs(a.length - b.length) > 1e-8 * max(a.length, b.length):
        raise TemplateError(f"Seam {spec['id']} has unequal lengths ({a.length:.6f}, {b.length:.6f} m); "
                            f"set allow_ease explicitly")
    knots = [(0.0, 0.0)]
    for kn in spec.get("notches", []):
        if any(not math.isfinite(kn[i]) or kn[i] <= knots[-1][i] + 1e-9 or kn[i] >= 1.0 - 1e-9 for i in (0, 1)):
            raise TemplateError(f"Seam {spec['id']} notches must be strictly increasing inside (0,1) on both edges")
        knots.append((kn[0], kn[1]))
    knots.append((1.0, 1.0))
    if not ease:
        for k0, k1 in zip(knots[:-1], knots[1:]):
            la, lb = a.length * (k1[0] - k0[0]), b.length * (k1[1] - k0[1])
            if abs(la - lb) > 1e-8 * max(la, lb):
                raise TemplateError(f"Seam {spec['id']} notches introduce local easing; set allow_ease explicitly")
    partition = a.breakpoints() + [_interpolate(knots, t, 1) for t in b.breakpoints()] + [kn[0] for kn in knots]
    partition = _dedup(sorted(partition))
    q = 1.0 / math.sqrt(3.0)
    samples = []
    for lo, hi in zip(partition[:-1], partition[1:]):
        blo, bhi = _interpolate(knots, lo, 0), _interpolate(knots, hi, 0)
        weight_m = (a.length * (hi - lo) + b.length * (bhi - blo)) / 4.0
        for g in (-q, q):
            s = (lo + hi) * 0.5 + g * (hi - lo) * 0.5
            samples.append((_stitch_point(a, b, s, _interpolate(knots, s, 0)), weight_m))
    probes = [_stitch_point(a, b, s, _interpolate(knots, s, 0)) for s in partition]
    joints = a.joints() + [_interpolate(knots, t, 1) for t in b.joints()]
    probes += [_stitch_point(a, b, s, _interpolate(knots, s, 0), left=True) for s in joints]
    return CompiledSeam(spec["id"], samples, probes, k, (a.length, b.length), list(spec.get("notches", [])))


# --------------------------------------------------------------- templates
def boundary_vertices(nx, ny, side, offset):
    """Panel::boundary_vertices (src/garment/template.rs:127-136)."""
    if side == "top":
        v = list(range(nx + 1))
    elif side == "right":
        v = [j * (nx + 1) + nx for j in range(ny + 1)]
    elif side == "bottom":
        v = [ny * (nx + 1) + i for i in range(nx, -1, -1)]
    elif side == "left":
        v = [j * (nx + 1) for j in range(ny, -1, -1)]
    else:
        raise TemplateError(f"unknown side {side}")
    return [i + offset for i in v]


CORNERS = {"top_left": lambda nx, ny: 0, "top_right": lambda nx, ny: nx,
           "bottom_right": lambda nx, ny: ny * (nx + 1) + nx, "bottom_left": lambda nx, ny: ny * (nx + 1)}

_ID = re.compile(r"^[A-Za-z0-9_-]{1,80}$")


def _unique_ids(ids, kind):
    seen = set()
    for i in ids:
        if not _ID.match(i) or i in seen:
            raise TemplateError(f"Invalid or duplicate {kind} id: {i}")
        seen.add(i)


def _references(spec):
    return [spec["a"]] + spec.get("a_tail", []) + [spec["b"]] + spec.get("b_tail", [])


@dataclass
class CompiledGarment:
    """The parts of CompiledGarment (src/garment/template.rs:170-180) the step uses."""
    garment: Garment
    state: State
    seams: list
    panels: list                   # (id, vertex range, triangle range, material index)
    warnings: list = field(default_factory=list)


def load_template(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def compile_template(t: dict) -> CompiledGarment:
    """GarmentTemplate::compile (src/garment/template.rs:214-451) for grid panels."""
    if t.get("format") != "marvelous-cloth-template" or t.get("version") not in (1, 2):
        raise TemplateError("Unsupported garment template format/version")
    panels, fabrics, seams = t["panels"], t["fabrics"], t["seams"]
    if t["version"] == 1 and (any(p.get("polygon") is not None or p.get("pin_points") for p in panels)
                              or any(r.get("edge") is not None for s in seams for r in _references(s))):
        raise TemplateError("Polygon geometry and material pins require template version 2")
    if not panels or len(panels) > 128 or not fabrics or len(seams) > 512:
        raise TemplateError("Template needs 1..128 panels, fabrics and at most 512 seams")
    _unique_ids([p["id"] for p in panels], "panel")
    _unique_ids([f["id"] for f in fabrics], "fabric")
    _unique_ids([s["id"] for s in seams], "seam")
    for f in fabrics:
        Material(**{k: f["properties"][k] for k in FABRIC_FIELDS}).validate()
    parts, ranges, x_all, materials = [], [], [], []
    nv = nt = 0
    for p in panels:
        if p.get("polygon") is not None:
            raise TemplateError(f"Panel {p['id']}: polygon panels need the reference's constrained-Delaunay "
                                f"mesher, which this build does not include")
        fab = next((f for f in fabrics if f["id"] == p["fabric"]), None)
        if fab is None:
            raise TemplateError(f"Panel {p['id']} references missing fabric {p['fabric']}")
        pl = p["placement"]
        if not all(math.isfinite(v) for v in list(pl["translation"]) + list(pl["rotation_degrees"])):
            raise TemplateError(f"Non-finite placement for {p['id']}")
        pins = p.get("pins", [])
        if len(set(pins)) != len(pins):
            raise TemplateError(f"Duplicate pin in {p['id']}")
        mat = Material(**{k: fab["properties"][k] for k in FABRIC_FIELDS}, grain_degrees=p["grain_degrees"])
        pat = p["pattern"]
        try:
            g, _ = pattern_build(pat["width"], pat["height"], pat["nx"], pat["ny"], mat, "sphere")
        except ValueError as e:
            raise TemplateError(str(e)) from None
        # x = R (u, -v, 0) + t, R = Rx Ry Rz (src/garment/template.rs:79-92)
        local = np.stack([g.uv[:, 0], -g.uv[:, 1], np.zeros(g.n)], axis=1)
        x = local @ euler_xyz(pl["rotation_degrees"]).T + np.asarray(pl["translation"], np.float64)
        if not np.isfinite(x).all():
            raise TemplateError("Placement overflow")
        g.pinned[:] = 0
        g.pin_targets[:] = 0
        for c in pins:
            i = CORNERS[c](pat["nx"], pat["ny"])
            g.pinned[i] = 1
            g.pin_targets[i] = x[i]
        parts.append((g, nv, nt, len(materials)))
        ranges.append((p["id"], (nv, nv + g.n), (nt, nt + len(g.tri_area)), len(materials)))
        x_all.append(x)
        materials.append(mat)
        nv += g.n
        nt += len(g.tri_area)
    uv = np.vstack([g.uv for g, *_ in parts])
    out = Garment(
        uv=uv, masses=np.concatenate([g.masses for g, *_ in parts]),
        pinned=np.concatenate([g.pinned for g, *_ in parts]),
        pin_targets=np.vstack([g.pin_targets for g, *_ in parts]),
        tri_vertices=np.vstack([g.tri_vertices + v0 for g, v0, _, _ in parts]).astype(np.uint32),
        tri_inv_dm=np.vstack([g.tri_inv_dm for g, *_ in parts]),
        tri_area=np.concatenate([g.tri_area for g, *_ in parts]),
        tri_material=np.concatenate([np.full(len(g.tri_area), m, np.uint32) for g, _, _, m in parts]),
        hinge_vertices=np.vstack([g.hinge_vertices + v0 for g, v0, _, _ in parts]).astype(np.uint32),
        hinge_weight=np.concatenate([g.hinge_weight for g, *_ in parts]),
        hinge_across=np.vstack([g.hinge_across for g, *_ in parts]),
        hinge_material=np.concatenate([np.full(len(g.hinge_weight), m, np.uint32) for g, _, _, m in parts]),
        materials=materials)
    state = State(np.ascontiguousarray(np.vstack(x_all)), np.zeros((nv, 3)), 0.0)
    # Seams (src/garment/template.rs:370-448).
    index = {p["id"]: i for i, p in enumerate(panels)}

    def curve(ref):
        if ref["panel"] not in index:
            raise TemplateError(f"Missing seam panel {ref['panel']}")
        i = index[ref["panel"]]
        pat = panels[i]["pattern"]
        verts = boundary_vertices(pat["nx"], pat["ny"], ref["side"], ranges[i][1][0])
        return BoundaryCurve.new(uv, verts, ref["range"], bool(ref.get("reversed", False)))

    compiled, warnings = [], []
    for si, spec in enumerate(seams):
        if len(spec.get("a_tail", [])) >= 128 or len(spec.get("b_tail", [])) >= 128:
            raise TemplateError("Sewing side needs 1..128 segments")
        refs = _references(spec)
        previous = [r for s in seams[:si] for r in _references(s)]
        for j, r in enumerate(refs):
            for o in previous + refs[:j]:
                same = (r.get("edge") == o.get("edge")) if (r.get("edge") or o.get("edge")) else r["side"] == o["side"]
                if r["panel"] == o["panel"] and same and \
                        max(r["range"][0], o["range"][0]) < min(r["range"][1], o["range"][1]) - 1e-12:
                    raise TemplateError(f"Overlapping seam ownership on {r['panel']} {r['side'].capitalize()}")
        a = BoundaryCurve.concatenate([curve(r) for r in [spec["a"]] + spec.get("a_tail", [])], False)
        b = BoundaryCurve.concatenate([curve(r) for r in [spec["b"]] + spec.get("b_tail", [])],
                                      spec["reverse_b"])
        seam = compile_seam(spec, a, b)
        if abs(seam.lengths[0] - seam.lengths[1]) > 1e-8 * max(seam.lengths) or \
                any(abs(kn[0] - kn[1]) > 1e-8 for kn in seam.notches):
            warnings.append(f"{seam.id}: unequal total or notch-interval lengths ({seam.lengths[0]:.3f} / "
                            f"{seam.lengths[1]:.3f} m total); no physical gathering model")
        compiled.append(seam)
    # VbdSolver::new's stitch terms (src/vbd.rs:128-144) and the probes of
    # CompiledGarment::max_seam_gap (src/garment/template.rs:536-538).
    so, sv, sw, sk = [0], [], [], []
    po, pv, pw = [0], [], []
    for seam in compiled:
        for terms, weight_m in seam.samples:
            sv += [v for v, _ in terms]
            sw += [c for _, c in terms]
            so.append(len(sv))
            sk.append(seam.stiffness * weight_m)
        for terms in seam.probes:
            pv += [v for v, _ in terms]
            pw += [c for _, c in terms]
            po.append(len(pv))
    out.stitch_offsets = np.array(so, np.uint32)
    out.stitch_vertices = np.array(sv, np.uint32)
    out.stitch_weights = np.array(sw, np.float64)
    out.stitch_stiffness = np.array(sk, np.float64)
    out.probe_offsets = np.array(po, np.uint32)
    out.probe_vertices = np.array(pv, np.uint32)
    out.probe_weights = np.array(pw, np.float64)
    return CompiledGarment(out, state, compiled, ranges, warnings)


def two_panel_demo() -> dict:
    """GarmentTemplate::two_panel_demo (src/garment/template.rs:453-533): 3x4
    and 4x7 cells, independent 3D placements, one notched centre seam."""
    props = {k: getattr(Material(), k) for k in FABRIC_FIELDS}

    def panel(pid, name, nx, ny, grain, translation, rotation, pins):
        return {"id": pid, "name": name, "pattern": {"width": 0.4, "height": 0.6, "nx": nx, "ny": ny},
                "fabric": "woven-demo", "grain_degrees": grain,
                "placement": {"translation": translation, "rotation_degrees": rotation}, "pins": pins}

    return {"format": "marvelous-cloth-template", "version": 1, "name": "Two panels / nonmatching seam",
            "fabrics": [{"id": "woven-demo", "name": "Uncalibrated woven reference", "properties": props}],
            "panels": [panel("left", "Left panel", 3, 4, 0.0, [-0.4, 1.0, 0.0], [0.0, 0.0, 0.0],
                             ["top_left", "top_right"]),
                       panel("right", "Right panel", 4, 7, 45.0, [0.08, 1.0, 0.03], [8.0, 12.0, 0.0], [])],
            "seams": [{"id": "centre-seam",
                       "a": {"panel": "left", "side": "right", "range": [0.0, 1.0]},
                       "b": {"panel": "right", "side": "left", "range": [0.0, 1.0]},
                       "reverse_b": True, "notches": [[0.5, 0.5]], "stiffness": 5000.0, "allow_ease": False}]}
