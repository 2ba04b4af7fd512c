"""Reference-format garment templates (GarmentTemplate::compile,
src/garment/template.rs:214-451, seams src/garment/seam.rs:80-407) for grid
panels.  Fixtures: the reference's own example templates
examples/templates/two-panel.json and local-ease.json (data files, copied
under tests/golden/).  Pinned by what the mathematics fixes: counts of the
grid meshes and of the common seam partition, sample weights that sum to the
mean seam length, zero gap where two nonmatching boundaries are placed on the
same segment, the two-point Gauss rule integrating a linear gap field's
squared norm exactly, and a rigid offset opening every probe by its length."""
import copy
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from clothcraft_amd import templates as T

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_two_panel_demo_equals_the_reference_example_file():
    demo, ref = T.two_panel_demo(), _load("two-panel.json")
    for key in ("format", "version", "name", "fabrics", "seams"):
        assert demo[key] == ref[key], key
    for p, q in zip(demo["panels"], ref["panels"]):
        for key in ("id", "name", "pattern", "fabric", "grain_degrees", "placement", "pins"):
            assert p[key] == q[key], key


def test_two_panel_counts_pins_and_partition():
    c = T.compile_template(_load("two-panel.json"))
    g = c.garment
    assert g.n == 4 * 5 + 5 * 8 and len(g.tri_area) == 2 * 12 + 2 * 28
    assert [r[1] for r in c.panels] == [(0, 20), (20, 60)]
    assert list(np.nonzero(g.pinned)[0]) == [0, 3]  # left panel's top corners
    assert np.allclose(g.pin_targets[[0, 3]], c.state.x[[0, 3]])
    assert set(g.tri_material[:24]) == {0} and set(g.tri_material[24:]) == {1}
    assert [m.grain_degrees for m in g.materials] == [0.0, 45.0]
    seam = c.seams[0]
    # partition: quarters of side a, sevenths of side b and the notch at 1/2
    assert len(seam.samples) == 2 * 10 and len(seam.probes) == 11
    assert seam.lengths == pytest.approx((0.6, 0.6))
    assert sum(w for _, w in seam.samples) == pytest.approx(0.6, rel=1e-14)  # (La + Lb) / 2
    for terms in [t for t, _ in seam.samples] + seam.probes:
        assert abs(sum(cf for _, cf in terms)) < 1e-14
    assert not c.warnings
    # placement: panel 2 rotated by Rx(8) Ry(12) about its origin, then translated
    R = T.euler_xyz((8.0, 12.0, 0.0))
    uv = g.uv[20:]
    want = np.stack([uv[:, 0], -uv[:, 1], 0 * uv[:, 0]], 1) @ R.T + [0.08, 1.0, 0.03]
    assert np.abs(c.state.x[20:] - want).max() < 1e-15


def _coincident(c):
    """Both panels in the plane z = 0 with the left panel's right side and the
    right panel's left side on the same segment x = 0.4, -0.6 <= y <= 0."""
    g = c.garment
    x = np.zeros((g.n, 3))
    x[:20, 0], x[:20, 1] = g.uv[:20, 0], -g.uv[:20, 1]
    x[20:, 0], x[20:, 1] = 0.4 + g.uv[20:, 0], -g.uv[20:, 1]
    return x


def _gap(terms, x):
    return sum(cf * x[v] for v, cf in terms)


def test_nonmatching_seam_gap_vanishes_on_a_common_segment_and_integrates_exactly():
    c = T.compile_template(_load("two-panel.json"))
    seam = c.seams[0]
    x = _coincident(c)
    assert max(np.linalg.norm(_gap(t, x)) for t in seam.probes) < 1e-15
    assert O.max_seam_gap(c.garment, x) < 1e-15
    # a gap field linear in the arclength s of side a (top to bottom):
    # b displaced by -(d0 + d1 s); the stitch energy 1/2 k sum w |gap|^2 is
    # 1/2 k L (|d0|^2 + d0.d1 + |d1|^2 / 3) for the two-point Gauss rule.
    d0, d1 = np.array([0.01, -0.02, 0.005]), np.array([-0.03, 0.01, 0.02])
    y = x.copy()
    s = -x[20:, 1] / 0.6
    y[20:] -= d0 + np.outer(s, d1)
    e = sum(0.5 * seam.stiffness * w * float(np.dot(_gap(t, y), _gap(t, y))) for t, w in seam.samples)
    want = 0.5 * seam.stiffness * 0.6 * (d0 @ d0 + d0 @ d1 + d1 @ d1 / 3.0)
    assert e == pytest.approx(want, rel=1e-12)
    # the same energy through the oracle's Newton objective (its seam term)
    g0 = copy.copy(c.garment)
    for k in ("stitch_offsets", "probe_offsets"):
        setattr(g0, k, np.zeros(1, np.uint32))
    for k in ("stitch_vertices", "probe_vertices"):
        setattr(g0, k, np.zeros(0, np.uint32))
    for k in ("stitch_weights", "stitch_stiffness", "probe_weights"):
        setattr(g0, k, np.zeros(0))
    old = T.State(y.copy(), np.zeros_like(y), 0.0)
    ns = O.newton_settings(gravity=(0.0, 0.0, 0.0))
    seam_e = O.newton_objective(c.garment, old, y, ns) - O.newton_objective(g0, old, y, ns)
    assert seam_e == pytest.approx(want, rel=1e-9)
    # a rigid offset of the right panel opens every probe by its length
    z = x.copy()
    z[20:] += [0.002, -0.001, 0.004]
    assert O.max_seam_gap(c.garment, z) == pytest.approx(math.sqrt(0.002**2 + 0.001**2 + 0.004**2), rel=1e-12)


def test_local_ease_template_warns_and_requires_allow_ease():
    t = _load("local-ease.json")
    c = T.compile_template(t)
    assert len(c.warnings) == 1 and "unequal total or notch-interval lengths" in c.warnings[0]
    t["seams"][0]["allow_ease"] = False
    with pytest.raises(T.TemplateError, match="notches introduce local easing; set allow_ease explicitly"):
        T.compile_template(t)


@pytest.mark.parametrize("edit,message", [
    (lambda t: t.update(version=3), "Unsupported garment template format/version"),
    (lambda t: t["panels"][0].update(fabric="silk"), "Panel left references missing fabric silk"),
    (lambda t: t["panels"][0].update(pins=["top_left", "top_left"]), "Duplicate pin in left"),
    (lambda t: t["panels"][1].update(id="left"), "Invalid or duplicate panel id: left"),
    (lambda t: t["panels"][0].update(polygon={"points": []}),
     "Polygon geometry and material pins require template version 2"),
    (lambda t: t["panels"][0]["pattern"].update(nx=0),
     "Pattern requires dimensions in 0.01..100 m and 1..128 cells per axis"),
    (lambda t: t["seams"][0].update(stiffness=-1.0), "Seam stiffness must be positive and finite"),
    (lambda t: t["seams"].append(dict(t["seams"][0], id="second")), "Overlapping seam ownership on left Right"),
])
def test_template_errors_carry_the_reference_messages(edit, message):
    t = _load("two-panel.json")
    edit(t)
    with pytest.raises(T.TemplateError, match=message):
        T.compile_template(t)
