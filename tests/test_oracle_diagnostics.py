"""Pins of the oracle's garment diagnostics and seam terms by what the
mathematics fixes (no reference vectors exist for them):

* extension_percent (src/garment/diagnostics.rs:7-28 with energy::stretches,
  src/energy.rs:80-114): homogeneous stretches have closed-form fibre
  stretches, and the area-weighted 95th percentile of a two-triangle garment is
  decided by which side of 5 % the stretched triangle's area falls;
* max_seam_gap (src/garment/template.rs:536-538, seam.rs:418-423): a rigid
  offset d of one panel opens every seam probe by exactly |d|, a translation
  of the whole garment or a rotation about the seam line opens none;
* the stitch energy (seam.rs:408-413 in the Newton objective, src/vbd.rs:217-227
  in the VBD local energy): 1/2 sum k |d|^2 under the rigid offset, translation
  invariance (the sample weights sum to zero), and the VBD local gradient of
  the stitches equals the central finite difference of the Newton objective's
  seam energy (two independent code paths of the oracle).
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from clothcraft_amd import garments, scenes  # noqa: E402


def _sheet(grain=0.0, nx=6, ny=5):
    m = scenes.Material.stvk(2e4, 8e3, 2e-5, 0.2, grain=grain)
    g, s = scenes.pattern_build(0.6, 0.5, nx, ny, m, "sphere")
    return g, s


def _deform(s, fu, fv):
    """Homogeneous deformation of the horizontal sheet: material u is world x,
    material v is world z (Pattern::build's 'sphere' placement)."""
    x = s.x.copy()
    c = x.mean(axis=0)
    x[:, 0] = c[0] + fu * (x[:, 0] - c[0])
    x[:, 2] = c[2] + fv * (x[:, 2] - c[2])
    return x


@pytest.mark.parametrize("s", [1.0, 1.07, 1.25])
def test_extension_uniform_stretch(s):
    g, st = _sheet()
    p95, mx = O.extension_percent(g, _deform(st, s, s))
    assert p95 == pytest.approx(100 * (s - 1), abs=1e-9)
    assert mx == pytest.approx(100 * (s - 1), abs=1e-9)


def test_extension_ignores_compression_and_rigid_motion():
    g, st = _sheet()
    assert O.extension_percent(g, _deform(st, 0.9, 0.8)) == pytest.approx([0.0, 0.0], abs=1e-12)
    R = scenes.euler_xyz((30.0, -50.0, 12.0))
    x = st.x @ R.T + np.array([0.3, -1.0, 2.0])
    assert O.extension_percent(g, x) == pytest.approx([0.0, 0.0], abs=1e-9)


@pytest.mark.parametrize("grain", [0.0, 90.0, 45.0, 30.0])
def test_extension_uniaxial_stretch_along_rotated_grain(grain):
    """F = diag(s, 1) in material coordinates: a fibre at angle a to u
    stretches by sqrt(s^2 cos^2 a + sin^2 a); warp lies at the grain angle
    and weft at grain + 90 degrees."""
    s = 1.2
    g, st = _sheet(grain)
    a = math.radians(grain)
    warp = math.sqrt(s * s * math.cos(a) ** 2 + math.sin(a) ** 2)
    weft = math.sqrt(s * s * math.sin(a) ** 2 + math.cos(a) ** 2)
    want = 100 * (max(warp, weft) - 1)
    p95, mx = O.extension_percent(g, _deform(st, s, 1.0))
    assert p95 == pytest.approx(want, abs=1e-9) and mx == pytest.approx(want, abs=1e-9)


@pytest.mark.parametrize("small_fraction,expect_small", [(0.04, False), (0.06, True)])
def test_extension_area_weighted_percentile(small_fraction, expect_small):
    """Two separate triangles; the smaller one is stretched more.  Sorted by
    extension, the big triangle's area reaches 95 % on its own exactly when
    the small one holds less than 5 % of the area."""
    big = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    k = math.sqrt(small_fraction / (1 - small_fraction))
    small = np.array([[0.0, 0.0], [k, 0.0], [0.0, k]])
    mat = scenes.Material.stvk(2e4, 8e3, 2e-5, 0.2)
    g = garments.compile_panels([(big, np.array([[0, 1, 2]]), 0), (small, np.array([[0, 1, 2]]), 0)],
                                lambda i: mat)
    g.materials = [mat]
    assert g.tri_area[1] / g.tri_area.sum() == pytest.approx(small_fraction)
    s_big, s_small = 1.03, 1.11
    x = np.zeros((6, 3))
    x[:3, 0], x[:3, 2] = s_big * big[:, 0], s_big * big[:, 1]
    x[3:, 0], x[3:, 2] = 5 + s_small * small[:, 0], s_small * small[:, 1]
    p95, mx = O.extension_percent(g, x)
    assert mx == pytest.approx(100 * (s_small - 1), abs=1e-9)
    assert p95 == pytest.approx(100 * ((s_small if expect_small else s_big) - 1), abs=1e-9)


def _seamed_pair(nx=4, ny=6, w=0.3, h=0.4, stiffness=1e4):
    """Two coplanar panels whose seam edges coincide: A spans x in [-w, 0], B
    spans [0, w]; A's right side is sewn to B's left side along z."""
    mat = scenes.Material.stvk(2e4, 8e3, 2e-5, 0.2)
    uv, tri = garments.grid_panel(w, h, nx, ny)
    g = garments.compile_panels([(uv, tri, 0), (uv, tri, 0)], lambda i: mat)
    g.materials = [mat]
    off = g.panel_offsets
    a = garments.boundary_chain(g.uv[off[0]:off[1]], "right", (0.0, h)) + off[0]
    b = garments.boundary_chain(g.uv[off[1]:off[2]], "left", (0.0, h)) + off[1]
    garments.add_seam(g, a, b, stiffness)
    x = np.zeros((g.n, 3))
    na = off[1]
    x[:na, 0] = g.uv[:na, 0] - w
    x[na:, 0] = g.uv[na:, 0]
    x[:, 2] = g.uv[:, 1]
    x[:, 1] = 1.0
    return g, x, na


def _no_seams(g):
    import copy
    h = copy.copy(g)
    h.stitch_offsets = np.zeros(1, np.uint32)
    h.stitch_vertices = np.zeros(0, np.uint32)
    h.stitch_weights = np.zeros(0)
    h.stitch_stiffness = np.zeros(0)
    h.probe_offsets = np.zeros(1, np.uint32)
    h.probe_vertices = np.zeros(0, np.uint32)
    h.probe_weights = np.zeros(0)
    return h


def _seam_energy(g, x):
    """The Newton objective's seam energy, isolated: the objective of the
    garment minus that of the same garment without stitches, from a state at
    rest at x (no gravity: the inertial term vanishes)."""
    old = scenes.State(x.copy(), np.zeros_like(x), 0.0)
    s = O.newton_settings(gravity=(0.0, 0.0, 0.0))
    return O.newton_objective(g, old, x, s) - O.newton_objective(_no_seams(g), old, x, s)


def test_seam_gap_and_energy_under_rigid_offset():
    g, x, na = _seamed_pair()
    assert O.max_seam_gap(g, x) < 1e-15
    assert abs(_seam_energy(g, x)) < 1e-20
    d = np.array([0.003, -0.004, 0.012])
    y = x.copy()
    y[na:] += d
    assert O.max_seam_gap(g, y) == pytest.approx(np.linalg.norm(d), rel=1e-12)
    assert _seam_energy(g, y) == pytest.approx(0.5 * d @ d * g.stitch_stiffness.sum(), rel=1e-9)


def test_seam_gap_and_energy_invariances():
    g, x, na = _seamed_pair()
    rng = np.random.default_rng(5)
    y = x + rng.uniform(-0.01, 0.01, x.shape)
    e, gap = _seam_energy(g, y), O.max_seam_gap(g, y)
    assert gap > 1e-3
    t = np.array([0.7, -0.2, 1.3])
    assert _seam_energy(g, y + t) == pytest.approx(e, rel=1e-9)
    assert O.max_seam_gap(g, y + t) == pytest.approx(gap, rel=1e-9)
    # Panel B folded about the seam line (the z axis through x = 0, y = 1):
    # the seam points lie on the axis and do not move.
    th = 0.7
    z = x.copy()
    px, py = z[na:, 0], z[na:, 1] - 1.0
    z[na:, 0], z[na:, 1] = px * math.cos(th) - py * math.sin(th), 1.0 + px * math.sin(th) + py * math.cos(th)
    assert O.max_seam_gap(g, z) < 1e-15


def test_vbd_local_stitch_gradient_is_the_seam_energy_derivative():
    """VbdSolver::local's stitch gradient (local gradient with seams minus
    without) equals central differences of the Newton objective's seam energy."""
    g, x, na = _seamed_pair()
    rng = np.random.default_rng(7)
    y = x + rng.uniform(-0.004, 0.004, x.shape)
    sep = x.copy()
    sep[na:, 1] += 0.05  # the start must pass OGC validation (no touching panels)
    st = scenes.State(sep, np.zeros_like(x), 0.0)
    s = scenes.default_settings(dt=1.0 / 60.0)
    inertia = np.ones(g.n)
    gs = O.Solver(g, None, st).local_gradient(y, y, inertia, s)[:, :3]
    g0 = O.Solver(_no_seams(g), None, st).local_gradient(y, y, inertia, s)[:, :3]
    stitch = gs - g0
    seam_vertices = np.unique(g.stitch_vertices)
    h = 1e-6
    for v in seam_vertices[::3]:
        for k in range(3):
            yp, ym = y.copy(), y.copy()
            yp[v, k] += h
            ym[v, k] -= h
            fd = (_seam_energy(g, yp) - _seam_energy(g, ym)) / (2 * h)
            assert stitch[v, k] == pytest.approx(fd, rel=2e-5, abs=1e-9)
    others = np.setdiff1d(np.arange(g.n), seam_vertices)
    assert np.abs(stitch[others]).max() == 0.0
