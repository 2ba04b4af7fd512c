"""Garment/state containers and deterministic scene generators.

`pattern_build` restates `Pattern::build` (src/model.rs:124-227) exactly — the
grid, alternating diagonals, BTreeMap edge-removal hinge order, lumped masses
and hanging pins — because term order fixes the f64 summation order of the
solver.  `place` restates the panel placement of `GarmentTemplate::compile`
(src/garment/template.rs:79-92,320-330).  The remaining generators (icosphere,
capsule-union mannequin, Poisson-disk panels, seams) build the benchmark
scenes; they have no reference counterpart.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


# --------------------------------------------------------------- data types
@dataclass
class Material:
    """src/model.rs:7-38 (defaults are Material::default())."""
    warp: float = 350.0
    weft: float = 180.0
    shear: float = 45.0
    compression_ratio: float = 0.15
    compression_transition: float = 0.005
    bend_warp: float = 0.00008
    bend_weft: float = 0.00004
    density: float = 0.2
    thickness: float = 0.001
    grain_degrees: float = 0.0

    def validate(self) -> None:
        """Material::validate (src/model.rs:40-63)."""
        for name in ("warp", "weft", "shear", "density", "thickness", "compression_transition"):
            v = getattr(self, name)
            if not math.isfinite(v) or v <= 0.0:
                raise ValueError(f"{name} must be finite and positive")
        if (not math.isfinite(self.grain_degrees) or not math.isfinite(self.compression_ratio)
                or not 0.0 <= self.compression_ratio <= 1.0
                or any(not math.isfinite(v) or v < 0.0 for v in (self.bend_warp, self.bend_weft))):
            raise ValueError("Invalid bending, compression or grain parameter")

    @staticmethod
    def stvk(mu: float, lam: float, bend: float, density: float, grain: float = 0.0) -> "Material":
        """An "StVK mu, lambda" material in the reference law: compression_ratio 1,
        warp = weft = 2 mu + lambda, shear = mu (isotropic StVK minus the
        lambda*E11*E22 Poisson coupling, which the orthotropic law has no term for)."""
        return Material(warp=2 * mu + lam, weft=2 * mu + lam, shear=mu, compression_ratio=1.0,
                        compression_transition=0.005, bend_warp=bend, bend_weft=bend,
                        density=density, thickness=0.001, grain_degrees=grain)

    def to_c(self) -> abi.cv_material:
        return abi.cv_material(self.warp, self.weft, self.shear, self.compression_ratio,
                               self.compression_transition, self.bend_warp, self.bend_weft,
                               self.density, self.thickness, self.grain_degrees)


@dataclass
class Garment:
    """Flattened CompiledGarment (src/garment/template.rs:170-180)."""
    uv: np.ndarray                      # [n,2] rest coordinates
    masses: np.ndarray                  # [n]
    pinned: np.ndarray                  # [n] uint8
    pin_targets: np.ndarray             # [n,3]
    tri_vertices: np.ndarray            # [T,3] uint32
    tri_inv_dm: np.ndarray              # [T,4]
    tri_area: np.ndarray                # [T]
    tri_material: np.ndarray            # [T] uint32
    hinge_vertices: np.ndarray          # [H,4] uint32
    hinge_weight: np.ndarray            # [H]
    hinge_across: np.ndarray            # [H,2]
    hinge_material: np.ndarray          # [H] uint32
    materials: list                     # [Material]
    stitch_offsets: np.ndarray = field(default_factory=lambda: np.zeros(1, np.uint32))
    stitch_vertices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    stitch_weights: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    stitch_stiffness: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    probe_offsets: np.ndarray = field(default_factory=lambda: np.zeros(1, np.uint32))
    probe_vertices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    probe_weights: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))

    @property
    def n(self) -> int:
        return int(self.masses.shape[0])

    def to_c(self) -> abi.cv_garment:
        """cv_garment view; the returned struct keeps the arrays alive."""
        arrs = {}

        def keep(name, a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            arrs[name] = a
            return a

        g = abi.cv_garment()
        g.n_vertices = self.n
        g.masses = abi.ptr(keep("m", self.masses, np.float64), C.c_double)
        g.pinned = abi.ptr(keep("p", self.pinned, np.uint8), C.c_uint8)
        g.pin_targets = abi.ptr(keep("pt", self.pin_targets, np.float64), C.c_double)
        g.n_triangles = len(self.tri_area)
        g.tri_vertices = abi.ptr(keep("tv", self.tri_vertices, np.uint32), C.c_uint32)
        g.tri_inv_dm = abi.ptr(keep("ti", self.tri_inv_dm, np.float64), C.c_double)
        g.tri_area = abi.ptr(keep("ta", self.tri_area, np.float64), C.c_double)
        g.tri_material = abi.ptr(keep("tm", self.tri_material, np.uint32), C.c_uint32)
        g.n_hinges = len(self.hinge_weight)
        g.hinge_vertices = abi.ptr(keep("hv", self.hinge_vertices, np.uint32), C.c_uint32)
        g.hinge_weight = abi.ptr(keep("hw", self.hinge_weight, np.float64), C.c_double)
        g.hinge_across = abi.ptr(keep("ha", self.hinge_across, np.float64), C.c_double)
        g.hinge_material = abi.ptr(keep("hm", self.hinge_material, np.uint32), C.c_uint32)
        mats = (abi.cv_material * max(1, len(self.materials)))(*[m.to_c() for m in self.materials])
        arrs["mats"] = mats
        g.n_materials = len(self.materials)
        g.materials = C.cast(mats, C.POINTER(abi.cv_material))
        g.n_stitches = len(self.stitch_stiffness)
        g.stitch_offsets = abi.ptr(keep("so", self.stitch_offsets, np.uint32), C.c_uint32)
        g.stitch_vertices = abi.ptr(keep("sv", self.stitch_vertices, np.uint32), C.c_uint32)
        g.stitch_weights = abi.ptr(keep("sw", self.stitch_weights, np.float64), C.c_double)
        g.stitch_stiffness = abi.ptr(keep("ss", self.stitch_stiffness, np.float64), C.c_double)
        g.n_probes = len(self.probe_offsets) - 1
        g.probe_offsets = abi.ptr(keep("po", self.probe_offsets, np.uint32), C.c_uint32)
        g.probe_vertices = abi.ptr(keep("pv", self.probe_vertices, np.uint32), C.c_uint32)
        g.probe_weights = abi.ptr(keep("pw", self.probe_weights, np.float64), C.c_double)
        g._keep = arrs  # type: ignore[attr-defined]
        return g


@dataclass
class Surface:
    """SurfaceMesh (src/avatar.rs:9-12)."""
    vertices: np.ndarray   # [V,3]
    triangles: np.ndarray  # [F,3] uint32

    def to_c(self) -> abi.cv_surface:
        v = np.ascontiguousarray(self.vertices, np.float64)
        t = np.ascontiguousarray(self.triangles, np.uint32)
        s = abi.cv_surface(len(v), abi.ptr(v, C.c_double), len(t), abi.ptr(t, C.c_uint32))
        s._keep = (v, t)  # type: ignore[attr-defined]
        return s


@dataclass
class State:
    """State (src/model.rs:116-121); x, v are [n,3] f64."""
    x: np.ndarray
    v: np.ndarray
    time: float = 0.0

    def copy(self) -> "State":
        return State(self.x.copy(), self.v.copy(), self.time)

    def to_c(self) -> abi.cv_state:
        assert self.x.dtype == np.float64 and self.x.flags["C_CONTIGUOUS"]
        assert self.v.dtype == np.float64 and self.v.flags["C_CONTIGUOUS"]
        return abi.cv_state(abi.ptr(self.x, C.c_double), abi.ptr(self.v, C.c_double),
                            self.x.shape[0], self.time)


@dataclass
class Scene:
    name: str
    garment: Garment
    state: State
    avatar: Surface | None
    settings: dict
    steps: int
    meta: dict = field(default_factory=dict)


# ------------------------------------------------------------ Pattern::build
def _inv2(c0, c1):
    """glam DMat2::inverse on columns c0, c1 -> flattened (x_axis.x, x_axis.y, y_axis.x, y_axis.y)."""
    det = c0[0] * c1[1] - c1[0] * c0[1]
    inv = 1.0 / det
    return (c1[1] * inv, c0[1] * -inv, c1[0] * -inv, c0[0] * inv), det


def pattern_build(width: float, height: float, nx: int, ny: int, material: Material,
                  scene: str = "hanging"):
    """Pattern::build (src/model.rs:124-227).  Returns (Garment, State)."""
    material.validate()
    if (not (0.01 <= width <= 100.0) or not (0.01 <= height <= 100.0)
            or nx == 0 or ny == 0 or nx > 128 or ny > 128):
        raise ValueError("Pattern requires dimensions in 0.01..100 m and 1..128 cells per axis")
    n = (nx + 1) * (ny + 1)
    uv = []
    x = []
    for j in range(ny + 1):
        for i in range(nx + 1):
            p = ((width * i) / nx, (height * j) / ny)
            uv.append(p)
            if scene == "hanging":
                x.append((p[0] - width * 0.5, 1.3 - p[1], 0.0))
            else:
                x.append((p[0] - width * 0.5, 0.65, p[1] - height * 0.5))
    rest_area = [0.0] * n
    tris, invs, areas = [], [], []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b = a + 1
            c = a + nx + 1
            d = c + 1
            faces = [[a, b, d], [a, d, c]] if (i + j) % 2 == 0 else [[a, b, c], [b, d, c]]
            for vs in faces:
                r = [uv[v] for v in vs]
                c0 = (r[1][0] - r[0][0], r[1][1] - r[0][1])
                c1 = (r[2][0] - r[0][0], r[2][1] - r[0][1])
                inv, det = _inv2(c0, c1)
                area = 0.5 * abs(det)
                for v in vs:
                    rest_area[v] += area / 3.0
                tris.append(vs)
                invs.append(inv)
                areas.append(area)
    # Hinges in BTreeMap edge-removal order (src/model.rs:193-211).
    edges: dict = {}
    hv, hw, hacross = [], [], []
    for t, vs in enumerate(tris):
        a0, b0, c0 = vs
        for (a, b, c) in ((a0, b0, c0), (b0, c0, a0), (c0, a0, b0)):
            key = (min(a, b), max(a, b))
            if key in edges:
                d, area = edges.pop(key)
                ex = uv[b][0] - uv[a][0]
                ey = uv[b][1] - uv[a][1]
                length = math.sqrt((-ey) * (-ey) + ex * ex)
                rl = 1.0 / length
                hv.append([a, b, c, d])
                hw.append(3.0 * (ex * ex + ey * ey) / (area + areas[t]))
                hacross.append((-ey * rl, ex * rl))
            else:
                edges[key] = (c, areas[t])
    masses = np.array([ra * material.density for ra in rest_area], np.float64)
    pinned = np.zeros(n, np.uint8)
    pin_targets = np.zeros((n, 3), np.float64)
    X = np.array(x, np.float64)
    if scene == "hanging":
        for i in (0, nx):
            pinned[i] = 1
            pin_targets[i] = X[i]
    g = Garment(
        uv=np.array(uv, np.float64), masses=masses, pinned=pinned, pin_targets=pin_targets,
        tri_vertices=np.array(tris, np.uint32).reshape(-1, 3),
        tri_inv_dm=np.array(invs, np.float64).reshape(-1, 4),
        tri_area=np.array(areas, np.float64),
        tri_material=np.zeros(len(tris), np.uint32),
        hinge_vertices=np.array(hv, np.uint32).reshape(-1, 4),
        hinge_weight=np.array(hw, np.float64),
        hinge_across=np.array(hacross, np.float64).reshape(-1, 2),
        hinge_material=np.zeros(len(hv), np.uint32),
        materials=[material])
    return g, State(X, np.zeros_like(X), 0.0)


def euler_xyz(deg) -> np.ndarray:
    """R = Rx * Ry * Rz (glam EulerRot::XYZ, src/garment/template.rs:79-92)."""
    ax, ay, az = (math.radians(d) for d in deg)
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rx @ ry @ rz


def place(g: Garment, translation, rotation_degrees=(0.0, 0.0, 0.0), pins=()) -> State:
    """Panel placement of GarmentTemplate::compile: x = R (u, -v, 0) + t, with
    the listed vertex indices pinned at their placed positions."""
    local = np.stack([g.uv[:, 0], -g.uv[:, 1], np.zeros(g.n)], axis=1)
    X = local @ euler_xyz(rotation_degrees).T + np.asarray(translation, np.float64)
    g.pinned[:] = 0
    g.pin_targets[:] = 0
    for i in pins:
        g.pinned[i] = 1
        g.pin_targets[i] = X[i]
    return State(np.ascontiguousarray(X), np.zeros_like(X), 0.0)


# ---------------------------------------------------------------- colliders
def icosphere(radius: float, subdivisions: int, center=(0.0, 0.0, 0.0)) -> Surface:
    """Outward-oriented icosphere: 20*4^s triangles (s=4: 5,120 tris, 2,562 vertices)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache: dict = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, np.float64)
    F = np.array(faces, np.uint32)
    # Orient outward.
    cen = np.asarray(center, np.float64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), (a + b + c) / 3 - cen) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return Surface(V, F)


# ------------------------------------------------------------- workloads
def default_settings(**kw) -> dict:
    """VbdSettings::default() (src/vbd.rs:26-36) with OgcSettings::default()."""
    s = dict(dt=1.0 / 240.0, iterations=16, damping=1.5, gravity=(0.0, -9.81, 0.0),
             radius=0.003, stiffness=100000.0, relaxation=0.45, refresh_fraction=0.01,
             residual_tolerance=0.0, flags=0)
    s.update(kw)
    return s


def hanging_sheet(seed: int = 0) -> Scene:
    """Plain VBD sheet: 1.0 x 1.0 m, 32x32 quads (1,089 vertices, 2,048 triangles),
    hanging from the two top corners, mu 2e4, lambda 8e3, bending 2e-5, density 0.2,
    dt 1/60, 10 iterations, 120 steps, no collider.  Unpinned vertices jittered by
    a seeded uniform +/-1e-4 m in every coordinate."""
    m = Material.stvk(2e4, 8e3, 2e-5, 0.2)
    g, s = pattern_build(1.0, 1.0, 32, 32, m, "hanging")
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-1e-4, 1e-4, size=s.x.shape)
    jit[g.pinned.astype(bool)] = 0.0
    s.x = np.ascontiguousarray(s.x + jit)
    return Scene("hanging_sheet_32x32", g, s, None,
                 default_settings(dt=1.0 / 60.0, iterations=10), 120, {"seed": seed})


def sheet_on_sphere(seed: int = 1, nx: int = 128) -> Scene:
    """Sheet on sphere: nx x nx quads over 1 m^2 (128: 16,641 vertices, 32,768
    triangles), horizontal, 0.3 m above the top of a static icosphere collider
    (radius 0.25 m at the origin, subdivision 4, 5,120 triangles).  Per-triangle
    mu scaled by a seeded uniform factor in [0.8, 1.2].  OGC radius 2 mm,
    stiffness 1e5 N/m, 20 iterations, dt 1/60, 300 steps."""
    base = Material.stvk(2e4, 8e3, 2e-5, 0.2)
    g, s = pattern_build(1.0, 1.0, nx, nx, base, "sphere")
    rng = np.random.default_rng(seed)
    factor = rng.uniform(0.8, 1.2, size=len(g.tri_area))
    g.materials = [Material.stvk(2e4 * f, 8e3, 2e-5, 0.2) for f in factor]
    g.tri_material = np.arange(len(g.tri_area), dtype=np.uint32)
    g.hinge_material = np.zeros(len(g.hinge_weight), np.uint32)
    s.x[:, 1] = 0.25 + 0.3
    sphere = icosphere(0.25, 4)
    return Scene(f"sheet_on_sphere_{nx}x{nx}", g, s, sphere,
                 default_settings(dt=1.0 / 60.0, iterations=20, radius=0.002, stiffness=1e5),
                 300, {"seed": seed, "sphere_center": (0.0, 0.0, 0.0), "sphere_radius": 0.25})


def small_sheet_on_sphere(nx: int = 16, height: float = 0.255, seed: int = 3) -> Scene:
    """The sheet-on-sphere scene at nx x nx quads, placed at the given height."""
    sc = sheet_on_sphere(seed=seed, nx=nx)
    sc.state.x[:, 1] = height
    sc.name = f"sheet_on_sphere_{nx}x{nx}"
    return sc


# ------------------------------------------------------ reference fixtures
def demo_fabric() -> Material:
    """FabricProperties::default() bound at grain 0 (src/garment/template.rs:31-35)."""
    return Material()


def free_panel(nx: int = 2, ny: int = 2) -> Scene:
    """tests/vbd.rs:71-79 fixture: the left panel of two_panel_demo
    (src/garment/template.rs:453-533, 0.4 x 0.6 m at (-0.4, 1, 0)) with nx x ny
    cells, seams and pins cleared."""
    g, _ = pattern_build(0.4, 0.6, nx, ny, demo_fabric(), "sphere")
    s = place(g, (-0.4, 1.0, 0.0))
    return Scene(f"free_panel_{nx}x{ny}", g, s, None, default_settings(), 1)


def resting_patch(offset=(0.0, 0.0, 0.0)) -> Scene:
    """tests/common/mod.rs:8-36: a 0.12 m 3x3 patch rotated 90 deg about x,
    50 micrometres outside a 3 mm shell above a 2 m floor triangle."""
    g, _ = pattern_build(0.12, 0.12, 3, 3, demo_fabric(), "sphere")
    off = np.asarray(offset, np.float64)
    s = place(g, off + np.array([-0.06, 0.00305, 0.06]), (90.0, 0.0, 0.0))
    floor = Surface(np.array([[-1.0, 0.0, -1.0], [1.0, 0.0, -1.0], [0.0, 0.0, 1.0]]) + off,
                    np.array([[0, 1, 2]], np.uint32))
    return Scene("resting_patch", g, s, floor, default_settings(dt=1.0 / 60.0), 60)
