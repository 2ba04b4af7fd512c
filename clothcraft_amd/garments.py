"""Synthetic sewn garments and body proxies for the larger workloads.

Panels are meshed from seeded Poisson-disk samples (grid-phase parallel dart
throwing, every accepted pair at least `spacing` apart, near-maximal) plus
boundary samples, triangulated by Delaunay.  A panel compiles like
GarmentTemplate::compile (src/garment/template.rs): rest coordinates in
metres, per-triangle inverse rest shape and area, lumped masses
(rest area / 3 x density), dihedral hinges in the order Pattern::build forms
them (src/model.rs:193-211).  Seams are penalty stitches between two boundary
curves, sampled at matching arc-length positions: each sample joins a point on
one curve (2 vertices, linear weights) to the matching point on the other
(2 vertices, negated weights), so the weights sum to zero as the reference's
StitchPoint samples do (src/garment/seam.rs:216-265), with coefficient
stiffness x sample length (src/vbd.rs:139-142).
"""
from __future__ import annotations

import math

import numpy as np

from .scenes import Garment, Material, Scene, State, Surface, default_settings


# --------------------------------------------------------------- sampling
def poisson_disk(width: float, height: float, spacing: float, seed: int, rounds: int = 40) -> np.ndarray:
    """Seeded Poisson-disk samples in [0, width] x [0, height], boundary
    included: boundary points every `spacing`, interior points by grid-phase
    dart throwing (cells of spacing/sqrt(2) hold at most one sample; cells of
    one phase are 3 cells apart, so a phase's proposals never conflict with
    each other and are accepted in parallel)."""
    rng = np.random.default_rng(seed)
    c = spacing / math.sqrt(2.0)
    nx, ny = int(math.ceil(width / c)), int(math.ceil(height / c))
    grid = -np.ones((nx + 4, ny + 4), np.int64)  # 2-cell pad on every side
    pts = []

    def cell(p):
        return min(int(p[0] / c), nx - 1) + 2, min(int(p[1] / c), ny - 1) + 2

    def border(n, a, b):
        t = np.linspace(0.0, 1.0, n + 1)
        return [(a[0] + (b[0] - a[0]) * s, a[1] + (b[1] - a[1]) * s) for s in t[:-1]]

    mx, my = max(1, int(round(width / spacing))), max(1, int(round(height / spacing)))
    boundary = (border(mx, (0, 0), (width, 0)) + border(my, (width, 0), (width, height))
                + border(mx, (width, height), (0, height)) + border(my, (0, height), (0, 0)))
    P = np.zeros((0, 2))
    for p in boundary:
        i, j = cell(p)
        # boundary samples may share a cell with a neighbour along the edge
        # only when spacing/sqrt(2) > spacing, which cannot happen
        grid[i, j] = len(pts)
        pts.append(p)
    P = np.array(pts, np.float64)
    offs = [(di, dj) for di in range(-2, 3) for dj in range(-2, 3) if (di, dj) != (0, 0)]
    for _ in range(rounds):
        added = 0
        for ph in range(9):
            pi, pj = divmod(ph, 3)
            ii, jj = np.meshgrid(np.arange(pi, nx, 3), np.arange(pj, ny, 3), indexing="ij")
            ii, jj = ii.ravel() + 2, jj.ravel() + 2
            free = grid[ii, jj] < 0
            ii, jj = ii[free], jj[free]
            if len(ii) == 0:
                continue
            cand = np.stack([(ii - 2 + rng.random(len(ii))) * c, (jj - 2 + rng.random(len(jj))) * c], 1)
            ok = (cand[:, 0] > 0) & (cand[:, 0] < width) & (cand[:, 1] > 0) & (cand[:, 1] < height)
            for di, dj in offs:
                nb = grid[ii + di, jj + dj]
                has = nb >= 0
                d2 = np.full(len(ii), np.inf)
                d2[has] = ((P[nb[has]] - cand[has]) ** 2).sum(1)
                ok &= d2 >= spacing * spacing
            sel = np.nonzero(ok)[0]
            if len(sel):
                base = len(P)
                P = np.vstack([P, cand[sel]])
                grid[ii[sel], jj[sel]] = base + np.arange(len(sel))
                added += len(sel)
        if added == 0:
            break
    return P


def mesh_panel(uv: np.ndarray) -> np.ndarray:
    """Delaunay triangles of the samples, counter-clockwise in uv, without
    degenerate slivers."""
    from scipy.spatial import Delaunay
    tri = Delaunay(uv).simplices.astype(np.int64)
    a, b, c = uv[tri[:, 0]], uv[tri[:, 1]], uv[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    keep = np.abs(cross) > 1e-6 * np.median(np.abs(cross))
    return tri[keep]


# --------------------------------------------------------------- compile
def compile_panels(panels, material_of, density_scale: float = 1.0) -> Garment:
    """Compile panels [(uv [k,2], triangles [t,3], material index)] into one
    garment; returns it with vertex offsets in `garment.panel_offsets`."""
    uv_all, tris, invs, areas, tmat = [], [], [], [], []
    hv, hw, hacross, hmat = [], [], [], []
    offsets = [0]
    for uv, tri, mi in panels:
        base = offsets[-1]
        a, b, c = uv[tri[:, 0]], uv[tri[:, 1]], uv[tri[:, 2]]
        c0, c1 = b - a, c - a
        det = c0[:, 0] * c1[:, 1] - c1[:, 0] * c0[:, 1]
        inv = 1.0 / det
        # glam DMat2::inverse, column-major (x_axis.x, x_axis.y, y_axis.x, y_axis.y)
        invs.append(np.stack([c1[:, 1] * inv, c0[:, 1] * -inv, c1[:, 0] * -inv, c0[:, 0] * inv], 1))
        area = 0.5 * np.abs(det)
        areas.append(area)
        tris.append(tri + base)
        tmat.append(np.full(len(tri), mi, np.uint32))
        # hinges in edge-removal order (src/model.rs:193-211)
        edges: dict = {}
        for t in range(len(tri)):
            a0, b0, c0_ = (int(v) for v in tri[t])
            for (p, q, r) in ((a0, b0, c0_), (b0, c0_, a0), (c0_, a0, b0)):
                key = (min(p, q), max(p, q))
                if key in edges:
                    d, area_d = edges.pop(key)
                    ex, ey = uv[q, 0] - uv[p, 0], uv[q, 1] - uv[p, 1]
                    rl = 1.0 / math.sqrt(ey * ey + ex * ex)
                    hv.append([p + base, q + base, r + base, d + base])
                    hw.append(3.0 * (ex * ex + ey * ey) / (area_d + area[t]))
                    hacross.append((-ey * rl, ex * rl))
                    hmat.append(mi)
                else:
                    edges[key] = (r, area[t])
        uv_all.append(uv)
        offsets.append(base + len(uv))
    uv = np.vstack(uv_all)
    tri = np.vstack(tris).astype(np.uint32)
    area = np.concatenate(areas)
    tm = np.concatenate(tmat)
    rest = np.zeros(len(uv))
    for k in range(3):
        np.add.at(rest, tri[:, k], area / 3.0)
    dens = np.array([material_of(int(m)).density for m in tm])
    masses = np.zeros(len(uv))
    for k in range(3):
        np.add.at(masses, tri[:, k], area / 3.0 * dens * density_scale)
    g = Garment(uv=uv, masses=masses, pinned=np.zeros(len(uv), np.uint8), pin_targets=np.zeros((len(uv), 3)),
                tri_vertices=tri, tri_inv_dm=np.vstack(invs), tri_area=area, tri_material=tm,
                hinge_vertices=np.array(hv, np.uint32).reshape(-1, 4), hinge_weight=np.array(hw, np.float64),
                hinge_across=np.array(hacross, np.float64).reshape(-1, 2),
                hinge_material=np.array(hmat, np.uint32), materials=[])
    g.panel_offsets = np.array(offsets)  # type: ignore[attr-defined]
    return g


def boundary_chain(uv: np.ndarray, side: str, extent, tol: float = 1e-9) -> np.ndarray:
    """Vertex indices on one straight side of a rectangular panel
    ('left' u=0, 'right' u=W, 'top' v=0, 'bottom' v=H), sorted along it and
    restricted to `extent` = (lo, hi) of the running coordinate."""
    W, H = uv[:, 0].max(), uv[:, 1].max()
    if side in ("left", "right"):
        on = np.abs(uv[:, 0] - (0.0 if side == "left" else W)) < tol
        run = uv[:, 1]
    else:
        on = np.abs(uv[:, 1] - (0.0 if side == "top" else H)) < tol
        run = uv[:, 0]
    idx = np.nonzero(on & (run >= extent[0] - tol) & (run <= extent[1] + tol))[0]
    return idx[np.argsort(run[idx])]


def add_seam(g: Garment, chain_a: np.ndarray, chain_b: np.ndarray, stiffness: float, samples_per_segment=2,
             reverse_b: bool = False):
    """Penalty stitches between two boundary chains (global vertex ids),
    matched by normalised arc length in rest coordinates."""
    def arclen(ch):
        p = g.uv[ch]
        s = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(p, axis=0), axis=1))])
        return s / s[-1], s[-1]

    def locate(ch, s_norm, t):
        k = int(np.clip(np.searchsorted(s_norm, t, side="right") - 1, 0, len(ch) - 2))
        w = (t - s_norm[k]) / (s_norm[k + 1] - s_norm[k])
        return (int(ch[k]), int(ch[k + 1])), (1.0 - w, w)

    if reverse_b:
        chain_b = chain_b[::-1]
    sa, la = arclen(chain_a)
    sb, _ = arclen(chain_b)
    n = samples_per_segment * (len(chain_a) - 1)
    off = list(g.stitch_offsets)
    verts, wts, ks = list(g.stitch_vertices), list(g.stitch_weights), list(g.stitch_stiffness)
    for i in range(n):
        t = (i + 0.5) / n
        (a0, a1), (wa0, wa1) = locate(chain_a, sa, t)
        (b0, b1), (wb0, wb1) = locate(chain_b, sb, t)
        verts += [a0, a1, b0, b1]
        wts += [wa0, wa1, -wb0, -wb1]
        ks.append(stiffness * la / n)
        off.append(off[-1] + 4)
    g.stitch_offsets = np.array(off, np.uint32)
    g.stitch_vertices = np.array(verts, np.uint32)
    g.stitch_weights = np.array(wts, np.float64)
    g.stitch_stiffness = np.array(ks, np.float64)
    g.probe_offsets, g.probe_vertices, g.probe_weights = g.stitch_offsets, g.stitch_vertices, g.stitch_weights


# --------------------------------------------------------------- bodies
def capsule(p0, p1, radius: float, around: int, along: int, cap_rings: int) -> Surface:
    """Closed triangulated capsule between centres p0 and p1."""
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    ax = p1 - p0
    L = np.linalg.norm(ax)
    ax /= L
    tmp = np.array([1.0, 0.0, 0.0]) if abs(ax[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(ax, tmp)
    u /= np.linalg.norm(u)
    w = np.cross(ax, u)
    rings = []  # (centre offset along axis, ring radius)
    for k in range(cap_rings, 0, -1):
        a = 0.5 * math.pi * k / (cap_rings + 1)
        rings.append((-radius * math.sin(a), radius * math.cos(a)))
    for k in range(along + 1):
        rings.append((L * k / along, radius))
    for k in range(1, cap_rings + 1):
        a = 0.5 * math.pi * k / (cap_rings + 1)
        rings.append((L + radius * math.sin(a), radius * math.cos(a)))
    th = 2 * math.pi * np.arange(around) / around
    V = [p0 - radius * ax]
    for s, r in rings:
        c = p0 + s * ax
        V += list(c + r * (np.cos(th)[:, None] * u + np.sin(th)[:, None] * w))
    V.append(p1 + radius * ax)
    V = np.array(V)
    T = []
    for j in range(around):  # bottom pole fan (outward normals)
        T.append([0, 1 + (j + 1) % around, 1 + j])
    for r in range(len(rings) - 1):
        b0, b1 = 1 + r * around, 1 + (r + 1) * around
        for j in range(around):
            j1 = (j + 1) % around
            T.append([b0 + j, b0 + j1, b1 + j])
            T.append([b0 + j1, b1 + j1, b1 + j])
    top, last = len(V) - 1, 1 + (len(rings) - 1) * around
    for j in range(around):
        T.append([top, last + j, last + (j + 1) % around])
    return Surface(V, np.array(T, np.uint32))


def merge_surfaces(parts) -> Surface:
    vs, ts, base = [], [], 0
    for s in parts:
        vs.append(s.vertices)
        ts.append(s.triangles.astype(np.int64) + base)
        base += len(s.vertices)
    return Surface(np.vstack(vs), np.vstack(ts).astype(np.uint32))


def mannequin(around: int = 112, along: int = 40, arm_x: float = 0.30) -> Surface:
    """Capsule-union body proxy: torso capsule r 0.15 m (centres y 0.85-1.30 m)
    and two hanging arm capsules r 0.05 m at x = +-arm_x (about 20k
    triangles at the default resolution)."""
    torso = capsule((0.0, 0.85, 0.0), (0.0, 1.30, 0.0), 0.15, around, along, 12)
    arm_l = capsule((-arm_x, 0.95, 0.0), (-arm_x, 1.30, 0.0), 0.05, around // 2, along // 3, 4)
    arm_r = capsule((arm_x, 0.95, 0.0), (arm_x, 1.30, 0.0), 0.05, around // 2, along // 3, 4)
    return merge_surfaces([torso, arm_l, arm_r])


def _half_ellipse(a: float, b: float, m: int = 20001):
    """Arc length s(theta) of x = a cos(theta), z = b sin(theta) on [0, pi]."""
    t = np.linspace(0.0, math.pi, m)
    ds = np.sqrt((a * np.sin(t)) ** 2 + (b * np.cos(t)) ** 2)
    s = np.concatenate([[0.0], np.cumsum(0.5 * (ds[1:] + ds[:-1]) * np.diff(t))])
    return t, s


# --------------------------------------------------------------- workloads
def sewn_garment_on_body(spacing: float = 0.0024, seed: int = 2, width: float = 0.6, height: float = 0.7,
                         seam_stiffness: float = 1e4, iterations: int = 40, radius: float = 0.001,
                         clearance: float = 0.01) -> Scene:
    """Two width x height panels (front and back, seeded Poisson-disk meshes;
    the default spacing gives about 50k vertices each), each wrapped without
    stretch on half of an elliptic cylinder around the mannequin's torso and
    joined by side seams (the full vertical edges) and shoulder seams (the
    outer 0.12 m of the top edges), penalty stitches of seam_stiffness N/m.
    The ellipse's front-back semi-axis is the torso radius plus `clearance`
    (1 cm: the panels start 1 cm outside the body), its side semi-axis makes
    the half perimeter the panel width plus the 4 mm the side seams start
    apart, and the arm capsules hang `clearance` outside the side seams.
    Material: StVK mu 2e4 / lambda 8e3 N/m, bending 2e-5 N m, 0.2 kg/m^2.
    OGC radius 1 mm, stiffness 1e5 N/m (the radius is below the rest
    distance of non-adjacent primitives of the 2.4 mm mesh, so the rest
    state has no in-plane contacts; the reference default of 3 mm is not); dt 1/60, 40
    iterations."""
    mat = Material.stvk(2e4, 8e3, 2e-5, 0.2)
    uv_f = poisson_disk(width, height, spacing, seed)
    uv_b = poisson_disk(width, height, spacing, seed + 1000)
    panels = [(uv_f, mesh_panel(uv_f), 0), (uv_b, mesh_panel(uv_b), 0)]
    g = compile_panels(panels, lambda i: mat)
    g.materials = [mat]
    off = g.panel_offsets
    torso_r, arm_r, gap, top = 0.15, 0.05, 0.004, 1.45
    b = torso_r + clearance
    lo, hi = b, 2.0 * width / math.pi  # half perimeter monotone in the side semi-axis
    for _ in range(60):
        a = 0.5 * (lo + hi)
        lo, hi = (a, hi) if _half_ellipse(a, b)[1][-1] < width + gap else (lo, a)
    t, arc = _half_ellipse(a, b)
    X = np.zeros((g.n, 3))
    for p, sign in ((0, 1.0), (1, -1.0)):
        sl = slice(off[p], off[p + 1])
        u, v = g.uv[sl, 0], g.uv[sl, 1]
        th = sign * np.interp(0.5 * gap + (width - u), arc, t)
        X[sl] = np.stack([a * np.cos(th), top - v, b * np.sin(th)], 1)
    for side in ("left", "right"):
        ca = boundary_chain(g.uv[off[0]:off[1]], side, (0.0, height)) + off[0]
        cb = boundary_chain(g.uv[off[1]:off[2]], side, (0.0, height)) + off[1]
        add_seam(g, ca, cb, seam_stiffness)
    for lo_, hi_ in ((0.0, 0.12), (width - 0.12, width)):
        ca = boundary_chain(g.uv[off[0]:off[1]], "top", (lo_, hi_)) + off[0]
        cb = boundary_chain(g.uv[off[1]:off[2]], "top", (lo_, hi_)) + off[1]
        add_seam(g, ca, cb, seam_stiffness)
    body = mannequin(arm_x=a + clearance + arm_r)
    st = State(X, np.zeros_like(X), 0.0)
    settings = default_settings(dt=1.0 / 60.0, iterations=iterations, radius=radius, stiffness=1e5)
    return Scene(f"sewn_garment_{g.n}v", g, st, body, settings, 240,
                 {"seed": seed, "spacing": spacing, "stitches": len(g.stitch_stiffness),
                  "clearance_m": clearance, "ellipse_semi_axes_m": (a, b)})


def grid_panel(width: float, height: float, nx: int, ny: int) -> tuple:
    """Regular grid rest coordinates and triangles with alternating diagonals
    (the layout of Pattern::build, src/model.rs:164-191, without its 128-cell
    limit)."""
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    uv = np.stack([width * ii.ravel() / nx, height * jj.ravel() / ny], 1)
    tris = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c = a + 1, a + nx + 1
            d = c + 1
            tris += [[a, b, d], [a, d, c]] if (i + j) % 2 == 0 else [[a, b, c], [b, d, c]]
    return uv, np.array(tris, np.int64)


def batched_scenes(count: int = 64, seed: int = 3, v_min: int = 1000, v_max: int = 50000,
                   iterations: int = 20) -> list:
    """Independent instances of uneven sizes: vertex counts log-uniform in
    [v_min, v_max]; each a horizontal sheet 5 cm above an icosphere
    (r 0.25 m) or capsule (r 0.15 m) collider, or a sewn tube (one Poisson-
    disk panel wrapped into a cylinder of radius 0.2 m, closed by a seam that
    starts 4 mm open) around a vertical capsule (r 0.15 m); per-instance StVK
    mu in [5e3, 5e4] N/m (lambda = 0.4 mu), bending in [5e-6, 1e-4] N m and
    density in [0.1, 0.4] kg/m^2 (log-uniform except density).  Shared
    settings: dt 1/60, `iterations` sweeps, OGC radius 2 mm, stiffness 1e5."""
    from .scenes import icosphere
    rng = np.random.default_rng(seed)
    out = []
    settings = default_settings(dt=1.0 / 60.0, iterations=iterations, radius=0.002, stiffness=1e5)
    for k in range(count):
        V = int(round(math.exp(rng.uniform(math.log(v_min), math.log(v_max)))))
        mu = math.exp(rng.uniform(math.log(5e3), math.log(5e4)))
        bend = math.exp(rng.uniform(math.log(5e-6), math.log(1e-4)))
        dens = rng.uniform(0.1, 0.4)
        mat = Material.stvk(mu, 0.4 * mu, bend, dens)
        tube = bool(rng.integers(2))
        if not tube:
            n = max(2, int(round(math.sqrt(V))) - 1)
            uv, tri = grid_panel(1.0, 1.0, n, n)
            g = compile_panels([(uv, tri, 0)], lambda i: mat)
            g.materials = [mat]
            if rng.integers(2):
                body = icosphere(0.25, 4)
                y = 0.25 + 0.05
            else:
                body = capsule((-0.3, 0.0, 0.0), (0.3, 0.0, 0.0), 0.15, 64, 16, 6)
                y = 0.15 + 0.05
            X = np.stack([g.uv[:, 0] - 0.5, np.full(g.n, y), g.uv[:, 1] - 0.5], 1)
            name = f"sheet_{g.n}v"
        else:
            R, H = 0.2, 0.6
            W = 2 * math.pi * R
            spacing = math.sqrt(0.7 * W * H / V)
            uv = poisson_disk(W, H, spacing, int(rng.integers(1 << 30)))
            g = compile_panels([(uv, mesh_panel(uv), 0)], lambda i: mat)
            g.materials = [mat]
            gap = 0.004
            eps = math.asin(0.5 * gap / R)
            th = eps + (2 * math.pi - 2 * eps) * g.uv[:, 0] / W
            X = np.stack([R * np.cos(th), 0.5 - g.uv[:, 1], R * np.sin(th)], 1)
            a = boundary_chain(g.uv, "left", (0.0, H))
            b = boundary_chain(g.uv, "right", (0.0, H))
            add_seam(g, a, b, 1e4)
            body = capsule((0.0, -0.3, 0.0), (0.0, 0.4, 0.0), 0.15, 64, 24, 6)
            name = f"tube_{g.n}v"
        out.append(Scene(name, g, State(X, np.zeros_like(X), 0.0), body, dict(settings), 120,
                         {"mu": mu, "bend": bend, "density": dens}))
    return out
