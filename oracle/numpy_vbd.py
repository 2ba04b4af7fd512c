"""ORACLE — test infrastructure only.

A second, independent restatement of `VbdSolver::step` (src/vbd.rs:250-422)
in Python/numpy, written apart from the C++ oracle and differentiating with
hyper-dual numbers instead of second-order jets.  It covers contact-free
steps on meshes whose features are all farther apart than the query radius
(so every trust-region ball has radius relaxation * rq; the constructor
asserts this); agreement with liboracle.so to ~1e-12 pins the transcription
of the step loop.  Small meshes only (pure Python per vertex)."""
from __future__ import annotations

import math

import numpy as np


class HD:
    """Hyper-dual number a + b1 e1 + b2 e2 + b12 e1e2 (e1^2 = e2^2 = 0)."""
    __slots__ = ("a", "b1", "b2", "b12")

    def __init__(self, a, b1=0.0, b2=0.0, b12=0.0):
        self.a, self.b1, self.b2, self.b12 = a, b1, b2, b12

    @staticmethod
    def lift(x):
        return x if isinstance(x, HD) else HD(float(x))

    def __add__(self, o):
        o = HD.lift(o)
        return HD(self.a + o.a, self.b1 + o.b1, self.b2 + o.b2, self.b12 + o.b12)

    __radd__ = __add__

    def __neg__(self):
        return HD(-self.a, -self.b1, -self.b2, -self.b12)

    def __sub__(self, o):
        return self + (-HD.lift(o))

    def __rsub__(self, o):
        return HD.lift(o) - self

    def __mul__(self, o):
        o = HD.lift(o)
        return HD(self.a * o.a, self.a * o.b1 + self.b1 * o.a, self.a * o.b2 + self.b2 * o.a,
                  self.a * o.b12 + self.b1 * o.b2 + self.b2 * o.b1 + self.b12 * o.a)

    __rmul__ = __mul__

    def _fn(self, f, d1, d2):
        return HD(f, d1 * self.b1, d1 * self.b2, d1 * self.b12 + d2 * self.b1 * self.b2)

    def recip(self):
        return self._fn(1.0 / self.a, -1.0 / self.a ** 2, 2.0 / self.a ** 3)

    def __truediv__(self, o):
        return self * HD.lift(o).recip()

    def __rtruediv__(self, o):
        return HD.lift(o) * self.recip()

    def sqrt(self):
        r = math.sqrt(self.a)
        return self._fn(r, 0.5 / r, -0.25 / r ** 3)


def hsqrt(x):
    return x.sqrt() if isinstance(x, HD) else math.sqrt(x)


def val(x):
    return x.a if isinstance(x, HD) else x


# Element energies, transcribed from src/energy.rs:10-46 for any number type.
def axial(s, k, m):
    d = m.compression_transition
    p = (s + hsqrt(s * s + d * d) - d) * 0.5
    return 0.5 * k * (m.compression_ratio * s * s + (1.0 - m.compression_ratio) * p * p)


def membrane(x, inv, area, m):
    e1 = [x[1][k] - x[0][k] for k in range(3)]
    e2 = [x[2][k] - x[0][k] for k in range(3)]
    f0 = [e1[k] * inv[0] + e2[k] * inv[1] for k in range(3)]
    f1 = [e1[k] * inv[2] + e2[k] * inv[3] for k in range(3)]
    th = math.radians(m.grain_degrees)
    s, c = math.sin(th), math.cos(th)
    w = [f0[k] * c + f1[k] * s for k in range(3)]
    v = [f0[k] * (-s) + f1[k] * c for k in range(3)]
    ew = (w[0] * w[0] + w[1] * w[1] + w[2] * w[2] - 1.0) * 0.5
    ev = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - 1.0) * 0.5
    sh = w[0] * v[0] + w[1] * v[1] + w[2] * v[2]
    return area * (axial(ew, m.warp, m) + axial(ev, m.weft, m) + 0.5 * m.shear * sh * sh)


def cross3(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def bending(x, weight, across, m):
    e = [x[1][k] - x[0][k] for k in range(3)]
    n1 = cross3(e, [x[2][k] - x[0][k] for k in range(3)])
    n2 = cross3([x[3][k] - x[0][k] for k in range(3)], e)
    l1 = hsqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2])
    l2 = hsqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2])
    d = [n1[k] / l1 - n2[k] / l2 for k in range(3)]
    th = math.radians(m.grain_degrees)
    s, c = math.sin(th), math.cos(th)
    aw = across[0] * c + across[1] * s
    af = -across[0] * s + across[1] * c
    coef = 0.5 * weight * (m.bend_warp * aw * aw + m.bend_weft * af * af)
    return coef * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2])


class NumpyVbd:
    def __init__(self, g):
        self.g = g
        n = g.n
        self.tri_of = [[] for _ in range(n)]
        self.hinge_of = [[] for _ in range(n)]
        nb = [set() for _ in range(n)]
        for t, vs in enumerate(g.tri_vertices):
            for v in vs:
                self.tri_of[int(v)].append(t)
            for a in vs:
                for b in vs:
                    if a != b:
                        nb[int(a)].add(int(b))
        for h, vs in enumerate(g.hinge_vertices):
            for v in vs:
                self.hinge_of[int(v)].append(h)
            for a in vs:
                for b in vs:
                    if a != b:
                        nb[int(a)].add(int(b))
        # Greedy colouring in vertex order over the static graph only.
        col = [-1] * n
        self.colors = []
        for v in range(n):
            used = {col[u] for u in nb[v] if col[u] >= 0}
            c = 0
            while c in used:
                c += 1
            if c == len(self.colors):
                self.colors.append([])
            self.colors[c].append(v)
            col[v] = c
        self.last = None
        self.acc = None

    def local(self, v, p, x):
        g = self.g
        get = lambda i: p if i == v else [float(c) for c in x[i]]  # noqa: E731
        e = 0.0
        for t in self.tri_of[v]:
            vs = g.tri_vertices[t]
            e = e + membrane([get(int(i)) for i in vs], g.tri_inv_dm[t], g.tri_area[t],
                             g.materials[int(g.tri_material[t])])
        for h in self.hinge_of[v]:
            vs = g.hinge_vertices[h]
            e = e + bending([get(int(i)) for i in vs], g.hinge_weight[h], g.hinge_across[h],
                            g.materials[int(g.hinge_material[h])])
        return e

    def derivs(self, v, x):
        """Energy, gradient and Hessian of local(v) by hyper-dual evaluation."""
        grad = np.zeros(3)
        H = np.zeros((3, 3))
        e = None
        for i in range(3):
            for j in range(i, 3):
                p = [HD(float(x[v][k]), 1.0 if k == i else 0.0, 1.0 if k == j else 0.0) for k in range(3)]
                r = self.local(v, p, x)
                e = r.a
                grad[i] = r.b1
                grad[j] = r.b2
                H[i, j] = H[j, i] = r.b12
        return e, grad, H

    def step(self, st, s):
        g = self.g
        n = g.n
        dt, c = s["dt"], s["damping"]
        grav = np.array(s["gravity"], float)
        damp = 1.0 + c * dt
        valid = (self.last is not None and self.last[2] == st.time and np.array_equal(self.last[0], st.x)
                 and np.array_equal(self.last[1], st.v))
        g2 = grav @ grav
        fac = np.ones(n)
        if valid and g2 > 1e-20:
            fac = np.clip(self.acc @ grav / g2, 0.0, 1.0)
        guess = st.x + (st.v * dt + fac[:, None] * grav * dt * dt) / (1.0 + c * dt)
        target = st.x + (st.v * dt + grav * dt * dt) / damp
        inertia = g.masses * damp / (dt * dt)
        motion = np.linalg.norm(target - st.x, axis=1).max()
        r = s["radius"]
        rq = min(max(r + motion, 2 * r), 4 * r)
        self._assert_far(st.x, rq)
        radius = s["relaxation"] * rq
        centers = st.x.copy()

        def trunc(i, p):
            d = p - centers[i]
            L = np.linalg.norm(d)
            return (centers[i] + d * (radius / L), True) if L > radius else (p, False)
        x = st.x.copy()
        clipped = 0
        for i in range(n):
            if not g.pinned[i]:
                x[i], b = trunc(i, guess[i])
                clipped += b
        for _ in range(s["iterations"]):
            if clipped > 0 and clipped >= s["refresh_fraction"] * n:
                # Re-detection (src/vbd.rs:326-334): balls re-centred at the
                # current iterate; this mesh has no feature within rq.
                self._assert_far(x, rq)
                centers = x.copy()
            clipped = 0
            for color in self.colors:
                upd = []
                for v in color:
                    if g.pinned[v]:
                        upd.append((v, x[v].copy()))
                        continue
                    e, gr, H = self.derivs(v, x)
                    gr = gr + (x[v] - target[v]) * inertia[v]
                    H = H + np.eye(3) * inertia[v]
                    lower = min(H[i, i] - sum(abs(H[j, i]) for j in range(3) if j != i) for i in range(3))
                    if lower < 1e-8:
                        H = H + np.eye(3) * (1e-8 - lower + 0.01 * inertia[v])
                    delta = -np.linalg.solve(H, gr)
                    old = e + 0.5 * inertia[v] * np.sum((x[v] - target[v]) ** 2)
                    acc = x[v].copy()
                    hit = False
                    for a in range(10):
                        cand, b = trunc(v, x[v] + delta * 0.5 ** a)
                        hit = hit or b
                        ec = self.local(v, list(cand), x) + 0.5 * inertia[v] * np.sum((cand - target[v]) ** 2)
                        if np.all(np.isfinite(cand)) and math.isfinite(ec) and ec <= old + 1e-14:
                            acc = cand
                            break
                    upd.append((v, acc))
                    clipped += hit
                for v, p in upd:
                    x[v] = p
        old_v = st.v.copy()
        st.v = np.ascontiguousarray((x - st.x) / dt)
        st.x = np.ascontiguousarray(x)
        st.time += dt
        self.acc = (st.v - old_v) / dt + c * st.v
        self.last = (st.x.copy(), st.v.copy(), st.time)

    def _assert_far(self, x, rq):
        """Every vertex farther than rq from every vertex it shares no face
        with, and the mesh coarse enough that no non-adjacent feature pair
        comes within rq (edge length > 4 rq)."""
        g = self.g
        e = g.tri_vertices
        lens = [np.linalg.norm(x[int(a)] - x[int(b)]) for t in e for a, b in ((t[0], t[1]), (t[1], t[2]))]
        assert min(lens) > 4 * rq, "numpy restatement covers only meshes coarser than 4 rq"
