// ORACLE — test infrastructure only.  A CPU f64 restatement of the reference
// Vertex Block Descent path, used as the parity checker and as the CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it; the product library never links or calls it.
//
// Each function cites the reference file:line it restates.  Operation order
// follows the reference where that fixes f64 rounding (sums left to right,
// products in written order), so the oracle is a bit-level restatement
// wherever the reference's own arithmetic is (glam/libm rounding aside).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cvo {

constexpr double INF = std::numeric_limits<double>::infinity();
constexpr double EPS64 = std::numeric_limits<double>::epsilon();

// ------------------------------------------------------------------ DVec3
// glam DVec3 arithmetic (component-wise, dot = (x*x)+(y*y)+(z*z)).
struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(V3 a, V3 b) { return !(a == b); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline double length2(V3 a) { return dot(a, a); }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline double distance(V3 a, V3 b) { return length(a - b); }
inline double distance2(V3 a, V3 b) { return length2(a - b); }
inline bool finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }  // f64::min ignores NaN
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline V3 vmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline double max_element(V3 a) { return fmax_(fmax_(a.x, a.y), a.z); }

// glam DMat3 (column-major) with the cofactor inverse used by VbdSolver::step.
struct M3 {
  V3 c[3];  // columns
};
inline M3 m3_inverse(const M3& m) {
  V3 t0 = cross(m.c[1], m.c[2]);
  V3 t1 = cross(m.c[2], m.c[0]);
  V3 t2 = cross(m.c[0], m.c[1]);
  double det = dot(m.c[2], t2);
  double inv = 1.0 / det;
  V3 a = t0 * inv, b = t1 * inv, c = t2 * inv;
  // from_cols(a,b,c).transpose()
  M3 r;
  r.c[0] = {a.x, b.x, c.x};
  r.c[1] = {a.y, b.y, c.y};
  r.c[2] = {a.z, b.z, c.z};
  return r;
}
inline V3 m3_mul(const M3& m, V3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }

// ------------------------------------------------------------------ Jet<N>
// Second-order forward AD — src/math.rs:33-122.
template <int N>
struct Jet {
  double v = 0;
  double g[N] = {};
  double h[N][N] = {};
  static Jet c(double x) {
    Jet r;
    r.v = x;
    return r;
  }
  static Jet variable(double x, int i) {
    Jet r = c(x);
    r.g[i] = 1.0;
    return r;
  }
  Jet chain(double value, double first, double second) const {  // src/math.rs:44-53
    Jet r;
    r.v = value;
    for (int i = 0; i < N; ++i) r.g[i] = first * g[i];
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) r.h[i][j] = first * h[i][j] + second * g[i] * g[j];
    return r;
  }
};
template <int N>
Jet<N> operator+(const Jet<N>& a, const Jet<N>& b) {  // src/math.rs:60-67
  Jet<N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j];
  return r;
}
template <int N>
Jet<N> operator*(const Jet<N>& a, const Jet<N>& b) {  // src/math.rs:84-100
  Jet<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      r.h[i][j] = a.h[i][j] * b.v + a.v * b.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
  return r;
}
template <int N>
Jet<N> operator-(const Jet<N>& a) {  // Neg = self * c(-1) (src/math.rs:70-74)
  return a * Jet<N>::c(-1.0);
}
template <int N>
Jet<N> operator-(const Jet<N>& a, const Jet<N>& b) {  // Sub = self + -b
  return a + (-b);
}
template <int N>
Jet<N> operator/(const Jet<N>& a, const Jet<N>& b) {  // src/math.rs:101-106
  double bv = b.v;
  return a * b.chain(1.0 / bv, -1.0 / (bv * bv), 2.0 / (bv * bv * bv));
}

// Scalar trait — src/math.rs:4-16.
template <class S>
struct Sc;
template <>
struct Sc<double> {
  static double c(double v) { return v; }
  static double value(double v) { return v; }
  static double sqrt(double v) { return std::sqrt(v); }
  static double ln(double v) { return std::log(v); }
};
template <int N>
struct Sc<Jet<N>> {
  static Jet<N> c(double v) { return Jet<N>::c(v); }
  static double value(const Jet<N>& v) { return v.v; }
  static Jet<N> sqrt(const Jet<N>& a) {  // src/math.rs:50-53
    double r = std::sqrt(a.v);
    return a.chain(r, 0.5 / r, -0.25 / (r * r * r));
  }
  static Jet<N> ln(const Jet<N>& a) {  // src/math.rs:54-56
    return a.chain(std::log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v));
  }
};
template <class S>
using A3 = std::array<S, 3>;
template <class S>
A3<S> sub(const A3<S>& a, const A3<S>& b) {  // src/math.rs:124-126
  return {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
}
template <class S>
S sdot(const A3<S>& a, const A3<S>& b) {  // src/math.rs:127-129
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class S>
A3<S> scross(const A3<S>& a, const A3<S>& b) {  // src/math.rs:130-136
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// ------------------------------------------------------------------ model
struct Material {  // src/model.rs:7-38
  double warp = 350.0, weft = 180.0, shear = 45.0, compression_ratio = 0.15,
         compression_transition = 0.005, bend_warp = 0.00008, bend_weft = 0.00004, density = 0.2,
         thickness = 0.001, grain_degrees = 0.0;
};
struct Triangle {  // src/model.rs:91-98
  std::array<int, 3> v;
  double inv_dm[4];  // x_axis.x, x_axis.y, y_axis.x, y_axis.y
  double area;
  int material;
};
struct Hinge {  // src/model.rs:99-106
  std::array<int, 4> v;
  double weight;
  double across[2];
  int material;
};
struct StitchPointT {  // src/garment/seam.rs:216-219
  std::vector<std::pair<int, double>> terms;
  V3 gap(const std::vector<V3>& x) const {  // src/garment/seam.rs:236-238 (Sum starts at 0)
    V3 s{0, 0, 0};
    for (auto& t : terms) s = s + x[t.first] * t.second;
    return s;
  }
};
struct Garment {
  int n = 0;
  std::vector<double> masses;
  std::vector<std::optional<V3>> pins;
  std::vector<Triangle> triangles;
  std::vector<Hinge> hinges;
  std::vector<Material> materials;
  std::vector<StitchPointT> stitch_points;  // seam samples
  std::vector<double> stitch_stiffness;     // seam.stiffness * weight_m
  std::vector<StitchPointT> probes;
  double max_seam_gap(const std::vector<V3>& x) const {  // src/garment/template.rs:536-538
    double m = 0.0;
    for (auto& p : probes) m = fmax_(m, length(p.gap(x)));
    return m;
  }
};

// Rust f64::to_radians then sin_cos.
inline void grain_sin_cos(double degrees, double& s, double& c) {
  double r = degrees * (M_PI / 180.0);
  s = std::sin(r);
  c = std::cos(r);
}

// ---------------------------------------------------------------- energy
// src/energy.rs:10-15
template <class S>
S axial(const S& strain, double stiffness, const Material& m) {
  using T = Sc<S>;
  S delta = T::c(m.compression_transition);
  S p = (strain + T::sqrt(strain * strain + delta * delta) - delta) * T::c(0.5);
  return T::c(0.5 * stiffness) *
         (T::c(m.compression_ratio) * strain * strain + T::c(1.0 - m.compression_ratio) * p * p);
}
// src/energy.rs:17-29
template <class S>
S membrane(const std::array<A3<S>, 3>& x, const Triangle& t, const Material& m) {
  using T = Sc<S>;
  A3<S> ds0 = sub(x[1], x[0]), ds1 = sub(x[2], x[0]);
  A3<S> f0, f1;
  for (int k = 0; k < 3; ++k) f0[k] = ds0[k] * T::c(t.inv_dm[0]) + ds1[k] * T::c(t.inv_dm[1]);
  for (int k = 0; k < 3; ++k) f1[k] = ds0[k] * T::c(t.inv_dm[2]) + ds1[k] * T::c(t.inv_dm[3]);
  double s, c;
  grain_sin_cos(m.grain_degrees, s, c);
  A3<S> w, v;
  for (int i = 0; i < 3; ++i) w[i] = f0[i] * T::c(c) + f1[i] * T::c(s);
  for (int i = 0; i < 3; ++i) v[i] = f0[i] * T::c(-s) + f1[i] * T::c(c);
  S ew = (sdot(w, w) - T::c(1.0)) * T::c(0.5);
  S ev = (sdot(v, v) - T::c(1.0)) * T::c(0.5);
  S shear = sdot(w, v);
  return T::c(t.area) *
         (axial(ew, m.warp, m) + axial(ev, m.weft, m) + T::c(0.5 * m.shear) * shear * shear);
}
// src/energy.rs:31-46
template <class S>
S bending(const std::array<A3<S>, 4>& x, const Hinge& h, const Material& m) {
  using T = Sc<S>;
  A3<S> edge = sub(x[1], x[0]);
  A3<S> n1 = scross(edge, sub(x[2], x[0]));
  A3<S> n2 = scross(sub(x[3], x[0]), edge);
  S l1 = T::sqrt(sdot(n1, n1));
  S l2 = T::sqrt(sdot(n2, n2));
  A3<S> d;
  for (int i = 0; i < 3; ++i) d[i] = n1[i] / l1 - n2[i] / l2;
  double s, c;
  grain_sin_cos(m.grain_degrees, s, c);
  double aw = h.across[0] * c + h.across[1] * s;
  double af = -h.across[0] * s + h.across[1] * c;
  return T::c(0.5 * h.weight * (m.bend_warp * aw * aw + m.bend_weft * af * af)) * sdot(d, d);
}

// ------------------------------------------------------------- potential
enum Kind : int { PointPoint = 0, PointEdge = 1, PointFace = 2, EdgeEdge = 3 };

// src/ogc/potential.rs:8-22
template <class S>
S activation(const S& d, double r, double k) {
  using T = Sc<S>;
  if (T::value(d) >= r) return T::c(0.0);
  if (T::value(d) <= 0.0) return T::c(INF);
  double tau = 0.5 * r;
  if (T::value(d) >= tau) {
    S g = T::c(r) - d;
    return T::c(0.5 * k) * g * g;
  }
  return T::c(0.5 * k * tau * tau) - T::c(k * tau * tau) * T::ln(d / T::c(tau));
}
// src/ogc/geometry.rs:111-127
template <class S>
S point_edge_distance(const A3<S>& p, const A3<S>& a, const A3<S>& b) {
  using T = Sc<S>;
  A3<S> u = sub(b, a), w = sub(p, a);
  S q = sdot(w, u) / sdot(u, u);
  A3<S> d;
  for (int k = 0; k < 3; ++k) d[k] = w[k] - q * u[k];
  return T::sqrt(sdot(d, d));
}
template <class S>
S point_face_distance(const A3<S>& p, const A3<S>& a, const A3<S>& b, const A3<S>& c) {
  using T = Sc<S>;
  A3<S> n = scross(sub(b, a), sub(c, a));
  S h = sdot(sub(p, a), n) / T::sqrt(sdot(n, n));
  return T::value(h) < 0.0 ? -h : h;
}
template <class S>
S edge_edge_distance(const A3<S>& a, const A3<S>& b, const A3<S>& c, const A3<S>& d) {
  using T = Sc<S>;
  A3<S> n = scross(sub(b, a), sub(d, c));
  S h = sdot(sub(a, c), n) / T::sqrt(sdot(n, n));
  return T::value(h) < 0.0 ? -h : h;
}

struct Contact {  // src/ogc/potential.rs:31-37
  int kind;
  std::array<int, 4> vertices;
  int count;
  double mollifier_epsilon;
  // src/ogc/potential.rs:39-65
  template <class S>
  S energy(const std::array<A3<S>, 4>& p, double r, double k) const {
    using T = Sc<S>;
    S d;
    switch (kind) {
      case PointPoint: {
        A3<S> v = sub(p[0], p[1]);
        d = T::sqrt(sdot(v, v));
        break;
      }
      case PointEdge:
        d = point_edge_distance(p[0], p[1], p[2]);
        break;
      case PointFace:
        d = point_face_distance(p[0], p[1], p[2], p[3]);
        break;
      default: {
        A3<S> n = scross(sub(p[1], p[0]), sub(p[3], p[2]));
        if (T::value(sdot(n, n)) < 1e-30) return T::c(0.0);
        d = edge_edge_distance(p[0], p[1], p[2], p[3]);
      }
    }
    S e = activation(d, r, k);
    if (kind == EdgeEdge) {
      A3<S> n = scross(sub(p[1], p[0]), sub(p[3], p[2]));
      S q = sdot(n, n) / T::c(mollifier_epsilon);
      if (T::value(q) < 1.0) e = e * q * (T::c(2.0) - q);
    }
    return e;
  }
  double value(const std::vector<V3>& x, double r, double k) const {  // :213-215
    std::array<A3<double>, 4> p;
    for (int i = 0; i < 4; ++i) p[i] = {x[vertices[i]].x, x[vertices[i]].y, x[vertices[i]].z};
    return energy<double>(p, r, k);
  }
  std::array<int, 5> key() const { return {kind, vertices[0], vertices[1], vertices[2], vertices[3]}; }
};

// -------------------------------------------------------------- geometry
enum FeatureKind { FVertex = 0, FEdge = 1, FFace = 2 };
struct Feature {
  int kind;
  int index;
};

// src/ogc/topology.rs:13-52
struct Topology {
  std::vector<std::array<int, 3>> triangles;
  std::vector<std::array<int, 2>> edges;
  std::vector<std::array<int, 3>> triangle_edges;
  std::vector<std::vector<int>> neighbors;
  std::vector<std::vector<int>> opposites;
  static Topology build(int n, const std::vector<std::array<int, 3>>& tris) {
    Topology t;
    std::map<std::pair<int, int>, int> map;
    std::vector<std::set<int>> nb(n);
    for (auto& tr : tris) {
      for (int k = 0; k < 3; ++k)
        if (tr[k] < 0 || tr[k] >= n) throw std::runtime_error("T|OGC: invalid triangle indices");
      if (tr[0] == tr[1] || tr[1] == tr[2] || tr[0] == tr[2])
        throw std::runtime_error("T|OGC: invalid triangle indices");
      std::array<int, 3> te;
      for (int k = 0; k < 3; ++k) {
        int a = tr[k], b = tr[(k + 1) % 3], c = tr[(k + 2) % 3];
        auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = map.find(key);
        int e;
        if (it == map.end()) {
          t.edges.push_back({key.first, key.second});
          t.opposites.push_back({});
          e = (int)t.edges.size() - 1;
          map.emplace(key, e);
        } else {
          e = it->second;
        }
        t.opposites[e].push_back(c);
        nb[a].insert(b);
        nb[b].insert(a);
        te[k] = e;
      }
      t.triangle_edges.push_back(te);
    }
    for (auto& o : t.opposites)
      if (o.size() > 2)
        throw std::runtime_error(
            "T|OGC requires manifold triangle edges (at most two incident faces)");
    t.triangles = tris;
    t.neighbors.resize(n);
    for (int i = 0; i < n; ++i) t.neighbors[i].assign(nb[i].begin(), nb[i].end());
    return t;
  }
};

// src/ogc/geometry.rs:13-17
inline bool vertex_feasible(V3 p, int v, const std::vector<V3>& x, const Topology& top) {
  for (int j : top.neighbors[v])
    if (!(dot(p - x[v], x[v] - x[j]) >= -1e-14)) return false;
  return true;
}
// src/ogc/geometry.rs:18-29
inline bool edge_feasible(V3 p, int e, const std::vector<V3>& x, const Topology& top) {
  V3 a = x[top.edges[e][0]], b = x[top.edges[e][1]];
  V3 u = b - a;
  double t = dot(p - a, u) / length2(u);
  if (!(t > 0.0 && t < 1.0)) return false;
  for (int i : top.opposites[e]) {
    V3 c = x[i];
    V3 q = a + u * (dot(c - a, u) / length2(u));
    if (!(dot(p - q, q - c) >= -1e-14)) return false;
  }
  return true;
}
// src/ogc/geometry.rs:31-77
inline std::pair<V3, Feature> closest_triangle(V3 p, int t, const std::vector<V3>& x,
                                               const Topology& top) {
  auto ids = top.triangles[t];
  V3 a = x[ids[0]], b = x[ids[1]], c = x[ids[2]];
  V3 ab = b - a, ac = c - a, ap = p - a;
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return {a, {FVertex, ids[0]}};
  V3 bp = p - b;
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return {b, {FVertex, ids[1]}};
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0)
    return {a + ab * (d1 / (d1 - d3)), {FEdge, top.triangle_edges[t][0]}};
  V3 cp = p - c;
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return {c, {FVertex, ids[2]}};
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0)
    return {a + ac * (d2 / (d2 - d6)), {FEdge, top.triangle_edges[t][2]}};
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0)
    return {b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6))),
            {FEdge, top.triangle_edges[t][1]}};
  double denom = 1.0 / (va + vb + vc);
  return {a + ab * (vb * denom) + ac * (vc * denom), {FFace, t}};
}
// src/ogc/geometry.rs:78-107
inline std::pair<double, double> closest_segments(V3 a, V3 b, V3 c, V3 d) {
  V3 u = b - a, v = d - c, w = a - c;
  double aa = length2(u), ee = length2(v);
  if (aa < 1e-24 || ee < 1e-24) return {0.0, 0.0};
  double bb = dot(u, v), cc = dot(u, w), ff = dot(v, w);
  double den = aa * ee - bb * bb;
  double s = den > 1e-14 * aa * ee ? std::clamp((bb * ff - cc * ee) / den, 0.0, 1.0) : 0.0;
  double t = (bb * s + ff) / ee;
  if (t < 0.0) {
    s = std::clamp(-cc / aa, 0.0, 1.0);
    t = 0.0;
  } else if (t > 1.0) {
    s = std::clamp((bb - cc) / aa, 0.0, 1.0);
    t = 1.0;
  }
  return {s, t};
}
// src/ogc/geometry.rs:131-147
inline bool segment_triangle(V3 a, V3 b, const std::array<V3, 3>& tri) {
  V3 p = tri[0], q = tri[1], r = tri[2];
  V3 e = q - p, f = r - p, d = b - a;
  V3 h = cross(d, f);
  double det = dot(e, h);
  if (std::fabs(det) < 1e-14) return false;
  V3 s = a - p;
  double u = dot(s, h) / det;
  V3 z = cross(s, e);
  double v = dot(d, z) / det;
  double k = dot(f, z) / det;
  return u >= -1e-10 && v >= -1e-10 && u + v <= 1.0 + 1e-10 && (k >= 1e-10 && k < 1.0 - 1e-10);
}

// ------------------------------------------------------------------- bvh
struct Bounds {  // src/ogc/bvh.rs:4-43
  V3 min{INF, INF, INF}, max{-INF, -INF, -INF};
  template <class It>
  static Bounds points(It begin, It end) {
    Bounds b;
    for (auto it = begin; it != end; ++it) {
      b.min = vmin(b.min, *it);
      b.max = vmax(b.max, *it);
    }
    return b;
  }
  static Bounds of(std::initializer_list<V3> p) { return points(p.begin(), p.end()); }
  Bounds inflate(double r) const {
    return {min - V3{r, r, r}, max + V3{r, r, r}};
  }
  bool overlaps(const Bounds& b) const {
    return min.x <= b.max.x && min.y <= b.max.y && min.z <= b.max.z && max.x >= b.min.x &&
           max.y >= b.min.y && max.z >= b.min.z;
  }
  double distance_squared(const Bounds& b) const {
    return length2(vmax(vmax(min - b.max, b.min - max), V3{0, 0, 0}));
  }
  Bounds unite(const Bounds& b) const { return {vmin(min, b.min), vmax(max, b.max)}; }
};

struct Bvh {  // src/ogc/bvh.rs:44-243
  struct Node {
    Bounds bounds;
    int children[2] = {-1, -1};
    int start = 0, end = 0;
    bool dynamic = false;
    bool leaf() const { return children[0] < 0; }
  };
  std::vector<Node> nodes;
  std::vector<int> order;
  std::vector<int> roots;
  std::vector<Bounds> primitive_bounds;
  int dynamic_nodes = 0;

  static Bvh build(const std::vector<Bounds>& bounds, const std::vector<bool>& dynamic) {
    Bvh b;
    int n = (int)bounds.size();
    for (int i = 0; i < n; ++i)
      if (dynamic[i]) b.order.push_back(i);
    for (int i = 0; i < n; ++i)
      if (!dynamic[i]) b.order.push_back(i);
    b.primitive_bounds = bounds;
    int split = 0;
    for (bool d : dynamic) split += d ? 1 : 0;
    std::pair<int, int> ranges[2] = {{0, split}, {split, n}};
    for (auto [start, end] : ranges) {
      if (start < end) {
        int root = b.build_node(bounds, dynamic, start, end);
        b.roots.push_back(root);
        if (start == 0 && split > 0) b.dynamic_nodes = (int)b.nodes.size();
      }
    }
    return b;
  }
  int build_node(const std::vector<Bounds>& bounds, const std::vector<bool>& dynamic, int start,
                 int end) {
    Bounds box = bounds[order[start]];
    for (int i = start + 1; i < end; ++i) box = box.unite(bounds[order[i]]);
    int id = (int)nodes.size();
    Node node;
    node.bounds = box;
    node.start = start;
    node.end = end;
    node.dynamic = false;
    for (int i = start; i < end; ++i) node.dynamic = node.dynamic || dynamic[order[i]];
    nodes.push_back(node);
    if (end - start > 8) {
      V3 d = box.max - box.min;
      int axis = (d.x > d.y && d.x > d.z) ? 0 : (d.y > d.z ? 1 : 2);
      std::sort(order.begin() + start, order.begin() + end, [&](int a, int b) {
        double ka = bounds[a].min[axis] + bounds[a].max[axis];
        double kb = bounds[b].min[axis] + bounds[b].max[axis];
        return ka < kb;
      });
      int mid = (start + end) / 2;
      int a = build_node(bounds, dynamic, start, mid);
      int c = build_node(bounds, dynamic, mid, end);
      nodes[id].children[0] = a;
      nodes[id].children[1] = c;
    }
    return id;
  }
  template <class F>
  void refit_dynamic(F&& bound) {  // src/ogc/bvh.rs:138-154
    for (int i = dynamic_nodes - 1; i >= 0; --i) {
      Node& n = nodes[i];
      if (!n.leaf()) {
        n.bounds = nodes[n.children[0]].bounds.unite(nodes[n.children[1]].bounds);
      } else {
        for (int k = n.start; k < n.end; ++k) primitive_bounds[order[k]] = bound(order[k]);
        Bounds b = primitive_bounds[order[n.start]];
        for (int k = n.start + 1; k < n.end; ++k) b = b.unite(primitive_bounds[order[k]]);
        n.bounds = b;
      }
    }
  }
  template <class F>
  void query(const Bounds& q, bool dynamic_only, F&& visit) const {  // :155-185
    if (nodes.empty()) return;
    std::vector<int> stack(roots.begin(), roots.end());
    while (!stack.empty()) {
      int i = stack.back();
      stack.pop_back();
      const Node& n = nodes[i];
      if ((dynamic_only && !n.dynamic) || !n.bounds.overlaps(q)) continue;
      if (!n.leaf()) {
        stack.push_back(n.children[0]);
        stack.push_back(n.children[1]);
      } else {
        for (int k = n.start; k < n.end; ++k)
          if (primitive_bounds[order[k]].overlaps(q)) visit(order[k]);
      }
    }
  }
  template <class F>
  void pairs(const Bvh& other, double radius, bool same, F&& visit) const {  // :198-242
    std::vector<std::pair<int, int>> stack;
    for (size_t i = 0; i < roots.size(); ++i)
      for (size_t j = same ? i : 0; j < other.roots.size(); ++j)
        stack.push_back({roots[i], other.roots[j]});
    double radius2 = radius * radius;
    while (!stack.empty()) {
      auto [a, b] = stack.back();
      stack.pop_back();
      const Node& na = nodes[a];
      const Node& nb = other.nodes[b];
      if ((!na.dynamic && !nb.dynamic) || na.bounds.distance_squared(nb.bounds) > radius2) continue;
      if (same && a == b) {
        if (!na.leaf()) {
          int l = na.children[0], r = na.children[1];
          stack.push_back({l, l});
          stack.push_back({l, r});
          stack.push_back({r, r});
          continue;
        }
      } else if (!na.leaf() && (nb.leaf() || na.end - na.start >= nb.end - nb.start)) {
        stack.push_back({na.children[0], b});
        stack.push_back({na.children[1], b});
        continue;
      } else if (!nb.leaf()) {
        stack.push_back({a, nb.children[0]});
        stack.push_back({a, nb.children[1]});
        continue;
      }
      for (int off = 0; off < na.end - na.start; ++off) {
        int i = order[na.start + off];
        int start = nb.start + ((same && a == b) ? off + 1 : 0);
        for (int k = start; k < nb.end; ++k) {
          int j = other.order[k];
          if (primitive_bounds[i].distance_squared(other.primitive_bounds[j]) <= radius2) {
            if (same)
              visit(std::min(i, j), std::max(i, j));
            else
              visit(i, j);
          }
        }
      }
    }
  }
};

// ------------------------------------------------------------ candidates
struct Candidates {  // src/ogc/candidates.rs:5-68
  std::vector<V3> centers;
  double radius = 0, skin = 0;
  std::vector<std::array<int, 2>> vf, ee;
  bool covers(const std::vector<V3>& x, int moving, double r) const {
    if (!(r <= radius) || moving != (int)centers.size()) return false;
    for (int i = 0; i < moving; ++i)
      if (!(distance2(x[i], centers[i]) <= skin * skin)) return false;
    return true;
  }
};

// --------------------------------------------------------------- settings
struct OgcSettings {  // src/ogc/world.rs:33-71
  double radius = 0.003, stiffness = 100000.0, relaxation = 0.45, refresh_fraction = 0.01;
  void validate() const {
    if (!std::isfinite(radius) || radius <= 0.0 || !std::isfinite(stiffness) || stiffness <= 0.0 ||
        !std::isfinite(relaxation) || relaxation <= 0.0 || relaxation >= 0.5 ||
        !std::isfinite(refresh_fraction) || !(refresh_fraction >= 0.0 && refresh_fraction <= 1.0))
      throw std::runtime_error("O|Invalid OGC radius, stiffness or trust-region parameters");
  }
};

struct DetectionProfile {
  double refit_ms = 0, candidates_ms = 0, narrow_ms = 0;
  uint64_t candidate_pairs = 0, aabb_rejections = 0;
};

struct ContactSet {  // src/ogc/world.rs:73-107
  DetectionProfile profile;
  std::vector<Contact> contacts;
  uint64_t active_contacts = 0;
  std::vector<std::vector<int>> incident;
  std::vector<V3> centers;
  std::vector<double> radii;
  double minimum_distance = 0;
  uint64_t vertex_face_candidates = 0, edge_edge_candidates = 0;
  bool broad_phase_rebuilt = false;
  std::pair<V3, bool> truncate(int v, V3 proposed) const {  // :91-100
    V3 d = proposed - centers[v];
    double len = length(d);
    double b = radii[v];
    if (len > b) return {centers[v] + d * (b / len), true};
    return {proposed, false};
  }
  double value(const std::vector<V3>& x, const OgcSettings& s) const {
    double sum = 0.0;
    for (auto& c : contacts) sum += c.value(x, s.radius, s.stiffness);
    return sum;
  }
};

double now_ms();

// ------------------------------------------------------------------ world
struct OgcWorld {  // src/ogc/world.rs:111-558
  Topology topology;
  int moving = 0;
  std::vector<V3> fixed;
  Bvh triangles, edges, points;
  std::vector<double> rest_edge_length2;
  std::set<std::pair<int, int>> excluded;
  std::vector<std::vector<int>> excluded_neighbors;
  std::vector<int> moving_faces;
  std::optional<Candidates> candidates;

  static OgcWorld create(const std::vector<V3>& x, const std::vector<std::array<int, 3>>& tris,
                         int moving, const std::vector<std::pair<int, int>>& excl) {
    OgcWorld w;
    if (moving > (int)x.size()) throw std::runtime_error("P|Invalid OGC positions");
    for (auto& p : x)
      if (!finite(p)) throw std::runtime_error("P|Invalid OGC positions");
    w.topology = Topology::build((int)x.size(), tris);
    for (auto& t : w.topology.triangles) {
      V3 a = x[t[0]], b = x[t[1]], c = x[t[2]];
      if (length2(cross(b - a, c - a)) < 1e-24)
        throw std::runtime_error("D|OGC degenerate initial face");
    }
    std::vector<bool> tdyn, edyn, pdyn;
    std::vector<Bounds> tb, eb, pb;
    for (size_t i = 0; i < w.topology.triangles.size(); ++i) {
      auto& t = w.topology.triangles[i];
      bool d = t[0] < moving || t[1] < moving || t[2] < moving;
      if (d) w.moving_faces.push_back((int)i);
      tdyn.push_back(d);
      tb.push_back(Bounds::of({x[t[0]], x[t[1]], x[t[2]]}));
    }
    for (auto& e : w.topology.edges) {
      edyn.push_back(e[0] < moving || e[1] < moving);
      eb.push_back(Bounds::of({x[e[0]], x[e[1]]}));
      w.rest_edge_length2.push_back(distance2(x[e[0]], x[e[1]]));
    }
    for (size_t i = 0; i < x.size(); ++i) {
      pdyn.push_back((int)i < moving);
      pb.push_back(Bounds::of({x[i]}));
    }
    w.triangles = Bvh::build(tb, tdyn);
    w.edges = Bvh::build(eb, edyn);
    w.points = Bvh::build(pb, pdyn);
    for (auto& [a, b] : excl) w.excluded.insert({std::min(a, b), std::max(a, b)});
    w.excluded_neighbors.assign(x.size(), {});
    for (auto& [a, b] : w.excluded) {
      if (b >= (int)x.size()) throw std::runtime_error("A|OGC invalid excluded vertex");
      w.excluded_neighbors[a].push_back(b);
      w.excluded_neighbors[b].push_back(a);
    }
    for (auto& v : w.excluded_neighbors) std::sort(v.begin(), v.end());
    w.moving = moving;
    w.fixed.assign(x.begin() + moving, x.end());
    return w;
  }
  std::vector<V3> positions(const std::vector<V3>& mov) const {
    std::vector<V3> r(mov);
    r.insert(r.end(), fixed.begin(), fixed.end());
    return r;
  }
  bool adjacent(int a, int b) const {
    return a == b ||
           std::binary_search(excluded_neighbors[a].begin(), excluded_neighbors[a].end(), b);
  }
  void refit(const std::vector<V3>& x) {  // :218-241
    if (x.size() != topology.neighbors.size())
      throw std::runtime_error("P|OGC invalid position array");
    for (auto& p : x)
      if (!finite(p)) throw std::runtime_error("P|OGC invalid position array");
    for (size_t i = 0; i < fixed.size(); ++i)
      if (x[moving + i] != fixed[i])
        throw std::runtime_error("F|OGC fixed geometry changed; rebuild the world");
    for (int f : moving_faces) {
      auto& t = topology.triangles[f];
      V3 a = x[t[0]], b = x[t[1]], c = x[t[2]];
      if (length2(cross(b - a, c - a)) < 1e-24)
        throw std::runtime_error("D|OGC degenerate face; step rejected");
    }
    triangles.refit_dynamic([&](int i) {
      auto& t = topology.triangles[i];
      return Bounds::of({x[t[0]], x[t[1]], x[t[2]]});
    });
    edges.refit_dynamic([&](int i) {
      auto& e = topology.edges[i];
      return Bounds::of({x[e[0]], x[e[1]]});
    });
    points.refit_dynamic([&](int i) { return Bounds::of({x[i]}); });
  }
  void validate_start(const std::vector<V3>& x);  // :242-307
  ContactSet detect(const std::vector<V3>& x, double rq, const OgcSettings& s, bool prune = true);
};

// -------------------------------------------------------------- coloring
// src/vbd/coloring.rs:5-43
struct GrabT {
  std::array<int, 3> vertices;
  std::array<double, 3> weights;
  V3 target;
  double stiffness;
  bool contains(int v) const { return vertices[0] == v || vertices[1] == v || vertices[2] == v; }
};
// `order` (optional): the vertex sequence to colour in; the reference colours
// in index order.  The fast-colouring extension passes smallest_last(graph).
inline std::vector<std::vector<int>> colors(const std::vector<std::vector<int>>& graph,
                                            const ContactSet& contact, const GrabT* grab,
                                            const std::vector<int>* order = nullptr) {
  int n = (int)graph.size();
  const int UNSET = -1;
  std::vector<int> assigned(n, UNSET);
  std::vector<std::vector<int>> out;
  std::vector<int> seen(n + 1, -1);
  for (int idx = 0; idx < n; ++idx) {
    const int v = order ? (*order)[idx] : idx;
    auto mark = [&](int i) {
      if (i < n && assigned[i] != UNSET) seen[assigned[i]] = v;
    };
    for (int i : graph[v]) mark(i);
    for (int idx : contact.incident[v]) {
      auto& c = contact.contacts[idx];
      for (int k = 0; k < c.count; ++k) mark(c.vertices[k]);
    }
    if (grab && grab->contains(v))
      for (int i : grab->vertices) mark(i);
    int c = 0;
    while (seen[c] == v) ++c;
    if (c == (int)out.size()) out.push_back({});
    out[c].push_back(v);
    assigned[v] = c;
  }
  return out;
}

// Smallest-last order (extension, no reference counterpart): remove a vertex
// of least remaining degree, ties to the smaller index, until none is left;
// colour in the reverse removal order.  Restated here with bucket lists.
inline std::vector<int> smallest_last(const std::vector<std::vector<int>>& graph) {
  int n = (int)graph.size();
  std::vector<int> deg(n);
  int maxd = 0;
  for (int v = 0; v < n; ++v) {
    deg[v] = (int)graph[v].size();
    maxd = std::max(maxd, deg[v]);
  }
  std::vector<std::set<int>> bucket(maxd + 1);
  for (int v = 0; v < n; ++v) bucket[deg[v]].insert(v);
  std::vector<char> gone(n, 0);
  std::vector<int> removal;
  for (int step = 0; step < n; ++step) {
    int d = 0;
    while (bucket[d].empty()) ++d;
    int v = *bucket[d].begin();
    bucket[d].erase(bucket[d].begin());
    gone[v] = 1;
    removal.push_back(v);
    for (int u : graph[v])
      if (u < n && !gone[u]) {
        bucket[deg[u]].erase(u);
        bucket[--deg[u]].insert(u);
      }
  }
  return std::vector<int>(removal.rbegin(), removal.rend());
}

// Fast colouring (extension, no reference counterpart; the product computes
// it on the device).  The static conflict graph is coloured once, greedily in
// smallest-last order (colours col0).  Each detection then colours the graph
// with its contacts and grab greedily in the order of the static colour
// classes: col0 ascending, ties by decreasing priority.  Without contacts this
// reproduces col0 exactly (each vertex of class c sees colours 0..c-1 on its
// earlier neighbours), and a greedy pass in the class order of a colouring
// needs no more colours than that colouring has, so contacts are absorbed by
// the whole graph rather than by new colours.  It is the result of
// Jones-Plassmann rounds with that priority (a vertex whose uncoloured
// neighbours all come later takes the smallest free colour), restated
// sequentially.  Priority: the lowbias32 integer hash of the vertex index (a
// bijection on 32 bits, so the order is total).
inline uint32_t repair_priority(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}
inline std::vector<std::vector<int>> fast_colors(const std::vector<std::vector<int>>& graph,
                                                 const ContactSet& contact, const GrabT* grab,
                                                 const std::vector<int>& sl) {
  const int n = (int)graph.size();
  std::vector<int> col0(n, -1), seen(n + 1, -1);
  for (int v : sl) {  // static colouring
    for (int u : graph[v])
      if (u < n && col0[u] >= 0) seen[col0[u]] = v;
    int c = 0;
    while (seen[c] == v) ++c;
    col0[v] = c;
  }
  std::vector<int> order(n);
  for (int v = 0; v < n; ++v) order[v] = v;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (col0[a] != col0[b]) return col0[a] < col0[b];
    return repair_priority(a) > repair_priority(b);
  });
  return colors(graph, contact, grab, &order);
}

// ------------------------------------------------------------ prediction
struct VbdSettings {  // src/vbd.rs:18-52
  double dt = 1.0 / 240.0;
  int iterations = 16;
  double damping = 1.5;
  V3 gravity{0.0, -9.81, 0.0};
  OgcSettings contact;
  // Extension with no reference counterpart: > 0 ends the iterations once the
  // post-sweep (true) residual is <= this value.
  double residual_tolerance = 0.0;
  uint32_t residual_interval = 0;  // residual after sweeps k, 2k, ... and the last (0, 1: every sweep)
  // Extension: the fast colouring (fast_colors) in place of the reference's.
  bool fast_coloring = false;
  void validate() const {
    contact.validate();
    if (!std::isfinite(dt) || !(dt >= 1e-6 && dt <= 0.025) || iterations == 0 || iterations > 200 ||
        !std::isfinite(damping) || damping < 0.0 || !finite(gravity) ||
        !std::isfinite(residual_tolerance) || residual_tolerance < 0.0)
      throw std::runtime_error("S|Invalid VBD settings");
  }
};
struct State {
  std::vector<V3> x, v;
  double time = 0;
};
// src/vbd/prediction.rs:15-54
struct Prediction {
  std::optional<State> last;
  std::vector<V3> acceleration;
  V3 gravity{0, 0, 0};
  double damping = 0;
  std::vector<V3> guess(const State& s, const VbdSettings& st) const {
    bool valid = last && last->time == s.time && last->x == s.x && last->v == s.v &&
                 gravity == st.gravity && damping == st.damping;
    double g2 = length2(st.gravity);
    std::vector<V3> out(s.x.size());
    for (size_t i = 0; i < s.x.size(); ++i) {
      double factor = 1.0;
      if (valid && g2 > 1e-20) factor = std::clamp(dot(acceleration[i], st.gravity) / g2, 0.0, 1.0);
      out[i] = s.x[i] + (s.v[i] * st.dt + (factor * st.gravity) * (st.dt * st.dt)) /
                            (1.0 + st.damping * st.dt);
    }
    return out;
  }
  void record(const std::vector<V3>& old_velocity, const State& s, const VbdSettings& st) {
    acceleration.resize(s.v.size());
    for (size_t i = 0; i < s.v.size(); ++i)
      acceleration[i] = (s.v[i] - old_velocity[i]) / st.dt + st.damping * s.v[i];
    last = s;
    gravity = st.gravity;
    damping = st.damping;
  }
};

struct VbdMetrics {  // src/vbd.rs:61-83
  double time_s = 0, dt_s = 0, step_ms = 0, detection_ms = 0, solve_ms = 0;
  DetectionProfile profile;
  int iterations = 0;
  uint64_t contacts = 0, contact_terms = 0, detection_updates = 0, broad_phase_rebuilds = 0,
           truncated_vertices = 0;
  double minimum_distance_m = 0, max_seam_gap_m = 0, extension_p95 = 0, extension_max = 0;
  double residual_rms_n = std::numeric_limits<double>::quiet_NaN();
  double true_residual_rms_n = std::numeric_limits<double>::quiet_NaN();
  uint32_t colors = 0;
  int iterations_run = 0;
};

struct Stitch {
  std::vector<std::pair<int, double>> terms;
  double stiffness;
};

// ---------------------------------------------------------- diagnostics
std::array<double, 2> extension_percent(const Garment& g, const std::vector<V3>& x);

// ------------------------------------------------------------ VbdSolver
struct VbdSolver {  // src/vbd.rs:89-422
  Prediction prediction;
  // Test hook: a guess for the next step in place of Prediction::guess (the
  // learned warm start's parity test installs the device network's guess).
  std::vector<V3> next_guess;
  Garment garment;
  OgcWorld world;
  std::vector<std::vector<int>> tri_of, hinge_of, stitch_of, graph;
  std::vector<Stitch> stitches;
  int threads = 1;
  // Introspection of the last step (test hooks).
  ContactSet last_contacts;
  std::vector<std::vector<int>> last_colors;
  std::vector<V3> last_x;  // final iterate of the last step (moving + fixed)
  std::vector<double> last_gsq;

  static VbdSolver create(const Garment& g, const std::vector<V3>* avatar_v,
                          const std::vector<std::array<int, 3>>* avatar_t, const State& s);
  template <class S>
  S local(int v, const A3<S>& p, const std::vector<V3>& x, const ContactSet& contact,
          const VbdSettings& st, const GrabT* grab) const;
  VbdMetrics step(State& state, const VbdSettings& st, const GrabT* grab);
  // True post-sweep residual sqrt(sum_free |g|^2 / 3 n_free) at positions x.
  double true_residual(const std::vector<V3>& x, const std::vector<V3>& target,
                       const std::vector<double>& inertia, const ContactSet& c,
                       const VbdSettings& st, const GrabT* grab) const;
};

// ------------------------------------------------------------- Newton
struct SolverSettings {  // src/solver.rs:13-58
  double dt = 1.0 / 60.0;
  V3 gravity{0.0, -9.81, 0.0};
  double damping = 0.4;
  int newton_iterations = 30, pcg_iterations = 250;
  double absolute_tolerance = 1e-6, relative_tolerance = 1e-4;
};
struct Sphere {  // src/contact.rs:37-89
  V3 center{0, 0, 0};
  double radius = 0.35, stiffness = 100000.0;
};
struct NewtonMetrics {
  int newton_iterations = 0, pcg_iterations = 0, line_search_backtracks = 0;
  bool converged = false;
  double residual_rms_n = 0, tolerance_n = 0, objective_initial_j = 0, objective_final_j = 0;
  double max_regularization = 0;
};
// Frozen OGC contacts: the extension that gives the Newton oracle the same
// objective as a contact VBD step (the reference Newton has no OGC term).
struct FrozenContacts {
  std::vector<Contact> contacts;
  double radius = 0, stiffness = 0;
  std::vector<V3> fixed;  // avatar positions appended after the garment
};
NewtonMetrics newton_step(const Garment& g, State& s, const SolverSettings& st,
                          const Sphere* sphere, const FrozenContacts* frozen);
double newton_objective(const Garment& g, const State& old, const std::vector<V3>& x,
                        const SolverSettings& st, const Sphere* sphere,
                        const FrozenContacts* frozen);

}  // namespace cvo
