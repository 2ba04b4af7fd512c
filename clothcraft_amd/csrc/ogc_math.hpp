// ogc_math.hpp — f64 narrow-phase geometry of Offset Geometric Contact,
// shared by the host start-state validation and the device detection
// kernels.  Translation units that include it are compiled with
// -ffp-contract=off (and each function also disables contraction) so a device
// evaluation rounds exactly like the reference's (uncontracted) f64 code;
// that is what makes the device candidate and contact sets bit-identical to
// the f64 restatement on identical inputs.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OG_HD __host__ __device__ __forceinline__
#else
#define OG_HD inline
#endif

namespace ogm {

struct D3 {
  double x, y, z;
};
OG_HD D3 d3(double x, double y, double z) { return D3{x, y, z}; }

OG_HD D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
OG_HD D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
OG_HD D3 scale(D3 a, double s) {
#pragma clang fp contract(off)
  return {a.x * s, a.y * s, a.z * s};
}
OG_HD double dot(D3 a, D3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
OG_HD D3 cross(D3 a, D3 b) {
#pragma clang fp contract(off)
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
OG_HD double len2(D3 a) { return dot(a, a); }
OG_HD double dist(D3 a, D3 b) { return sqrt(len2(sub(a, b))); }

// Bounds::distance_squared (src/ogc/bvh.rs:31-36).
struct Box {
  D3 lo, hi;
};
OG_HD double fmax0(double a, double b) { return a > b ? a : (b != b ? a : b); }
OG_HD double box_dist2(const Box& a, const Box& b) {
#pragma clang fp contract(off)
  double x = fmax0(fmax0(a.lo.x - b.hi.x, b.lo.x - a.hi.x), 0.0);
  double y = fmax0(fmax0(a.lo.y - b.hi.y, b.lo.y - a.hi.y), 0.0);
  double z = fmax0(fmax0(a.lo.z - b.hi.z, b.lo.z - a.hi.z), 0.0);
  return x * x + y * y + z * z;
}
OG_HD Box box_of(D3 a) { return {a, a}; }
OG_HD Box box_of(D3 a, D3 b) {
  return {{fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)},
          {fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)}};
}
OG_HD Box box_of(D3 a, D3 b, D3 c) {
  Box r = box_of(a, b);
  r.lo = {fmin(r.lo.x, c.x), fmin(r.lo.y, c.y), fmin(r.lo.z, c.z)};
  r.hi = {fmax(r.hi.x, c.x), fmax(r.hi.y, c.y), fmax(r.hi.z, c.z)};
  return r;
}

// Feature of the closest point (src/ogc/geometry.rs:6-11).
enum { F_VERTEX = 0, F_EDGE = 1, F_FACE = 2 };

// closest_triangle (src/ogc/geometry.rs:31-77).  `edge_slot` returns which
// triangle edge (0: ab, 1: bc, 2: ca) for an edge feature; `vert` the local
// vertex (0..2) for a vertex feature.
OG_HD void closest_triangle(D3 p, D3 a, D3 b, D3 c, D3& q, int& feature, int& local) {
#pragma clang fp contract(off)
  D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) {
    q = a;
    feature = F_VERTEX;
    local = 0;
    return;
  }
  D3 bp = sub(p, b);
  double d3_ = dot(ab, bp), d4 = dot(ac, bp);
  if (d3_ >= 0.0 && d4 <= d3_) {
    q = b;
    feature = F_VERTEX;
    local = 1;
    return;
  }
  double vc = d1 * d4 - d3_ * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) {
    q = add(a, scale(ab, d1 / (d1 - d3_)));
    feature = F_EDGE;
    local = 0;
    return;
  }
  D3 cp = sub(p, c);
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) {
    q = c;
    feature = F_VERTEX;
    local = 2;
    return;
  }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    q = add(a, scale(ac, d2 / (d2 - d6)));
    feature = F_EDGE;
    local = 2;
    return;
  }
  double va = d3_ * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3_ >= 0.0 && d5 - d6 >= 0.0) {
    q = add(b, scale(sub(c, b), (d4 - d3_) / ((d4 - d3_) + (d5 - d6))));
    feature = F_EDGE;
    local = 1;
    return;
  }
  double denom = 1.0 / (va + vb + vc);
  q = add(add(a, scale(ab, vb * denom)), scale(ac, vc * denom));
  feature = F_FACE;
  local = 0;
}

OG_HD double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// closest_segments (src/ogc/geometry.rs:78-107).
OG_HD void closest_segments(D3 a, D3 b, D3 c, D3 d, double& s_out, double& t_out) {
#pragma clang fp contract(off)
  D3 u = sub(b, a), v = sub(d, c), w = sub(a, c);
  double aa = len2(u), ee = len2(v);
  if (aa < 1e-24 || ee < 1e-24) {
    s_out = 0.0;
    t_out = 0.0;
    return;
  }
  double bb = dot(u, v), cc = dot(u, w), ff = dot(v, w);
  double den = aa * ee - bb * bb;
  double s = den > 1e-14 * aa * ee ? clamp01((bb * ff - cc * ee) / den) : 0.0;
  double t = (bb * s + ff) / ee;
  if (t < 0.0) {
    s = clamp01(-cc / aa);
    t = 0.0;
  } else if (t > 1.0) {
    s = clamp01((bb - cc) / aa);
    t = 1.0;
  }
  s_out = s;
  t_out = t;
}

// vertex_feasible (src/ogc/geometry.rs:13-17): neighbours from a CSR list.
OG_HD bool vertex_feasible(D3 p, int v, const D3* x, const uint32_t* nb_off, const uint32_t* nb) {
#pragma clang fp contract(off)
  D3 xv = x[v];
  D3 pv = sub(p, xv);
  for (uint32_t k = nb_off[v]; k < nb_off[v + 1]; ++k) {
    D3 e = sub(xv, x[nb[k]]);
    if (!(dot(pv, e) >= -1e-14)) return false;
  }
  return true;
}
// edge_feasible (src/ogc/geometry.rs:18-29); opposites: two slots, -1 when absent.
OG_HD bool edge_feasible(D3 p, D3 a, D3 b, const int* opp2, const D3* x) {
#pragma clang fp contract(off)
  D3 u = sub(b, a);
  double uu = len2(u);
  double t = dot(sub(p, a), u) / uu;
  if (!(t > 0.0 && t < 1.0)) return false;
  for (int k = 0; k < 2; ++k) {
    if (opp2[k] < 0) continue;
    D3 c = x[opp2[k]];
    D3 q = add(a, scale(u, dot(sub(c, a), u) / uu));
    if (!(dot(sub(p, q), sub(q, c)) >= -1e-14)) return false;
  }
  return true;
}

// Canonical id order of a contact (src/ogc/world.rs:435-452).
OG_HD void canonical(int kind, uint32_t* ids) {
  auto sw = [](uint32_t& a, uint32_t& b) {
    uint32_t t = a;
    a = b;
    b = t;
  };
  if (kind == 0 && ids[0] > ids[1]) sw(ids[0], ids[1]);
  if (kind == 1 && ids[1] > ids[2]) sw(ids[1], ids[2]);
  if (kind == 2) {
    if (ids[1] > ids[2]) sw(ids[1], ids[2]);
    if (ids[2] > ids[3]) sw(ids[2], ids[3]);
    if (ids[1] > ids[2]) sw(ids[1], ids[2]);
  }
  if (kind == 3) {
    if (ids[0] > ids[1]) sw(ids[0], ids[1]);
    if (ids[2] > ids[3]) sw(ids[2], ids[3]);
    if (ids[0] > ids[2]) {
      sw(ids[0], ids[2]);
      sw(ids[1], ids[3]);
    }
  }
}

// segment_triangle (src/ogc/geometry.rs:131-147).
OG_HD bool segment_triangle(D3 a, D3 b, D3 p, D3 q, D3 r) {
#pragma clang fp contract(off)
  D3 e = sub(q, p), f = sub(r, p), d = sub(b, a);
  D3 h = cross(d, f);
  double det = dot(e, h);
  if (fabs(det) < 1e-14) return false;
  D3 s = sub(a, p);
  double u = dot(s, h) / det;
  D3 z = cross(s, e);
  double v = dot(d, z) / det;
  double k = dot(f, z) / det;
  return u >= -1e-10 && v >= -1e-10 && u + v <= 1.0 + 1e-10 && k >= 1e-10 && k < 1.0 - 1e-10;
}

// activation and Contact::energy in f64 (src/ogc/potential.rs:8-65), used for
// the active-contact count of a detection.
OG_HD double activation(double d, double r, double k) {
#pragma clang fp contract(off)
  if (d >= r) return 0.0;
  if (d <= 0.0) return INFINITY;
  double tau = 0.5 * r;
  if (d >= tau) {
    double g = r - d;
    return 0.5 * k * g * g;
  }
  return 0.5 * k * tau * tau - k * tau * tau * log(d / tau);
}
OG_HD double contact_value(int kind, const D3* p, double eps, double r, double k) {
#pragma clang fp contract(off)
  double d;
  if (kind == 0) {
    d = sqrt(len2(sub(p[0], p[1])));
  } else if (kind == 1) {
    D3 u = sub(p[2], p[1]), w = sub(p[0], p[1]);
    double q = dot(w, u) / dot(u, u);
    D3 dd = sub(w, scale(u, q));
    d = sqrt(len2(dd));
  } else if (kind == 2) {
    D3 n = cross(sub(p[2], p[1]), sub(p[3], p[1]));
    double h = dot(sub(p[0], p[1]), n) / sqrt(len2(n));
    d = h < 0.0 ? -h : h;
  } else {
    D3 n = cross(sub(p[1], p[0]), sub(p[3], p[2]));
    if (len2(n) < 1e-30) return 0.0;
    double h = dot(sub(p[0], p[2]), n) / sqrt(len2(n));
    d = h < 0.0 ? -h : h;
  }
  double e = activation(d, r, k);
  if (kind == 3) {
    D3 n = cross(sub(p[1], p[0]), sub(p[3], p[2]));
    double q = len2(n) / eps;
    if (q < 1.0) e = e * q * (2.0 - q);
  }
  return e;
}
}  // namespace ogm
