// vbd_math.hpp — per-vertex local energies, gradients and exact Hessians of
// the VBD local solve, written once for the device kernels and for a host
// build that the CPU tests compare against the f64 oracle.
//
// Terms (all with respect to the position of ONE vertex v, as in
// VbdSolver::local, src/vbd.rs:196-248):
//   membrane  — src/energy.rs:10-29, closed form in the vertex-relative
//               deformation vectors w = sum_a cw_a (x_a - x_v);
//   bending   — src/energy.rs:31-46, closed form with the exact second
//               derivative of the unit normals (not the Gauss-Newton part only);
//   stitch    — src/vbd.rs:217-227 (seams) and :237-246 (grab);
//   contacts  — src/ogc/potential.rs:8-65 via a 3-variable second-order jet,
//               the same forward-mode construction as src/math.rs:33-122.
// Positions are relative to a per-step origin (the garment centroid,
// src/gpu_vbd.rs:154-156) so f32 differences stay accurate.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CV_HD __host__ __device__ __forceinline__
#else
#define CV_HD inline
#endif

namespace cvm {

#if !defined(__HIPCC__)
using std::fabs;
using std::fmin;
using std::isfinite;
using std::log;
using std::sqrt;
#endif

template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
CV_HD V3<R> mk(R x, R y, R z) {
  return V3<R>{x, y, z};
}
template <class R>
CV_HD V3<R> operator+(V3<R> a, V3<R> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
CV_HD V3<R> operator-(V3<R> a, V3<R> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
CV_HD V3<R> operator-(V3<R> a) {
  return {-a.x, -a.y, -a.z};
}
template <class R>
CV_HD V3<R> operator*(V3<R> a, R s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class R>
CV_HD V3<R> operator*(R s, V3<R> a) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class R>
CV_HD R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
CV_HD V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Symmetric 3x3: xx xy xz yy yz zz.
template <class R>
struct S3 {
  R a[6];
};
template <class R>
CV_HD S3<R> s3_zero() {
  S3<R> s;
  for (int i = 0; i < 6; ++i) s.a[i] = R(0);
  return s;
}
template <class R>
CV_HD void s3_add_outer(S3<R>& s, V3<R> u, R c) {  // s += c u u^T
  s.a[0] += c * u.x * u.x;
  s.a[1] += c * u.x * u.y;
  s.a[2] += c * u.x * u.z;
  s.a[3] += c * u.y * u.y;
  s.a[4] += c * u.y * u.z;
  s.a[5] += c * u.z * u.z;
}
template <class R>
CV_HD void s3_add_diag(S3<R>& s, R c) {
  s.a[0] += c;
  s.a[3] += c;
  s.a[5] += c;
}

// Full 3x3 row-major for the bending Hessian algebra.
template <class R>
struct M3 {
  R m[3][3];
};
template <class R>
CV_HD M3<R> skew(V3<R> a) {  // skew(a) * d = a x d
  M3<R> s;
  s.m[0][0] = R(0);
  s.m[0][1] = -a.z;
  s.m[0][2] = a.y;
  s.m[1][0] = a.z;
  s.m[1][1] = R(0);
  s.m[1][2] = -a.x;
  s.m[2][0] = -a.y;
  s.m[2][1] = a.x;
  s.m[2][2] = R(0);
  return s;
}
template <class R>
CV_HD M3<R> mul(const M3<R>& a, const M3<R>& b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}
template <class R>
CV_HD M3<R> mul_tn(const M3<R>& a, const M3<R>& b) {  // a^T b
  M3<R> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[0][i] * b.m[0][j] + a.m[1][i] * b.m[1][j] + a.m[2][i] * b.m[2][j];
  return r;
}

// -------------------------------------------------------------- 3-var jet
// Second-order forward AD over the 3 coordinates of one vertex; symmetric
// Hessian storage.  Same product/chain rules as src/math.rs:44-106.
template <class R>
struct J3 {
  R v;
  R g[3];
  R h[6];
};
template <class R>
CV_HD J3<R> jc(R v) {
  J3<R> j;
  j.v = v;
  for (int i = 0; i < 3; ++i) j.g[i] = R(0);
  for (int i = 0; i < 6; ++i) j.h[i] = R(0);
  return j;
}
template <class R>
CV_HD J3<R> jvar(R v, int i) {
  J3<R> j = jc<R>(v);
  j.g[i] = R(1);
  return j;
}
#define CV_HIDX(i, j) ((i) == 0 ? (j) : ((i) == 1 ? 2 + (j) : 5))  // valid for i <= j
template <class R>
CV_HD J3<R> operator+(const J3<R>& a, const J3<R>& b) {
  J3<R> r;
  r.v = a.v + b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < 6; ++i) r.h[i] = a.h[i] + b.h[i];
  return r;
}
template <class R>
CV_HD J3<R> operator-(const J3<R>& a, const J3<R>& b) {
  J3<R> r;
  r.v = a.v - b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] - b.g[i];
  for (int i = 0; i < 6; ++i) r.h[i] = a.h[i] - b.h[i];
  return r;
}
template <class R>
CV_HD J3<R> operator-(const J3<R>& a) {
  J3<R> r;
  r.v = -a.v;
  for (int i = 0; i < 3; ++i) r.g[i] = -a.g[i];
  for (int i = 0; i < 6; ++i) r.h[i] = -a.h[i];
  return r;
}
template <class R>
CV_HD J3<R> operator*(const J3<R>& a, const J3<R>& b) {
  J3<R> r;
  r.v = a.v * b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  int k = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j, ++k)
      r.h[k] = a.h[k] * b.v + a.v * b.h[k] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
  return r;
}
template <class R>
CV_HD J3<R> operator*(const J3<R>& a, R s) {
  J3<R> r;
  r.v = a.v * s;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * s;
  for (int i = 0; i < 6; ++i) r.h[i] = a.h[i] * s;
  return r;
}
template <class R>
CV_HD J3<R> chain(const J3<R>& a, R value, R first, R second) {
  J3<R> r;
  r.v = value;
  for (int i = 0; i < 3; ++i) r.g[i] = first * a.g[i];
  int k = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j, ++k) r.h[k] = first * a.h[k] + second * a.g[i] * a.g[j];
  return r;
}
template <class R>
CV_HD J3<R> operator/(const J3<R>& a, const J3<R>& b) {
  R bv = b.v;
  return a * chain(b, R(1) / bv, R(-1) / (bv * bv), R(2) / (bv * bv * bv));
}
template <class R>
CV_HD J3<R> jsqrt(const J3<R>& a) {
  R r = sqrt(a.v);
  return chain(a, r, R(0.5) / r, R(-0.25) / (r * r * r));
}
template <class R>
CV_HD J3<R> jlog(const J3<R>& a) {
  return chain(a, R(log(a.v)), R(1) / a.v, R(-1) / (a.v * a.v));
}
template <class R>
struct JV {  // jet 3-vector
  J3<R> c[3];
};
template <class R>
CV_HD JV<R> jsub(const JV<R>& a, const JV<R>& b) {
  return {{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}};
}
template <class R>
CV_HD J3<R> jdot(const JV<R>& a, const JV<R>& b) {
  return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2];
}
template <class R>
CV_HD JV<R> jcross(const JV<R>& a, const JV<R>& b) {
  return {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2],
           a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
}
template <class R>
CV_HD JV<R> jpoint(V3<R> p, bool variable) {
  if (variable) return {{jvar<R>(p.x, 0), jvar<R>(p.y, 1), jvar<R>(p.z, 2)}};
  return {{jc<R>(p.x), jc<R>(p.y), jc<R>(p.z)}};
}

// Output of one term for vertex v: energy, gradient, Hessian (symmetric).
template <class R>
struct Deriv {
  R e;
  V3<R> g;
  S3<R> h;
};
template <class R>
CV_HD Deriv<R> deriv_zero() {
  Deriv<R> d;
  d.e = R(0);
  d.g = mk<R>(0, 0, 0);
  d.h = s3_zero<R>();
  return d;
}
template <class R>
CV_HD void deriv_add(Deriv<R>& a, const Deriv<R>& b) {
  a.e += b.e;
  a.g = a.g + b.g;
  for (int i = 0; i < 6; ++i) a.h.a[i] += b.h.a[i];
}

// ------------------------------------------------------------- membrane
// src/energy.rs:10-15 with derivatives of the smooth tension/compression split.
template <class R>
CV_HD void axial3(R s, R k, R c, R delta, R& e, R& e1, R& e2) {
  if (c == R(1) && delta > R(0)) {
    // Pure tension law: the smoothed-compression part is multiplied by
    // 1 - c = 0 and finite (r >= delta > 0), so these are the same values.
    e = R(0.5) * k * (c * s * s);
    e1 = k * (c * s);
    e2 = k * c;
    return;
  }
  R r = sqrt(s * s + delta * delta);
  R p = (s + r - delta) * R(0.5);
  R p1 = R(0.5) * (R(1) + s / r);
  R p2 = R(0.5) * delta * delta / (r * r * r);
  e = R(0.5) * k * (c * s * s + (R(1) - c) * p * p);
  e1 = k * (c * s + (R(1) - c) * p * p1);
  e2 = k * (c + (R(1) - c) * (p1 * p1 + p * p2));
}
template <class R>
CV_HD R axial_e(R s, R k, R c, R delta) {
  if (c == R(1) && delta > R(0)) return R(0.5) * k * (c * s * s);  // as in axial3
  R r = sqrt(s * s + delta * delta);
  R p = (s + r - delta) * R(0.5);
  return R(0.5) * k * (c * s * s + (R(1) - c) * p * p);
}
// Record layout (f[]): a1 a2 | b1 b2 | area | warp weft shear | c delta.
// a_i, b_i are the grain-rotated Dm^-1 coefficients of the two other vertices
// (w = a1 (x1 - xv) + a2 (x2 - xv), v = b1 (x1 - xv) + b2 (x2 - xv)).
template <class R, class F>
CV_HD R membrane_e(V3<R> xv, V3<R> x1, V3<R> x2, const F* f) {
  V3<R> e1 = x1 - xv, e2 = x2 - xv;
  V3<R> w = e1 * R(f[0]) + e2 * R(f[1]);
  V3<R> u = e1 * R(f[2]) + e2 * R(f[3]);
  R ew = (dot(w, w) - R(1)) * R(0.5);
  R ev = (dot(u, u) - R(1)) * R(0.5);
  R sh = dot(w, u);
  return R(f[4]) * (axial_e(ew, R(f[5]), R(f[8]), R(f[9])) + axial_e(ev, R(f[6]), R(f[8]), R(f[9])) +
                    R(0.5) * R(f[7]) * sh * sh);
}
template <class R, class F>
CV_HD Deriv<R> membrane_d(V3<R> xv, V3<R> x1, V3<R> x2, const F* f) {
  V3<R> e1 = x1 - xv, e2 = x2 - xv;
  V3<R> w = e1 * R(f[0]) + e2 * R(f[1]);
  V3<R> u = e1 * R(f[2]) + e2 * R(f[3]);
  R cw = -(R(f[0]) + R(f[1]));
  R cu = -(R(f[2]) + R(f[3]));
  R ew = (dot(w, w) - R(1)) * R(0.5);
  R ev = (dot(u, u) - R(1)) * R(0.5);
  R sh = dot(w, u);
  R area = R(f[4]), ks = R(f[7]);
  R aw, aw1, aw2, av, av1, av2;
  axial3(ew, R(f[5]), R(f[8]), R(f[9]), aw, aw1, aw2);
  axial3(ev, R(f[6]), R(f[8]), R(f[9]), av, av1, av2);
  Deriv<R> d;
  d.e = area * (aw + av + R(0.5) * ks * sh * sh);
  V3<R> gs = u * cw + w * cu;
  d.g = (w * (aw1 * cw) + u * (av1 * cu) + gs * (ks * sh)) * area;
  d.h = s3_zero<R>();
  s3_add_outer(d.h, w, area * aw2 * cw * cw);
  s3_add_outer(d.h, u, area * av2 * cu * cu);
  s3_add_outer(d.h, gs, area * ks);
  s3_add_diag(d.h, area * (aw1 * cw * cw + av1 * cu * cu + R(2) * ks * sh * cw * cu));
  return d;
}

// --------------------------------------------------------------- bending
// E = 1/2 kappa |n1/|n1| - n2/|n2||^2, n1 = (x1-x0)x(x2-x0), n2 = (x3-x0)x(x1-x0).
// f[0] = kappa = weight (bend_warp aw^2 + bend_weft af^2) (src/energy.rs:40-45).
template <class R, class F>
CV_HD R bending_e(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, const F* f) {
  V3<R> e = x1 - x0;
  V3<R> n1 = cross(e, x2 - x0);
  V3<R> n2 = cross(x3 - x0, e);
  R l1 = sqrt(dot(n1, n1)), l2 = sqrt(dot(n2, n2));
  V3<R> d = n1 * (R(1) / l1) - n2 * (R(1) / l2);
  return R(0.5) * R(f[0]) * dot(d, d);
}
template <class R, class F>
CV_HD Deriv<R> bending_d(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, int slot, const F* f) {
  R kappa = R(f[0]);
  V3<R> e = x1 - x0, u = x2 - x0, w = x3 - x0;
  V3<R> n1 = cross(e, u);
  V3<R> n2 = cross(w, e);
  R l1 = sqrt(dot(n1, n1)), l2 = sqrt(dot(n2, n2));
  // One division per normal; the Jacobian and curvature reuse the reciprocals.
  R il1 = R(1) / l1, il2 = R(1) / l2;
  V3<R> a = n1 * il1, b = n2 * il2;
  V3<R> D = a - b;
  // dn1 = a1 x delta, dn2 = a2 x delta for the moving slot.
  V3<R> a1, a2;
  V3<R> z = mk<R>(0, 0, 0);
  if (slot == 0) {
    a1 = u - e;
    a2 = e - w;
  } else if (slot == 1) {
    a1 = -u;
    a2 = w;
  } else if (slot == 2) {
    a1 = e;
    a2 = z;
  } else {
    a1 = z;
    a2 = -e;
  }
  // J = P1 [a1]x / l1 - P2 [a2]x / l2 with P = I - n n^T.  Since
  // n^T [a1]x = (n x a1)^T, P1 [a1]x = [a1]x - a c1^T with c1 = a x a1.
  V3<R> c1 = cross(a, a1), c2 = cross(b, a2);
  R ax[3] = {a.x, a.y, a.z}, bx[3] = {b.x, b.y, b.z};
  R c1x[3] = {c1.x, c1.y, c1.z}, c2x[3] = {c2.x, c2.y, c2.z};
  R s1[3][3] = {{R(0), -a1.z, a1.y}, {a1.z, R(0), -a1.x}, {-a1.y, a1.x, R(0)}};
  R s2[3][3] = {{R(0), -a2.z, a2.y}, {a2.z, R(0), -a2.x}, {-a2.y, a2.x, R(0)}};
  R J[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) J[i][j] = (s1[i][j] - ax[i] * c1x[j]) * il1 - (s2[i][j] - bx[i] * c2x[j]) * il2;
  Deriv<R> d;
  d.e = R(0.5) * kappa * dot(D, D);
  R Dn[3] = {D.x, D.y, D.z};
  R g[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) g[j] = kappa * (J[0][j] * Dn[0] + J[1][j] * Dn[1] + J[2][j] * Dn[2]);
  d.g = mk<R>(g[0], g[1], g[2]);
  // Curvature part [a]x^T T [a]x for T = (3 yn n n^T - y n^T - n y^T - yn I) / l^2,
  // using [a]x^T u = u x a:  = (3 yn c c^T - q c^T - c q^T - yn (|a|^2 I - a a^T)) / l^2
  // with c = n x a (c1, c2 above), q = D x a, yn = D . n.
  V3<R> q1 = cross(D, a1), q2 = cross(D, a2);
  R q1x[3] = {q1.x, q1.y, q1.z}, q2x[3] = {q2.x, q2.y, q2.z};
  R a1x[3] = {a1.x, a1.y, a1.z}, a2x[3] = {a2.x, a2.y, a2.z};
  R yn1 = dot(D, a), yn2 = dot(D, b);
  R n1a = dot(a1, a1), n2a = dot(a2, a2);
  R k1 = il1 * il1, k2 = il2 * il2;
  int k = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j, ++k) {
      R jj = J[0][i] * J[0][j] + J[1][i] * J[1][j] + J[2][i] * J[2][j];
      R e1 = R(3) * yn1 * c1x[i] * c1x[j] - q1x[i] * c1x[j] - c1x[i] * q1x[j] -
             yn1 * ((i == j ? n1a : R(0)) - a1x[i] * a1x[j]);
      R e2 = R(3) * yn2 * c2x[i] * c2x[j] - q2x[i] * c2x[j] - c2x[i] * q2x[j] -
             yn2 * ((i == j ? n2a : R(0)) - a2x[i] * a2x[j]);
      d.h.a[k] = kappa * (jj + e1 * k1 - e2 * k2);
    }
  return d;
}

// ---------------------------------------------------------------- stitch
// E = 1/2 k |d|^2 with d = rel + w_v (x_v - anchor) where the caller formed
// rel = sum_{j != v} w_j (x_j - x_v) (+ (x_v - target) for a grab).
template <class R>
CV_HD Deriv<R> stitch_d(V3<R> d, R wv, R k) {
  Deriv<R> o;
  o.e = R(0.5) * k * dot(d, d);
  o.g = d * (k * wv);
  o.h = s3_zero<R>();
  s3_add_diag(o.h, k * wv * wv);
  return o;
}

// -------------------------------------------------------------- contacts
// Kinds follow src/ogc/potential.rs:171-177.
enum { KIND_PP = 0, KIND_PE = 1, KIND_PF = 2, KIND_EE = 3 };

// src/ogc/potential.rs:8-22 on a jet or a plain value.
template <class R>
CV_HD J3<R> activation_j(const J3<R>& d, R r, R k) {
  if (d.v >= r) return jc<R>(R(0));
  if (d.v <= R(0)) return jc<R>(R(INFINITY));
  R tau = R(0.5) * r;
  if (d.v >= tau) {
    J3<R> g = jc<R>(r) - d;
    return (jc<R>(R(0.5) * k) * g) * g;
  }
  return jc<R>(R(0.5) * k * tau * tau) - jc<R>(k * tau * tau) * jlog(d / jc<R>(tau));
}
template <class R>
CV_HD R activation_v(R d, R r, R k) {
  if (d >= r) return R(0);
  if (d <= R(0)) return R(INFINITY);
  R tau = R(0.5) * r;
  if (d >= tau) {
    R g = r - d;
    return R(0.5) * k * g * g;
  }
  return R(0.5) * k * tau * tau - k * tau * tau * R(log(d / tau));
}
// Contact::energy (src/ogc/potential.rs:186-212) with vertex `slot` as the variable.
template <class R>
CV_HD J3<R> contact_j(int kind, const V3<R>* p, int slot, R eps, R r, R k) {
  JV<R> q[4];
  for (int i = 0; i < 4; ++i) q[i] = jpoint(p[i], i == slot);
  J3<R> d;
  if (kind == KIND_PP) {
    JV<R> v = jsub(q[0], q[1]);
    d = jsqrt(jdot(v, v));
  } else if (kind == KIND_PE) {
    JV<R> u = jsub(q[2], q[1]), w = jsub(q[0], q[1]);
    J3<R> s = jdot(w, u) / jdot(u, u);
    JV<R> dd = {{w.c[0] - s * u.c[0], w.c[1] - s * u.c[1], w.c[2] - s * u.c[2]}};
    d = jsqrt(jdot(dd, dd));
  } else if (kind == KIND_PF) {
    JV<R> n = jcross(jsub(q[2], q[1]), jsub(q[3], q[1]));
    J3<R> h = jdot(jsub(q[0], q[1]), n) / jsqrt(jdot(n, n));
    d = h.v < R(0) ? -h : h;
  } else {
    JV<R> n = jcross(jsub(q[1], q[0]), jsub(q[3], q[2]));
    J3<R> nn = jdot(n, n);
    if (nn.v < R(1e-30)) return jc<R>(R(0));
    J3<R> h = jdot(jsub(q[0], q[2]), n) / jsqrt(nn);
    d = h.v < R(0) ? -h : h;
  }
  J3<R> e = activation_j(d, r, k);
  if (kind == KIND_EE) {
    JV<R> n = jcross(jsub(q[1], q[0]), jsub(q[3], q[2]));
    J3<R> qq = jdot(n, n) * (R(1) / eps);
    if (qq.v < R(1)) e = e * qq * (jc<R>(R(2)) - qq);
  }
  return e;
}
template <class R>
CV_HD R contact_e(int kind, const V3<R>* p, R eps, R r, R k) {
  R d;
  if (kind == KIND_PP) {
    V3<R> v = p[0] - p[1];
    d = sqrt(dot(v, v));
  } else if (kind == KIND_PE) {
    V3<R> u = p[2] - p[1], w = p[0] - p[1];
    R s = dot(w, u) / dot(u, u);
    V3<R> dd = w - u * s;
    d = sqrt(dot(dd, dd));
  } else if (kind == KIND_PF) {
    V3<R> n = cross(p[2] - p[1], p[3] - p[1]);
    R h = dot(p[0] - p[1], n) / sqrt(dot(n, n));
    d = h < R(0) ? -h : h;
  } else {
    V3<R> n = cross(p[1] - p[0], p[3] - p[2]);
    R nn = dot(n, n);
    if (nn < R(1e-30)) return R(0);
    R h = dot(p[0] - p[2], n) / sqrt(nn);
    d = h < R(0) ? -h : h;
  }
  R e = activation_v(d, r, k);
  if (kind == KIND_EE) {
    V3<R> n = cross(p[1] - p[0], p[3] - p[2]);
    R qq = dot(n, n) * (R(1) / eps);
    if (qq < R(1)) e = e * qq * (R(2) - qq);
  }
  return e;
}
template <class R>
CV_HD Deriv<R> contact_d(int kind, const V3<R>* p, int slot, R eps, R r, R k) {
  J3<R> j = contact_j(kind, p, slot, eps, r, k);
  Deriv<R> d;
  d.e = j.v;
  d.g = mk<R>(j.g[0], j.g[1], j.g[2]);
  for (int i = 0; i < 6; ++i) d.h.a[i] = j.h[i];
  return d;
}

// ------------------------------------------------------- term records
// One 64-byte record per (term, incident vertex), stored contiguously per
// vertex so the lanes serving a vertex read one coalesced run.
enum : uint32_t {
  TERM_MEMBRANE = 0,
  TERM_HINGE = 1,
  TERM_STITCH = 2,
  TERM_CONTACT = 3,
};
template <class F>
struct alignas(16) TermRecT {
  uint32_t head;     // kind | slot << 4 | count << 8 | contact kind << 12
  uint32_t ids[4];   // membrane: 2 others; hinge: x0..x3; stitch/contact: all vertices
  F f[11];
};
using TermRec = TermRecT<float>;     // f32 sweep: 64 bytes
using TermRec64 = TermRecT<double>;  // f64 sweep: 112 bytes
static_assert(sizeof(TermRec) == 64, "TermRec must be 64 bytes");
static_assert(sizeof(TermRec64) == 112, "TermRec64 must be 112 bytes");
CV_HD uint32_t rec_kind(uint32_t h) { return h & 0xFu; }
CV_HD int rec_slot(uint32_t h) { return (int)((h >> 4) & 0xFu); }
CV_HD int rec_count(uint32_t h) { return (int)((h >> 8) & 0xFu); }
CV_HD int rec_ckind(uint32_t h) { return (int)((h >> 12) & 0xFu); }
CV_HD uint32_t rec_head(uint32_t kind, int slot, int count, int ckind) {
  return kind | ((uint32_t)slot << 4) | ((uint32_t)count << 8) | ((uint32_t)ckind << 12);
}

// Position lookup abstraction: P(i) returns the relative position of vertex i.
template <class R, class F, class PosFn>
CV_HD R term_energy(const TermRecT<F>& t, V3<R> xv, R cr, R ck, PosFn&& P) {
  uint32_t kind = rec_kind(t.head);
  if (kind == TERM_MEMBRANE) return membrane_e<R, F>(xv, P(t.ids[0]), P(t.ids[1]), t.f);
  if (kind == TERM_HINGE) {
    int s = rec_slot(t.head);
    V3<R> x[4];
    for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(t.ids[i]);
    return bending_e<R, F>(x[0], x[1], x[2], x[3], t.f);
  }
  if (kind == TERM_STITCH) {
    // Constant indices only (fully unrolled): a record held in registers is
    // never indexed dynamically.
    int s = rec_slot(t.head), n = rec_count(t.head);
    V3<R> d = mk<R>(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < n && i != s) d = d + (P(t.ids[i]) - xv) * R(t.f[i]);
    return R(0.5) * R(t.f[4]) * dot(d, d);
  }
  int s = rec_slot(t.head);
  V3<R> x[4];
  for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(t.ids[i]);
  return contact_e<R>(rec_ckind(t.head), x, R(t.f[0]), cr, ck);
}
template <class R, class F, class PosFn>
CV_HD Deriv<R> term_deriv(const TermRecT<F>& t, V3<R> xv, R cr, R ck, PosFn&& P) {
  uint32_t kind = rec_kind(t.head);
  if (kind == TERM_MEMBRANE) return membrane_d<R, F>(xv, P(t.ids[0]), P(t.ids[1]), t.f);
  int s = rec_slot(t.head);
  if (kind == TERM_HINGE) {
    V3<R> x[4];
    for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(t.ids[i]);
    return bending_d<R, F>(x[0], x[1], x[2], x[3], s, t.f);
  }
  if (kind == TERM_STITCH) {
    int n = rec_count(t.head);
    V3<R> d = mk<R>(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < n && i != s) d = d + (P(t.ids[i]) - xv) * R(t.f[i]);
    R ws = s == 0 ? R(t.f[0]) : s == 1 ? R(t.f[1]) : s == 2 ? R(t.f[2]) : R(t.f[3]);
    return stitch_d<R>(d, ws, R(t.f[4]));
  }
  V3<R> x[4];
  for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(t.ids[i]);
  return contact_d<R>(rec_ckind(t.head), x, s, R(t.f[0]), cr, ck);
}

// ------------------------------------------------------ local Newton
// src/vbd.rs:353-371: inertia, Gershgorin shift and the cofactor solve.
template <class R>
CV_HD V3<R> newton_direction(Deriv<R>& tot, V3<R> x, V3<R> target, R inertia) {
  V3<R> dx = x - target;
  tot.g = tot.g + dx * inertia;
  s3_add_diag(tot.h, inertia);
  R* h = tot.h.a;
  R l0 = h[0] - (fabs(h[1]) + fabs(h[2]));
  R l1 = h[3] - (fabs(h[1]) + fabs(h[4]));
  R l2 = h[5] - (fabs(h[2]) + fabs(h[4]));
  R lower = fmin(l0, fmin(l1, l2));
  if (lower < R(1e-8)) s3_add_diag(tot.h, R(1e-8) - lower + R(0.01) * inertia);
  // Symmetric cofactor inverse.
  R a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5];
  R c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
  R c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
  R det = a * c00 + b * c01 + c * c02;
  R inv = R(1) / det;
  V3<R> g = tot.g;
  return mk<R>(-(c00 * g.x + c01 * g.y + c02 * g.z) * inv, -(c01 * g.x + c11 * g.y + c12 * g.z) * inv,
               -(c02 * g.x + c12 * g.y + c22 * g.z) * inv);
}
// ContactSet::truncate (src/ogc/world.rs:91-100) against a ball (center, radius).
template <class R>
CV_HD V3<R> truncate_ball(V3<R> p, V3<R> c, R radius, bool& clipped) {
  V3<R> d = p - c;
  R len = sqrt(dot(d, d));
  if (len > radius) {
    clipped = true;
    return c + d * (radius / len);
  }
  return p;
}
template <class R>
CV_HD bool finite3(V3<R> p) {
  return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}

}  // namespace cvm
