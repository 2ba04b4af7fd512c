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
//   contacts  — src/ogc/potential.rs:8-65, closed-form derivatives of the
//               four distance kinds and the EE mollifier (the reference
//               differentiates them with Jet<3>, src/math.rs:33-122).
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
template <class R>
CV_HD void s3_add_sym(S3<R>& s, V3<R> x, V3<R> y, R c) {  // s += c (x y^T + y x^T)
  s.a[0] += c * (R(2) * x.x * y.x);
  s.a[1] += c * (x.x * y.y + y.x * x.y);
  s.a[2] += c * (x.x * y.z + y.x * x.z);
  s.a[3] += c * (R(2) * x.y * y.y);
  s.a[4] += c * (x.y * y.z + y.y * x.z);
  s.a[5] += c * (R(2) * x.z * y.z);
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
// Reciprocals, quotients and square roots of the term math.  f64 keeps IEEE
// division and square root (the reference's rounding).  The f32 sweep uses
// the hardware reciprocal and square root (1 ulp each) on the device, a
// fraction of the instructions of the correctly rounded sequences; its
// parity bounds are set by the reference step's conditioning, far above an
// ulp.  The host build (CPU tests) divides exactly.
template <class R>
CV_HD R recip(R x) {
  return R(1) / x;
}
template <class R>
CV_HD R quot(R a, R b) {
  return a / b;
}
template <class R>
CV_HD R root(R x) {
  return sqrt(x);
}
#if defined(__HIP_DEVICE_COMPILE__)
template <>
CV_HD float recip<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <>
CV_HD float quot<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <>
CV_HD float root<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
#endif

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
  R r = root(s * s + delta * delta);
  R p = (s + r - delta) * R(0.5);
  R p1 = R(0.5) * (R(1) + quot(s, r));
  R p2 = quot(R(0.5) * delta * delta, r * r * r);
  e = R(0.5) * k * (c * s * s + (R(1) - c) * p * p);
  e1 = k * (c * s + (R(1) - c) * p * p1);
  e2 = k * (c + (R(1) - c) * (p1 * p1 + p * p2));
}
template <class R>
CV_HD R axial_e(R s, R k, R c, R delta) {
  if (c == R(1) && delta > R(0)) return R(0.5) * k * (c * s * s);  // as in axial3
  R r = root(s * s + delta * delta);
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
  R l1 = root(dot(n1, n1)), l2 = root(dot(n2, n2));
  V3<R> d = n1 * recip(l1) - n2 * recip(l2);
  return R(0.5) * R(f[0]) * dot(d, d);
}
template <class R, class F>
CV_HD Deriv<R> bending_d(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, int slot, const F* f) {
  R kappa = R(f[0]);
  V3<R> e = x1 - x0, u = x2 - x0, w = x3 - x0;
  V3<R> n1 = cross(e, u);
  V3<R> n2 = cross(w, e);
  R l1 = root(dot(n1, n1)), l2 = root(dot(n2, n2));
  // One division per normal; the Jacobian and curvature reuse the reciprocals.
  R il1 = recip(l1), il2 = recip(l2);
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
// Kinds follow src/ogc/potential.rs:25-30.
enum { KIND_PP = 0, KIND_PE = 1, KIND_PF = 2, KIND_EE = 3 };

// src/ogc/potential.rs:8-22.
template <class R>
CV_HD R activation_v(R d, R r, R k) {
  if (d >= r) return R(0);
  if (d <= R(0)) return R(INFINITY);
  R tau = R(0.5) * r;
  if (d >= tau) {
    R g = r - d;
    return R(0.5) * k * g * g;
  }
  return R(0.5) * k * tau * tau - k * tau * tau * R(log(quot(d, tau)));
}
// Contact::energy (src/ogc/potential.rs:39-65).
template <class R>
CV_HD R contact_e(int kind, const V3<R>* p, R eps, R r, R k) {
  R d, nn = R(0);
  if (kind == KIND_PP) {
    V3<R> v = p[0] - p[1];
    d = root(dot(v, v));
  } else if (kind == KIND_PE) {
    V3<R> u = p[2] - p[1], w = p[0] - p[1];
    R s = quot(dot(w, u), dot(u, u));
    V3<R> dd = w - u * s;
    d = root(dot(dd, dd));
  } else {
    // PF: n = (p2 - p1) x (p3 - p1), r = p0 - p1; EE: n = (p1 - p0) x (p3 - p2),
    // r = p0 - p2 (one code path for both: a wavefront holding both kinds
    // runs it once).
    const bool pf = kind == KIND_PF;
    V3<R> n = cross(pf ? p[2] - p[1] : p[1] - p[0], pf ? p[3] - p[1] : p[3] - p[2]);
    nn = dot(n, n);
    if (!pf && nn < R(1e-30)) return R(0);
    R h = quot(dot(pf ? p[0] - p[1] : p[0] - p[2], n), root(nn));
    d = h < R(0) ? -h : h;
  }
  R e = activation_v(d, r, k);
  if (kind == KIND_EE) {
    R qq = nn * recip(eps);
    if (qq < R(1)) e = e * qq * (R(2) - qq);
  }
  return e;
}
// First and second derivatives of the activation at d in (0, r).
template <class R>
CV_HD void activation_d(R d, R r, R k, R& e1, R& e2) {
  R tau = R(0.5) * r;
  if (d >= tau) {
    e1 = -k * (r - d);
    e2 = k;
  } else {
    R kt = k * tau * tau;
    e1 = -quot(kt, d);
    e2 = quot(kt, d * d);
  }
}
// Unsigned plane distance d = |r . n| / |n| and its derivatives in one
// vertex x, where that vertex moves r by c dx and the normal by a x dx.
// With m = |n|, n^ = n / m, h = r . n^, rho = r - h n^ (the in-plane part of
// r) and c_a = n^ x a:
//   grad h = c n^ + (rho x a) / m
//   hess h = -(c / m)(n^ c_a^T + c_a n^T) - (c_a w^T + w c_a^T) / m^2
//            - (h / m^2)(|a|^2 I - a a^T - c_a c_a^T),   w = rho x a,
// from dn^ = (I - n^ n^T) dn / m and dm = c_a . dx; d = sign(h) h.
template <class R>
CV_HD void plane_dist_d(V3<R> r, V3<R> n, R nn, V3<R> a, R c, R& d, V3<R>& g, S3<R>& H) {
  R m = root(nn);
  R h = quot(dot(r, n), m);  // as contact_e
  R sg = h < R(0) ? R(-1) : R(1);
  d = h < R(0) ? -h : h;
  R im = recip(m);
  V3<R> nh = n * im;
  V3<R> ca = cross(nh, a);
  V3<R> w = cross(r - nh * h, a);
  g = (nh * c + w * im) * sg;
  H = s3_zero<R>();
  R im2 = im * im;
  s3_add_sym(H, nh, ca, -sg * c * im);
  s3_add_sym(H, ca, w, -sg * im2);
  R q = -sg * h * im2;
  s3_add_diag(H, q * dot(a, a));
  s3_add_outer(H, a, -q);
  s3_add_outer(H, ca, -q);
}
// Contact energy of vertex `slot` with its gradient and Hessian in closed
// form: E = phi(d) (times the EE mollifier M(q) = q (2 - q) while q < 1), so
// grad E = phi' grad d (+ phi M' grad q) and hess E = phi'' grad d grad d^T +
// phi' hess d (+ the mollifier's product-rule terms), with the distance
// derivatives of each kind:
//   PP  d = |p0 - p1|:             grad = +-u^, hess = (I - u^ u^T) / d;
//   PE  distance of p0 to the line p1 p2, t = w - s u (w = p0 - p1,
//       u = p2 - p1, s = w.u / u.u), t^ = t / d; for the point
//       grad = t^, hess = (I - u u^T/u.u - t^ t^T) / d; for an edge end with
//       weight b (1 - s at p1, s at p2) grad = -b t^ and
//       hess = b^2 (I - u u^T/u.u - t^ t^T) / d + g (t^ u^T + u t^T) / u.u
//              - d t^ t^T / u.u, g = -(1 - s) at p1 and s at p2;
//   PF, EE  plane distances (plane_dist_d) with (a, c) of the moving vertex;
//   EE  q = |n|^2 / eps: grad q = 2 (n x a) / eps, hess q = 2 (|a|^2 I - a a^T) / eps.
// Values follow contact_e's operations, so the energy is the one the line
// search evaluates.  Degenerate cases as the energy: zero beyond r and for
// parallel edges, infinite energy (no derivatives) at d <= 0.
template <class R>
CV_HD Deriv<R> contact_d(int kind, const V3<R>* p, int slot, R eps, R r, R k) {
  Deriv<R> o = deriv_zero<R>();
  R d;
  V3<R> gd;
  S3<R> hd = s3_zero<R>();
  V3<R> n = mk<R>(0, 0, 0), a = mk<R>(0, 0, 0);
  R nn = R(0);
  if (kind == KIND_PP) {
    V3<R> v = p[0] - p[1];
    d = root(dot(v, v));
    R id = recip(d);
    V3<R> uh = v * id;
    gd = slot == 0 ? uh : -uh;
    s3_add_diag(hd, id);
    s3_add_outer(hd, uh, -id);
  } else if (kind == KIND_PE) {
    V3<R> u = p[2] - p[1], w = p[0] - p[1];
    R uu = dot(u, u);
    R s = quot(dot(w, u), uu);
    V3<R> t = w - u * s;
    d = root(dot(t, t));
    R id = recip(d);
    V3<R> th = t * id;
    R b = slot == 0 ? R(1) : (slot == 1 ? R(1) - s : s);
    gd = slot == 0 ? th : th * (-b);
    R bb = b * b * id;
    s3_add_diag(hd, bb);
    s3_add_outer(hd, u, -quot(bb, uu));
    s3_add_outer(hd, th, -bb);
    if (slot != 0) {
      R gam = slot == 1 ? s - R(1) : s;
      s3_add_sym(hd, th, u, quot(gam, uu));
      s3_add_outer(hd, th, -quot(d, uu));
    }
  } else {  // PF and EE share one path (see contact_e)
    const bool pf = kind == KIND_PF;
    n = cross(pf ? p[2] - p[1] : p[1] - p[0], pf ? p[3] - p[1] : p[3] - p[2]);
    nn = dot(n, n);
    if (!pf && nn < R(1e-30)) return o;
    R c;
    if (pf) {
      c = slot == 0 ? R(1) : (slot == 1 ? R(-1) : R(0));
      a = slot == 1 ? p[3] - p[2] : (slot == 2 ? p[1] - p[3] : (slot == 3 ? p[2] - p[1] : a));
    } else {
      c = slot == 0 ? R(1) : (slot == 2 ? R(-1) : R(0));
      a = slot == 0 ? p[3] - p[2] : (slot == 1 ? p[2] - p[3] : (slot == 2 ? p[0] - p[1] : p[1] - p[0]));
    }
    plane_dist_d(pf ? p[0] - p[1] : p[0] - p[2], n, nn, a, c, d, gd, hd);
  }
  R phi = activation_v(d, r, k);
  R q = kind == KIND_EE ? nn * recip(eps) : R(1);
  bool moll = kind == KIND_EE && q < R(1);
  o.e = moll ? phi * q * (R(2) - q) : phi;
  if (d >= r) return o;
  if (d <= R(0)) {
    if (moll) {  // infinite energy times the mollifier's derivatives
      R nan = R(NAN);
      o.g = mk<R>(nan, nan, nan);
      for (int i = 0; i < 6; ++i) o.h.a[i] = nan;
    }
    return o;
  }
  R e1, e2;
  activation_d(d, r, k, e1, e2);
  if (!moll) {
    o.g = gd * e1;
    s3_add_outer(o.h, gd, e2);
    for (int i = 0; i < 6; ++i) o.h.a[i] += e1 * hd.a[i];
    return o;
  }
  R M = q * (R(2) - q), M1 = R(2) - R(2) * q;
  R te = quot(R(2), eps);
  V3<R> gq = cross(n, a) * te;
  o.g = gd * (e1 * M) + gq * (phi * M1);
  s3_add_outer(o.h, gd, e2 * M);
  for (int i = 0; i < 6; ++i) o.h.a[i] += (e1 * M) * hd.a[i];
  s3_add_sym(o.h, gd, gq, e1 * M1);
  s3_add_outer(o.h, gq, R(-2) * phi);
  R pm = phi * M1 * te;
  s3_add_diag(o.h, pm * dot(a, a));
  s3_add_outer(o.h, a, -pm);
  return o;
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

// Position lookup abstraction: P(k, i) returns the relative position of vertex
// i = t.ids[k], the record's k-th vertex (k is a compile-time constant at
// every call, so a caller may serve it from registers).
template <class R, class F, class PosFn>
CV_HD R term_energy(const TermRecT<F>& t, V3<R> xv, R cr, R ck, PosFn&& P) {
  uint32_t kind = rec_kind(t.head);
  if (kind == TERM_MEMBRANE) return membrane_e<R, F>(xv, P(0, t.ids[0]), P(1, t.ids[1]), t.f);
  if (kind == TERM_HINGE) {
    int s = rec_slot(t.head);
    V3<R> x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(i, t.ids[i]);
    return bending_e<R, F>(x[0], x[1], x[2], x[3], t.f);
  }
  if (kind == TERM_STITCH) {
    // Constant indices only (fully unrolled): a record held in registers is
    // never indexed dynamically.
    int s = rec_slot(t.head), n = rec_count(t.head);
    V3<R> d = mk<R>(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < n && i != s) d = d + (P(i, t.ids[i]) - xv) * R(t.f[i]);
    return R(0.5) * R(t.f[4]) * dot(d, d);
  }
  int s = rec_slot(t.head);
  V3<R> x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(i, t.ids[i]);
  return contact_e<R>(rec_ckind(t.head), x, R(t.f[0]), cr, ck);
}
template <class R, class F, class PosFn>
CV_HD Deriv<R> term_deriv(const TermRecT<F>& t, V3<R> xv, R cr, R ck, PosFn&& P) {
  uint32_t kind = rec_kind(t.head);
  if (kind == TERM_MEMBRANE) return membrane_d<R, F>(xv, P(0, t.ids[0]), P(1, t.ids[1]), t.f);
  int s = rec_slot(t.head);
  if (kind == TERM_HINGE) {
    V3<R> x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(i, t.ids[i]);
    return bending_d<R, F>(x[0], x[1], x[2], x[3], s, t.f);
  }
  if (kind == TERM_STITCH) {
    int n = rec_count(t.head);
    V3<R> d = mk<R>(0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < n && i != s) d = d + (P(i, t.ids[i]) - xv) * R(t.f[i]);
    R ws = s == 0 ? R(t.f[0]) : s == 1 ? R(t.f[1]) : s == 2 ? R(t.f[2]) : R(t.f[3]);
    return stitch_d<R>(d, ws, R(t.f[4]));
  }
  V3<R> x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = i == s ? xv : P(i, t.ids[i]);
  return contact_d<R>(rec_ckind(t.head), x, s, R(t.f[0]), cr, ck);
}

// Energy changes for the f32 line search.  A stitch (or grab) energy
// 1/2 k |d|^2 is linear in the vertex through d = rel + w_v x_v, so moving
// the vertex by dx changes it by k w_v (d . dx) + 1/2 k w_v^2 |dx|^2 exactly:
// no difference of two large energies (a seam still open by centimetres
// holds ~1e2 J against ~1e-7 J of membrane energy, far beyond f32's reach).
template <class R>
CV_HD R quad_change(V3<R> d, R wv, R k, V3<R> dx) {
  return k * wv * dot(d, dx) + R(0.5) * k * wv * wv * dot(dx, dx);
}
// A record's contribution to the change of the local energy from xv to xc:
// the exact change for a stitch; for the other kinds the energy at xc (the
// caller subtracts their energies at xv, which it already holds).
template <class R, class F, class PosFn>
CV_HD R term_change_part(const TermRecT<F>& t, V3<R> xv, V3<R> xc, R cr, R ck, PosFn&& P) {
  if (rec_kind(t.head) != TERM_STITCH) return term_energy<R>(t, xc, cr, ck, P);
  int s = rec_slot(t.head), n = rec_count(t.head);
  V3<R> d = mk<R>(0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < n && i != s) d = d + (P(i, t.ids[i]) - xv) * R(t.f[i]);
  R ws = s == 0 ? R(t.f[0]) : s == 1 ? R(t.f[1]) : s == 2 ? R(t.f[2]) : R(t.f[3]);
  return quad_change<R>(d, ws, R(t.f[4]), xc - xv);
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
  R inv = recip(det);
  V3<R> g = tot.g;
  return mk<R>(-(c00 * g.x + c01 * g.y + c02 * g.z) * inv, -(c01 * g.x + c11 * g.y + c12 * g.z) * inv,
               -(c02 * g.x + c12 * g.y + c22 * g.z) * inv);
}
// Slack of the line search's accept test `E <= E_old + slack`
// (src/vbd.rs:375-386).  In f64 the reference's 1e-14.  In f32 that absolute
// slack lies below one ulp of a local energy, so candidates whose energy
// differs from E_old only by rounding would be rejected and converged
// vertices would run all ten halvings; the f32 slack adds four ulps of E_old
// (local energies are sums of non-negative terms), the rounding level of an
// f32 sum, as the reference's slack sits above the rounding level of f64.
template <class R>
CV_HD R accept_slack(R old);
template <>
CV_HD double accept_slack<double>(double) {
  return 1e-14;
}
template <>
CV_HD float accept_slack<float>(float old) {
  return 1e-14f + 4.0f * 1.1920928955078125e-07f * fabs(old);
}

// ContactSet::truncate (src/ogc/world.rs:91-100) against a ball (center, radius).
template <class R>
CV_HD V3<R> truncate_ball(V3<R> p, V3<R> c, R radius, bool& clipped) {
  V3<R> d = p - c;
  R len = root(dot(d, d));
  if (len > radius) {
    clipped = true;
    return c + d * quot(radius, len);
  }
  return p;
}
template <class R>
CV_HD bool finite3(V3<R> p) {
  return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}

}  // namespace cvm
