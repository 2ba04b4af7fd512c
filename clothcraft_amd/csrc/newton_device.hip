// newton_device.hip — the Newton energy reference of one step on the device:
// ReferenceSolver::step (src/solver.rs:422-647) with its System/pcg/rms
// (src/solver.rs:142-231), in f64, over the objective of src/solver.rs:243-296
// (inertial form, membrane, bending, seams) plus the OGC contacts of the
// handle's last detection, frozen (the reference Newton has no OGC term; the
// contacts make its objective the one a contact step minimises).
//
// Matrix-free on the term records the sweep already holds:
//  * gradient and the Jacobi diagonal of each vertex from the closed-form
//    per-vertex term derivatives (exact), plus the inertial term;
//  * the objective with every term counted once, at its smallest garment
//    vertex (every garment vertex of a term carries a record of it), so the
//    value is the reference objective's up to summation order; a collapsed
//    triangle makes it infinite as in src/solver.rs:235-242;
//  * Hessian-vector products H p as central differences of that exact
//    gradient, (g(x + h p) - g(x - h p)) / 2h, with h p at most 1e-7 of the
//    scene scale (truncation (h p / d)^2 for contact distances d >= 1e-5 of the
//    scale, round-off ~1e-9 relative).  The converged point is fixed by the
//    exact gradient, so the inexact products change the iteration count, not
//    the minimiser.
// The Newton loop, PCG, regularisation shifts and Armijo backtracking are the
// reference's; reductions run as fixed-size per-block partials reduced in a
// fixed order (deterministic), and the PCG scalars stay on the device with
// the host reading the PCG state once per batch of iterations.
#include <cmath>
#include <vector>

#include "gpu.hpp"

using cvm::V3;

namespace {

constexpr int NT_TPB = 256;
constexpr int NT_L = 32;           // lanes per vertex in the gradient pass
constexpr uint32_t NT_PARTS = 256; // blocks of the reductions
constexpr double NT_DELTA = 1e-7;  // largest finite-difference displacement / scene scale

struct NtArgs {
  const uint32_t* term_off;
  const cvm::TermRec64* terms;
  const uint32_t* coff;
  const cvm::TermRec64* crec;
  const double* x0;  // step start positions [3n]
  const double* v0;  // step start velocities [3n]
  const double* masses;
  uint32_t n;
  double dt, damping, gx, gy, gz, cr, ck;
};

__device__ __forceinline__ V3<double> ldx(const double4* X, uint32_t i) {
  double4 q = X[i];
  return V3<double>{q.x, q.y, q.z};
}

// The term's smallest garment vertex (v itself for a record of v's own term).
__device__ __forceinline__ uint32_t term_owner(const cvm::TermRec64& t, uint32_t v, uint32_t n) {
  const uint32_t kind = cvm::rec_kind(t.head);
  if (kind == cvm::TERM_MEMBRANE) return min(v, min(t.ids[0], t.ids[1]));
  const int cnt = kind == cvm::TERM_HINGE ? 4 : cvm::rec_count(t.head);
  uint32_t o = v;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q < cnt && t.ids[q] < n) o = min(o, t.ids[q]);
  return o;
}

template <int L>
__device__ __forceinline__ double lane_sum(double x) {
#pragma unroll
  for (int m = 1; m < L; m <<= 1) x += __shfl_xor(x, m, L);
  return x;
}

// Per garment vertex: gradient of the objective (0 on pinned vertices), the
// diagonal of its Hessian block plus the inertial mass term, and the vertex's
// share of the objective (its inertial term and the terms it owns).
__global__ __launch_bounds__(NT_TPB) void k_nt_eval(NtArgs a, const double4* X, double* grad, double* diag,
                                                    double* epart, const NtScal* sc) {
  if (sc && sc->state) return;  // a stopped PCG's queued difference passes
  const int lane = threadIdx.x & (NT_L - 1);
  const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / NT_L;
  if (v >= a.n) return;  // whole groups leave together
  const double4 p4 = X[v];
  const V3<double> x{p4.x, p4.y, p4.z};
  const bool fixed = p4.w > 0.5;
  const uint32_t s0 = a.term_off[v], ns = a.term_off[v + 1] - s0;
  const uint32_t c0 = a.coff[v], nt = ns + (a.coff[v + 1] - c0);
  auto P = [&](int, uint32_t i) { return ldx(X, i); };
  double e = 0, g0 = 0, g1 = 0, g2 = 0, h0 = 0, h1 = 0, h2 = 0;
  for (uint32_t i = lane; i < nt; i += NT_L) {
    const cvm::TermRec64& t = i < ns ? a.terms[s0 + i] : a.crec[c0 + (i - ns)];
    const cvm::Deriv<double> d = cvm::term_deriv<double>(t, x, a.cr, a.ck, P);
    g0 += d.g.x;
    g1 += d.g.y;
    g2 += d.g.z;
    h0 += d.h.a[0];
    h1 += d.h.a[3];
    h2 += d.h.a[5];
    if (term_owner(t, v, a.n) == v) {
      double te = d.e;
      if (cvm::rec_kind(t.head) == cvm::TERM_MEMBRANE) {
        // src/solver.rs:235-242: |(b - a) x (c - a)| > 2 area 1e-6
        const V3<double> b = ldx(X, t.ids[0]) - x, c = ldx(X, t.ids[1]) - x;
        const V3<double> n = cvm::cross(b, c);
        if (!(sqrt(cvm::dot(n, n)) > 2.0 * t.f[4] * 1e-6)) te = INFINITY;
      }
      e += te;
    }
  }
  e = lane_sum<NT_L>(e);
  g0 = lane_sum<NT_L>(g0);
  g1 = lane_sum<NT_L>(g1);
  g2 = lane_sum<NT_L>(g2);
  h0 = lane_sum<NT_L>(h0);
  h1 = lane_sum<NT_L>(h1);
  h2 = lane_sum<NT_L>(h2);
  if (lane != 0) return;
  // Inertial term of src/solver.rs:243-254 and its gradient (:560-570 form).
  const double m = a.masses[v], dt = a.dt;
  const double xs[3] = {x.x, x.y, x.z}, gr[3] = {a.gx, a.gy, a.gz};
  const double gs[3] = {g0, g1, g2}, hs[3] = {h0, h1, h2};
  double sy = 0, so = 0, sg = 0;
  const double md = m * (1.0 / (dt * dt) + a.damping / dt);
  if (!isfinite(x.x) || !isfinite(x.y) || !isfinite(x.z)) e = NAN;
  for (int k = 0; k < 3; ++k) {
    const double o = a.x0[3 * v + k], w = a.v0[3 * v + k];
    const double y = o + dt * w;
    sy += (xs[k] - y) * (xs[k] - y);
    so += (xs[k] - o) * (xs[k] - o);
    sg += gr[k] * xs[k];
    const double delta = m * ((xs[k] - o - dt * w) / (dt * dt) + a.damping / dt * (xs[k] - o) - gr[k]);
    grad[3 * v + k] = fixed ? 0.0 : gs[k] + delta;
    diag[3 * v + k] = md + hs[k];
  }
  epart[v] = e + m * (0.5 * sy / (dt * dt) + 0.5 * a.damping / dt * so - sg);
}

// Block partials of K values into parts[K * block + j]: sums, or maxima for
// the entries whose bit is set in MAXMASK.  Every thread of the block calls.
template <int K, unsigned MAXMASK = 0>
__device__ __forceinline__ void block_parts(const double (&v)[K], double* parts) {
  __shared__ double red[K][NT_TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double x = v[j];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = ((MAXMASK >> j) & 1u) ? fmax(x, y) : x + y;
    }
    if (lane == 0) red[j][w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      double s = red[j][0];
      for (int q = 1; q < NT_TPB / 64; ++q) s = ((MAXMASK >> j) & 1u) ? fmax(s, red[j][q]) : s + red[j][q];
      parts[K * blockIdx.x + j] = s;
    }
  }
}

#define NT_LOOP(i, m) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (m); i += gridDim.x * blockDim.x)

__device__ __forceinline__ double precond(const double* diag0, const double4* X, uint32_t i, double shift) {
  return X[i / 3].w > 0.5 ? 1.0 : fmax(diag0[i] + shift, 1e-12);
}

__global__ __launch_bounds__(NT_TPB) void k_nt_sum(const double* a, uint32_t m, double* parts) {
  double v[1] = {0};
  NT_LOOP(i, m) v[0] += a[i];
  block_parts<1>(v, parts);
}
__global__ __launch_bounds__(NT_TPB) void k_nt_dot(const double* a, const double* b, uint32_t m, double* parts) {
  double v[1] = {0};
  NT_LOOP(i, m) v[0] += a[i] * b[i];
  block_parts<1>(v, parts);
}
__global__ __launch_bounds__(NT_TPB) void k_nt_max_diag(const double* diag0, const double4* X, uint32_t m,
                                                        double* parts) {
  double v[1] = {0};
  NT_LOOP(i, m) v[0] = fmax(v[0], precond(diag0, X, i, 0.0));
  block_parts<1, 1u>(v, parts);
}
// PCG start: x = 0, r = -g, z = r / d, p = z; partials r.r, r.z, max |p|.
__global__ __launch_bounds__(NT_TPB) void k_nt_init(const double* g, const double* diag0, const double4* X,
                                                    uint32_t m, double shift, double* sol, double* r, double* z,
                                                    double* p, double* parts) {
  double v[3] = {0, 0, 0};
  NT_LOOP(i, m) {
    const double ri = -g[i], zi = ri / precond(diag0, X, i, shift);
    sol[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    v[0] += ri * ri;
    v[1] += ri * zi;
    v[2] = fmax(v[2], fabs(zi));
  }
  block_parts<3, 4u>(v, parts);
}
// Xt = X + s p on the free garment vertices (pinned ones and the avatar stay).
__global__ __launch_bounds__(NT_TPB) void k_nt_move(const double4* X, const double* p, double s, uint32_t n,
                                                    double4* Xt) {
  NT_LOOP(v, n) {
    const double4 q = X[v];
    Xt[v] = q.w > 0.5 ? q : make_double4(q.x + s * p[3 * v], q.y + s * p[3 * v + 1], q.z + s * p[3 * v + 2], 0.0);
  }
}
// The PCG kernels below read their scalars from NtScal and do nothing once
// its state is set, so the host can queue several iterations between
// read-backs.  The *_fin kernels (one block of NT_PARTS threads) reduce the
// block partials in a fixed tree order (deterministic) and thread 0 updates
// the scalars.
template <bool MAX>
__device__ __forceinline__ double parts_tree(const double* parts, int K, int j) {
  __shared__ double t[NT_PARTS];
  t[threadIdx.x] = parts[(size_t)K * threadIdx.x + j];
  __syncthreads();
  for (uint32_t s = NT_PARTS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) t[threadIdx.x] = MAX ? fmax(t[threadIdx.x], t[threadIdx.x + s]) : t[threadIdx.x] + t[threadIdx.x + s];
    __syncthreads();
  }
  const double r = t[0];
  __syncthreads();  // t is reused by the next call
  return r;
}
// Difference step h = delta / max|p| (delta = 1e-7 of the scene scale).
__device__ __forceinline__ void set_step(NtScal* sc, double pmax, double delta) {
  const double h = delta / (pmax > 0.0 ? pmax : 1.0);
  sc->h = h;
  sc->inv2h = 0.5 / h;
}
__global__ __launch_bounds__(NT_PARTS) void k_nt_init_fin(const double* parts, double delta, NtScal* sc) {
  const double rr = parts_tree<false>(parts, 3, 0), rz = parts_tree<false>(parts, 3, 1);
  const double pmax = parts_tree<true>(parts, 3, 2);
  if (threadIdx.x) return;
  sc->rz = rz;
  sc->tol = 1e-5 * sqrt(rr);
  set_step(sc, pmax, delta);
  sc->state = 0;
  sc->its = 0;
}
// Xt = X +- h p on the free garment vertices.
__global__ __launch_bounds__(NT_TPB) void k_nt_probe(const double4* X, const double* p, const NtScal* sc,
                                                     double sign, uint32_t n, double4* Xt) {
  if (sc->state) return;
  const double s = sign * sc->h;
  NT_LOOP(v, n) {
    const double4 q = X[v];
    Xt[v] = q.w > 0.5 ? q : make_double4(q.x + s * p[3 * v], q.y + s * p[3 * v + 1], q.z + s * p[3 * v + 2], 0.0);
  }
}
// A p = (g(x + h p) - g(x - h p)) / 2h + shift p (p on pinned degrees of
// freedom, where p is 0); partial p.Ap.
__global__ __launch_bounds__(NT_TPB) void k_nt_ap(const double* gp, const double* gm, const double* p,
                                                  const double4* X, const NtScal* sc, double shift, uint32_t m,
                                                  double* ap, double* parts) {
  if (sc->state) return;
  const double inv2h = sc->inv2h;
  double v[1] = {0};
  NT_LOOP(i, m) {
    const double a = X[i / 3].w > 0.5 ? p[i] : (gp[i] - gm[i]) * inv2h + shift * p[i];
    ap[i] = a;
    v[0] += p[i] * a;
  }
  block_parts<1>(v, parts);
}
__global__ __launch_bounds__(NT_PARTS) void k_nt_ap_fin(const double* parts, uint32_t it, NtScal* sc) {
  if (sc->state) return;
  const double curv = parts_tree<false>(parts, 1, 0);
  if (threadIdx.x) return;
  if (!isfinite(curv) || curv <= 0.0) {
    sc->state = 2;
    sc->its = it + 1;
    return;
  }
  sc->alpha = sc->rz / curv;
}
// x += alpha p, r -= alpha Ap, z = r / d; partials r.r, r.z.
__global__ __launch_bounds__(NT_TPB) void k_nt_update(const NtScal* sc, const double* p, const double* ap,
                                                      const double* diag0, const double4* X, double shift,
                                                      uint32_t m, double* sol, double* r, double* z, double* parts) {
  if (sc->state) return;
  const double alpha = sc->alpha;
  double v[2] = {0, 0};
  NT_LOOP(i, m) {
    sol[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i], zi = ri / precond(diag0, X, i, shift);
    r[i] = ri;
    z[i] = zi;
    v[0] += ri * ri;
    v[1] += ri * zi;
  }
  block_parts<2>(v, parts);
}
__global__ __launch_bounds__(NT_PARTS) void k_nt_update_fin(const double* parts, uint32_t it, NtScal* sc) {
  if (sc->state) return;
  const double rr = parts_tree<false>(parts, 2, 0), rz = parts_tree<false>(parts, 2, 1);
  if (threadIdx.x) return;
  if (sqrt(rr) <= sc->tol) {
    sc->state = 1;
    sc->its = it + 1;
    return;
  }
  sc->beta = rz / sc->rz;
  sc->rz = rz;
}
// p = z + beta p; partial max |p|.
__global__ __launch_bounds__(NT_TPB) void k_nt_dir(const NtScal* sc, const double* z, uint32_t m, double* p,
                                                   double* parts) {
  if (sc->state) return;
  const double beta = sc->beta;
  double v[1] = {0};
  NT_LOOP(i, m) {
    const double pi = z[i] + beta * p[i];
    p[i] = pi;
    v[0] = fmax(v[0], fabs(pi));
  }
  block_parts<1, 1u>(v, parts);
}
__global__ __launch_bounds__(NT_PARTS) void k_nt_dir_fin(const double* parts, double delta, NtScal* sc) {
  if (sc->state) return;
  const double pmax = parts_tree<true>(parts, 1, 0);
  if (threadIdx.x == 0) set_step(sc, pmax, delta);
}

struct Newton {
  cv_gpu& g;
  NtArgs a;
  uint32_t n, n3;
  double scale;  // scene scale for the difference step
  NewtonBufs& w;
  std::vector<double> host;

  // Reduces the K partials of the last reduction launch on the host.
  template <int K>
  void fetch(double (&out)[K], unsigned maxmask = 0) {
    CV_HIP(hipGetLastError());
    host.resize((size_t)K * NT_PARTS);
    CV_HIP(hipMemcpyAsync(host.data(), w.parts.p, host.size() * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    CV_HIP(hipStreamSynchronize(g.stream));
    for (int j = 0; j < K; ++j) {
      double s = host[j];
      for (uint32_t b = 1; b < NT_PARTS; ++b) {
        const double x = host[(size_t)K * b + j];
        s = ((maxmask >> j) & 1u) ? std::fmax(s, x) : s + x;
      }
      out[j] = s;
    }
  }
  double fetch1(unsigned maxmask = 0) {
    double o[1];
    fetch<1>(o, maxmask);
    return o[0];
  }
  void eval(const double4* X, double* grad) {
    k_nt_eval<<<(unsigned)(((uint64_t)n * NT_L + NT_TPB - 1) / NT_TPB), NT_TPB, 0, g.stream>>>(a, X, grad, w.diag.p,
                                                                                             w.epart.p, nullptr);
  }
  double objective(const double4* X, double* grad) {
    eval(X, grad);
    k_nt_sum<<<NT_PARTS, NT_TPB, 0, g.stream>>>(w.epart.p, n, w.parts.p);
    return fetch1();
  }
  double dot(const double* x, const double* y) {
    k_nt_dot<<<NT_PARTS, NT_TPB, 0, g.stream>>>(x, y, n3, w.parts.p);
    return fetch1();
  }
  // src/solver.rs:189-224 with A p by differences of the gradient.  The
  // diagonal of the current assembly is in w.diag0.  Returns false when the
  // curvature along a direction is not positive.  Iterations are queued in
  // batches of kBatch between read-backs of the state.
  bool pcg(double shift, uint32_t limit, uint32_t& its) {
    constexpr uint32_t kBatch = 16;
    const double delta = NT_DELTA * scale;
    NtScal* sc = w.sc.p;
    hipStream_t st = g.stream;
    k_nt_init<<<NT_PARTS, NT_TPB, 0, st>>>(w.g.p, w.diag0.p, w.X.p, n3, shift, w.sol.p, w.r.p, w.z.p, w.p.p,
                                           w.parts.p);
    k_nt_init_fin<<<1, NT_PARTS, 0, st>>>(w.parts.p, delta, sc);
    const unsigned eb = (unsigned)(((uint64_t)n * NT_L + NT_TPB - 1) / NT_TPB);
    NtScal h{};
    for (uint32_t it0 = 0; it0 < limit; it0 += kBatch) {
      for (uint32_t it = it0; it < std::min(limit, it0 + kBatch); ++it) {
        k_nt_probe<<<NT_PARTS, NT_TPB, 0, st>>>(w.X.p, w.p.p, sc, 1.0, n, w.Xt.p);
        k_nt_eval<<<eb, NT_TPB, 0, st>>>(a, w.Xt.p, w.gp.p, w.diag.p, w.epart.p, sc);
        k_nt_probe<<<NT_PARTS, NT_TPB, 0, st>>>(w.X.p, w.p.p, sc, -1.0, n, w.Xt.p);
        k_nt_eval<<<eb, NT_TPB, 0, st>>>(a, w.Xt.p, w.gm.p, w.diag.p, w.epart.p, sc);
        k_nt_ap<<<NT_PARTS, NT_TPB, 0, st>>>(w.gp.p, w.gm.p, w.p.p, w.X.p, sc, shift, n3, w.ap.p, w.parts.p);
        k_nt_ap_fin<<<1, NT_PARTS, 0, st>>>(w.parts.p, it, sc);
        k_nt_update<<<NT_PARTS, NT_TPB, 0, st>>>(sc, w.p.p, w.ap.p, w.diag0.p, w.X.p, shift, n3, w.sol.p, w.r.p,
                                                 w.z.p, w.parts.p);
        k_nt_update_fin<<<1, NT_PARTS, 0, st>>>(w.parts.p, it, sc);
        k_nt_dir<<<NT_PARTS, NT_TPB, 0, st>>>(sc, w.z.p, n3, w.p.p, w.parts.p);
        k_nt_dir_fin<<<1, NT_PARTS, 0, st>>>(w.parts.p, delta, sc);
      }
      CV_HIP(hipGetLastError());
      CV_HIP(hipMemcpyAsync(&h, sc, sizeof(NtScal), hipMemcpyDeviceToHost, st));
      CV_HIP(hipStreamSynchronize(st));
      if (h.state) {
        its = h.its;
        return h.state == 1;
      }
    }
    its = limit;
    return true;
  }
};

}  // namespace

// src/solver.rs:422-647 on the device (see the file header).
void newton_reference(cv_gpu& g, const cv_solver& s, const cv_state* st, const cv_settings* set,
                      const cv_newton_settings* ns, double* x_out, cv_newton_metrics* out) {
  const uint32_t n = s.n;
  NewtonBufs& w = g.nt;
  hipStream_t strm = g.stream;
  const DevSolver& d = g.dev;
  const size_t nt = (size_t)d.n_total;
  w.X.reserve(nt);
  w.Xt.reserve(nt);
  for (DBuf<double>* b : {&w.x0, &w.v0, &w.g, &w.gp, &w.gm, &w.diag, &w.diag0, &w.sol, &w.r, &w.z, &w.p, &w.ap})
    b->reserve(3 * (size_t)n);
  w.epart.reserve(n);
  w.parts.reserve(3 * NT_PARTS);
  w.sc.reserve(1);

  // Start positions with pins at their targets (src/solver.rs:437-440); the
  // avatar as fixed vertices after the garment.
  std::vector<double4> X(nt);
  double scale = 1.0;
  for (uint32_t i = 0; i < n; ++i) {
    const bool pin = s.pinned[i] != 0;
    const double* p = pin ? &s.pin_targets[3 * i] : &st->x[3 * i];
    X[i] = make_double4(p[0], p[1], p[2], pin ? 1.0 : 0.0);
    scale = std::fmax(scale, std::fmax(std::fabs(p[0]), std::fmax(std::fabs(p[1]), std::fabs(p[2]))));
  }
  for (size_t j = 0; j < s.avatar.size() / 3; ++j)
    X[n + j] = make_double4(s.avatar[3 * j], s.avatar[3 * j + 1], s.avatar[3 * j + 2], 1.0);
  w.X.upload(X, strm);
  w.Xt.upload(X, strm);  // the fixed part of every trial and difference point
  w.x0.upload(st->x, 3 * (size_t)n, strm);
  w.v0.upload(st->v, 3 * (size_t)n, strm);

  Newton nw{g, NtArgs{}, n, 3 * n, scale, w, {}};
  nw.a.term_off = d.term_off.p;
  nw.a.terms = d.terms_d.p;
  nw.a.coff = g.b.coff.p;
  nw.a.crec = g.b.crec_d.p;
  nw.a.x0 = w.x0.p;
  nw.a.v0 = w.v0.p;
  nw.a.masses = d.masses.p;
  nw.a.n = n;
  nw.a.dt = set->dt;
  nw.a.damping = set->damping;
  nw.a.gx = set->gravity[0];
  nw.a.gy = set->gravity[1];
  nw.a.gz = set->gravity[2];
  nw.a.cr = set->contact.radius;
  nw.a.ck = set->contact.stiffness;

  cv_newton_metrics m{};
  const double free_dofs = (double)std::max<uint32_t>(3 * s.n_free, 1);
  double tolerance = 0.0;
  for (uint32_t iteration = 0; iteration <= ns->newton_iterations; ++iteration) {
    const double energy = nw.objective(w.X.p, w.g.p);
    CV_HIP(hipMemcpyAsync(w.diag0.p, w.diag.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, strm));
    const double residual = std::sqrt(nw.dot(w.g.p, w.g.p) / free_dofs);
    if (iteration == 0 && !std::isfinite(energy))
      fail(CV_ERR_INVALID_STATE, "Invalid state, collapsed triangle or vertex at sphere centre");
    if (!std::isfinite(residual) || !std::isfinite(energy))
      fail(CV_ERR_NONFINITE_RESULT, "Non-finite FEM assembly; step rolled back");
    if (iteration == 0) {
      m.objective_initial_j = energy;
      tolerance = std::fmax(ns->absolute_tolerance, ns->relative_tolerance * residual);
    }
    m.residual_rms_n = residual;
    m.tolerance_n = tolerance;
    m.objective_final_j = energy;
    m.newton_iterations = iteration;
    if (residual <= tolerance) {
      m.converged = 1;
      break;
    }
    if (iteration == ns->newton_iterations) break;
    k_nt_max_diag<<<NT_PARTS, NT_TPB, 0, strm>>>(w.diag0.p, w.X.p, 3 * n, w.parts.p);
    const double scale_d = std::fmax(1.0, nw.fetch1(1u));
    bool accepted = false;
    double shift = 0.0;
    for (int rep = 0; rep < 12; ++rep) {
      uint32_t its = 0;
      const bool ok = nw.pcg(shift, ns->pcg_iterations, its);
      m.pcg_iterations += its;
      if (ok) {
        const double slope = nw.dot(w.g.p, w.sol.p);
        if (std::isfinite(slope) && slope < 0.0) {
          double alpha = 1.0;
          for (int k = 0; k < 24; ++k) {
            k_nt_move<<<NT_PARTS, NT_TPB, 0, strm>>>(w.X.p, w.sol.p, alpha, n, w.Xt.p);
            const double obj = nw.objective(w.Xt.p, w.gp.p);
            if (std::isfinite(obj) && obj <= energy + 1e-4 * alpha * slope) {
              accepted = true;
              break;
            }
            alpha *= 0.5;
            m.line_search_backtracks++;
          }
        }
      }
      if (accepted) break;
      shift = shift == 0.0 ? scale_d * 1e-6 : shift * 10.0;
    }
    m.max_regularization = std::fmax(m.max_regularization, shift);
    if (!accepted) fail(CV_ERR_LINE_SEARCH, "Newton line search failed; state unchanged. Reduce time step.");
    CV_HIP(hipMemcpyAsync(w.X.p, w.Xt.p, (size_t)n * sizeof(double4), hipMemcpyDeviceToDevice, strm));
  }
  CV_HIP(hipMemcpyAsync(X.data(), w.X.p, (size_t)n * sizeof(double4), hipMemcpyDeviceToHost, strm));
  CV_HIP(hipStreamSynchronize(strm));
  for (uint32_t i = 0; i < n; ++i) {
    x_out[3 * i] = X[i].x;
    x_out[3 * i + 1] = X[i].y;
    x_out[3 * i + 2] = X[i].z;
  }
  if (out) *out = m;
}

extern "C" cv_status cv_gpu_newton(cv_gpu* g, cv_solver* s, const cv_state* st, const cv_settings* set,
                                   const cv_newton_settings* ns, double* x_out, cv_newton_metrics* metrics) {
  if (!g || !s || !set || !ns || !x_out) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    // SolverSettings::validate (src/solver.rs:38-58).
    const double dt = set->dt;
    bool bad = !std::isfinite(dt) || !(dt >= 1e-6 && dt <= 0.1) || !std::isfinite(set->damping) ||
               set->damping < 0.0 || ns->newton_iterations == 0 || ns->newton_iterations > 200 ||
               ns->pcg_iterations == 0 || ns->pcg_iterations > 5000 || !std::isfinite(ns->absolute_tolerance) ||
               ns->absolute_tolerance <= 0.0 || !std::isfinite(ns->relative_tolerance) ||
               !(ns->relative_tolerance >= 0.0 && ns->relative_tolerance < 1.0);
    for (int k = 0; k < 3; ++k) bad = bad || !std::isfinite(set->gravity[k]);
    if (bad) fail(CV_ERR_INVALID_SETTINGS, "Invalid solver settings");
    const cv_ogc_settings& c = set->contact;
    if (!std::isfinite(c.radius) || c.radius <= 0.0 || !std::isfinite(c.stiffness) || c.stiffness <= 0.0)
      fail(CV_ERR_INVALID_OGC_SETTINGS, "Invalid OGC radius, stiffness or trust-region parameters");
    validate_state(*s, st);
    if (g->dev.generation != s->generation || !g->detected || !g->f64)
      fail(CV_ERR_INVALID_ARGUMENT,
           "cv_gpu_newton needs a cv_gpu_step or cv_gpu_detect of this solver with the f64 sweep first");
    newton_reference(*g, *s, st, set, ns, x_out, metrics);
  });
}
