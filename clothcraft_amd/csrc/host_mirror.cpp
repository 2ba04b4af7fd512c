// host_mirror.cpp — host (x86) build of the product's solver set-up and of
// the device math headers, behind a small C ABI.  The CPU test suite uses it
// to check, without a GPU, the exact per-term energies/gradients/Hessians the
// kernels evaluate (vbd_math.hpp, f32 and f64 instantiations), the packed
// term records, the colouring and the start-state validation.
#include <cstdio>
#include <cstring>
#include <vector>

#include "ogc_math.hpp"
#include "solver.hpp"
#include "vbd_math.hpp"

namespace {
template <class R>
struct Pos {
  const double* x;
  cvm::V3<R> operator()(int, uint32_t i) const { return {(R)x[3 * i], (R)x[3 * i + 1], (R)x[3 * i + 2]}; }
};
template <class R>
void deriv_out(const cvm::Deriv<R>& d, double* out) {
  out[0] = d.e;
  out[1] = d.g.x;
  out[2] = d.g.y;
  out[3] = d.g.z;
  for (int k = 0; k < 6; ++k) out[4 + k] = d.h.a[k];
}
}  // namespace

extern "C" {

int cvh_solver_create(const cv_garment* g, const cv_surface* avatar, const cv_state* st, cv_solver** out,
                      char* err, size_t errn) {
  cv_solver* s = new cv_solver();
  try {
    build_solver(*s, g, avatar, st);
  } catch (const CvError& e) {
    if (err && errn) std::snprintf(err, errn, "%s", e.message.c_str());
    delete s;
    return e.status;
  }
  *out = s;
  return 0;
}
void cvh_solver_destroy(cv_solver* s) { delete s; }

// Host half of the runtime gradient check (check_probe) on given inputs.
int cvh_check_probe(const cv_solver* s, const double* x, const double* t, const uint32_t* coff,
                    const cvm::TermRec64* crec, const double* origin, const cv_settings* set, const cv_grab* grab,
                    const double* probe, double* max_error, char* err, size_t errn) {
  try {
    *max_error = check_probe(*s, x, t, coff, crec, origin, set, grab, probe);
  } catch (const CvError& e) {
    if (err && errn) std::snprintf(err, errn, "%s", e.message.c_str());
    return e.status;
  }
  return 0;
}

uint64_t cvh_solver_terms(const cv_solver* s, uint32_t* term_off, cvm::TermRec* recs, uint64_t cap) {
  if (term_off) std::memcpy(term_off, s->term_off.data(), s->term_off.size() * sizeof(uint32_t));
  if (recs && cap >= s->terms.size()) std::memcpy(recs, s->terms.data(), s->terms.size() * sizeof(cvm::TermRec));
  return s->terms.size();
}
uint64_t cvh_solver_terms64(const cv_solver* s, cvm::TermRec64* recs, uint64_t cap) {
  if (recs && cap >= s->terms64.size())
    std::memcpy(recs, s->terms64.data(), s->terms64.size() * sizeof(cvm::TermRec64));
  return s->terms64.size();
}
uint32_t cvh_solver_edges(const cv_solver* s, uint32_t* out, uint32_t cap) {
  uint32_t E = (uint32_t)(s->top.edges.size() / 2);
  if (out && cap >= E) std::memcpy(out, s->top.edges.data(), 2 * E * sizeof(uint32_t));
  return E;
}

}  // extern "C"

extern "C" {

// Incremental colouring state for tests of update_colors.
void* cvh_coloring_new() { return new ColoringState(); }
void cvh_coloring_free(void* p) { delete (ColoringState*)p; }
uint32_t cvh_coloring_update(const cv_solver* s, void* state, const uint32_t* kind_ids, const uint32_t* counts,
                             uint64_t m, const cv_grab* grab, int smallest_last, uint32_t* color_of) {
  std::vector<uint32_t> coff;
  ColoringState& cs = *(ColoringState*)state;
  uint32_t k = update_colors(*s, cs, kind_ids, counts, m, grab, smallest_last != 0, coff);
  std::memcpy(color_of, cs.color_of.data(), s->n * sizeof(uint32_t));
  return k;
}

// Greedy colouring with contact cliques given as (kind, ids[4], count) rows.
uint32_t cvh_greedy_colors(const cv_solver* s, const uint32_t* kind_ids, const uint32_t* counts,
                           uint64_t m, const cv_grab* grab, uint32_t* color_of, int smallest_last) {
  uint32_t n = s->n;
  std::vector<uint32_t> off(n + 1, 0), nb;
  for (uint64_t c = 0; c < m; ++c)
    for (uint32_t q = 0; q < counts[c]; ++q) {
      uint32_t v = kind_ids[5 * c + 1 + q];
      if (v < n) off[v + 1] += counts[c];
    }
  for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
  nb.resize(off[n]);
  std::vector<uint32_t> cur(off.begin(), off.end() - 1);
  for (uint64_t c = 0; c < m; ++c)
    for (uint32_t q = 0; q < counts[c]; ++q) {
      uint32_t v = kind_ids[5 * c + 1 + q];
      if (v < n)
        for (uint32_t r = 0; r < counts[c]; ++r) nb[cur[v]++] = kind_ids[5 * c + 1 + r];
    }
  std::vector<uint32_t> col, order, coff;
  uint32_t k = greedy_colors(*s, off.data(), nb.data(), grab, col, order, coff, smallest_last != 0);
  std::memcpy(color_of, col.data(), n * sizeof(uint32_t));
  return k;
}

// One term record evaluated for its vertex at xv with the other vertices at x
// (absolute f64 positions, rounded to R first).  out = e, g[3], H[6].
void cvh_term_deriv_f64(const cvm::TermRec* r, const double* xv, const double* x, double cr, double ck,
                        double* out) {
  Pos<double> P{x};
  deriv_out(cvm::term_deriv<double>(*r, cvm::V3<double>{xv[0], xv[1], xv[2]}, cr, ck, P), out);
}
void cvh_term_deriv_f32(const cvm::TermRec* r, const double* xv, const double* x, double cr, double ck,
                        double* out) {
  Pos<float> P{x};
  deriv_out(cvm::term_deriv<float>(*r, cvm::V3<float>{(float)xv[0], (float)xv[1], (float)xv[2]},
                                   (float)cr, (float)ck, P),
            out);
}
double cvh_term_energy_f64(const cvm::TermRec* r, const double* xv, const double* x, double cr, double ck) {
  Pos<double> P{x};
  return cvm::term_energy<double>(*r, cvm::V3<double>{xv[0], xv[1], xv[2]}, cr, ck, P);
}
double cvh_term_energy_f32(const cvm::TermRec* r, const double* xv, const double* x, double cr, double ck) {
  Pos<float> P{x};
  return cvm::term_energy<float>(*r, cvm::V3<float>{(float)xv[0], (float)xv[1], (float)xv[2]}, (float)cr,
                                 (float)ck, P);
}
// Contact record builder matching the device (k_contact_records).
void cvh_contact_record(int kind, const uint32_t* ids, int count, double eps, uint32_t v, cvm::TermRec* out) {
  cvm::TermRec r{};
  int slot = 0;
  for (int q = 0; q < 4; ++q) {
    r.ids[q] = ids[q];
    if (q < count && ids[q] == v) slot = q;
  }
  r.head = cvm::rec_head(cvm::TERM_CONTACT, slot, count, kind);
  r.f[0] = (float)eps;
  *out = r;
}
// Newton direction of the local solve (src/vbd.rs:353-371) for a summed
// derivative in[10] (e, g, H) plus inertia.
void cvh_newton_direction_f64(const double* in, const double* x, const double* target, double inertia,
                              double* delta) {
  cvm::Deriv<double> d;
  d.e = in[0];
  d.g = {in[1], in[2], in[3]};
  for (int k = 0; k < 6; ++k) d.h.a[k] = in[4 + k];
  cvm::V3<double> r = cvm::newton_direction<double>(d, {x[0], x[1], x[2]}, {target[0], target[1], target[2]},
                                                    inertia);
  delta[0] = r.x;
  delta[1] = r.y;
  delta[2] = r.z;
}
// f64 narrow-phase pieces (ogc_math.hpp) for comparison with the oracle.
void cvh_closest_triangle(const double* p, const double* t9, double* q, int* feature, int* local) {
  ogm::D3 Q;
  ogm::closest_triangle({p[0], p[1], p[2]}, {t9[0], t9[1], t9[2]}, {t9[3], t9[4], t9[5]},
                        {t9[6], t9[7], t9[8]}, Q, *feature, *local);
  q[0] = Q.x;
  q[1] = Q.y;
  q[2] = Q.z;
}
void cvh_closest_segments(const double* a, const double* b, const double* c, const double* d, double* s,
                          double* t) {
  ogm::closest_segments({a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, {d[0], d[1], d[2]}, *s,
                        *t);
}

}  // extern "C"

// Sequential host mirrors of one colour of the device sweep (k_sweep) in f64
// and in f32: same record math, Newton step, truncation and accept test; the
// only difference from the device is the order of the per-vertex term sums.
namespace {
template <class R>
struct PosArr {
  const R* p;  // [4 * n] xyz w
  cvm::V3<R> operator()(int, uint32_t i) const { return {p[4 * i], p[4 * i + 1], p[4 * i + 2]}; }
};
template <class R>
void sweep_color(uint32_t count, const uint32_t* order, R* pos, const R* tgt, const R* ball,
                 const uint32_t* term_off, const cvm::TermRecT<R>* terms, const uint32_t* coff,
                 const cvm::TermRecT<R>* crec, R cr, R ck, R* gsq, uint64_t* clipped, uint32_t* attempts) {
  PosArr<R> P{pos};
  std::vector<std::pair<uint32_t, cvm::V3<R>>> upd;
  for (uint32_t gi = 0; gi < count; ++gi) {
    uint32_t v = order[gi];
    if (pos[4 * v + 3] > R(0.5)) {
      gsq[v] = 0;
      continue;
    }
    cvm::V3<R> x{pos[4 * v], pos[4 * v + 1], pos[4 * v + 2]};
    uint32_t s0 = term_off[v], ns = term_off[v + 1] - s0, c0 = coff[v], nt = ns + (coff[v + 1] - c0);
    cvm::Deriv<R> acc = cvm::deriv_zero<R>();
    auto energy = [&](cvm::V3<R> p) {
      R e = 0;
      for (uint32_t i = 0; i < nt; ++i)
        e += cvm::term_energy<R>(i < ns ? terms[s0 + i] : crec[c0 + i - ns], p, cr, ck, P);
      return e;
    };
    for (uint32_t i = 0; i < nt; ++i)
      cvm::deriv_add(acc, cvm::term_deriv<R>(i < ns ? terms[s0 + i] : crec[c0 + i - ns], x, cr, ck, P));
    cvm::V3<R> target{tgt[4 * v], tgt[4 * v + 1], tgt[4 * v + 2]};
    R inertia = tgt[4 * v + 3];
    cvm::V3<R> dx = x - target;
    R old = acc.e + R(0.5) * inertia * cvm::dot(dx, dx);  // the terms' own values, as the device
    cvm::V3<R> delta = cvm::newton_direction<R>(acc, x, target, inertia);
    gsq[v] = cvm::dot(acc.g, acc.g);
    cvm::V3<R> center{ball[4 * v], ball[4 * v + 1], ball[4 * v + 2]};
    bool tr = false;
    cvm::V3<R> accepted = x;
    R scale = 1;
    int a = 0;
    for (; a < 10; ++a) {
      cvm::V3<R> cand = cvm::truncate_ball<R>(x + delta * scale, center, ball[4 * v + 3], tr);
      cvm::V3<R> dc = cand - target;
      R e = energy(cand) + R(0.5) * inertia * cvm::dot(dc, dc);
      if (cvm::finite3(cand) && std::isfinite(e) && e <= old + cvm::accept_slack<R>(old)) {
        accepted = cand;
        break;
      }
      scale *= R(0.5);
    }
    if (attempts) attempts[v] = a;
    upd.push_back({v, accepted});
    if (tr) (*clipped)++;
  }
  for (auto& [v, p] : upd) {
    pos[4 * v] = p.x;
    pos[4 * v + 1] = p.y;
    pos[4 * v + 2] = p.z;
  }
}
// The f32 sweep as the device runs it: positions, targets and ball centres
// stay f64; each vertex solve works in f32 on coordinates relative to that
// vertex's own position (every term, the inertia and the ball depend on
// differences only), so f32 rounds lengths of the one-ring's size, not of
// the scene's.
struct PosRel {
  const double* p;
  double bx, by, bz;
  cvm::V3<float> operator()(int, uint32_t i) const {
    return {(float)(p[4 * i] - bx), (float)(p[4 * i + 1] - by), (float)(p[4 * i + 2] - bz)};
  }
};
void sweep_color_local(uint32_t count, const uint32_t* order, double* pos, const double* tgt, const double* ball,
                       const uint32_t* term_off, const cvm::TermRec* terms, const uint32_t* coff,
                       const cvm::TermRec* crec, float cr, float ck, double* gsq, uint64_t* clipped,
                       uint32_t* attempts) {
  std::vector<std::pair<uint32_t, cvm::V3<double>>> upd;
  for (uint32_t gi = 0; gi < count; ++gi) {
    uint32_t v = order[gi];
    if (pos[4 * v + 3] > 0.5) {
      gsq[v] = 0;
      continue;
    }
    const double bx = pos[4 * v], by = pos[4 * v + 1], bz = pos[4 * v + 2];
    PosRel P{pos, bx, by, bz};
    cvm::V3<float> x{0.f, 0.f, 0.f};
    uint32_t s0 = term_off[v], ns = term_off[v + 1] - s0, c0 = coff[v], nt = ns + (coff[v + 1] - c0);
    cvm::Deriv<float> acc = cvm::deriv_zero<float>();
    // The line search accepts on the change of the local energy: records'
    // energies at the candidate less their energies at x (ex), stitches in
    // the exact difference form, the inertia as m/2 dx.(c + x - 2 t).
    float ex = 0, emag = 0;
    auto change = [&](cvm::V3<float> p) {
      float e = -ex;
      for (uint32_t i = 0; i < nt; ++i)
        e += cvm::term_change_part<float>(i < ns ? terms[s0 + i] : crec[c0 + i - ns], x, p, cr, ck, P);
      return e;
    };
    for (uint32_t i = 0; i < nt; ++i) {
      const cvm::TermRec& t = i < ns ? terms[s0 + i] : crec[c0 + i - ns];
      cvm::Deriv<float> d = cvm::term_deriv<float>(t, x, cr, ck, P);
      if (cvm::rec_kind(t.head) != cvm::TERM_STITCH) {
        ex += d.e;
        emag += std::fabs(d.e);
      }
      cvm::deriv_add(acc, d);
    }
    cvm::V3<float> target{(float)(tgt[4 * v] - bx), (float)(tgt[4 * v + 1] - by), (float)(tgt[4 * v + 2] - bz)};
    float inertia = (float)tgt[4 * v + 3];
    cvm::V3<float> dx = x - target;
    const float slack = 1e-14f + 4.0f * 1.1920928955078125e-07f * emag;
    cvm::V3<float> delta = cvm::newton_direction<float>(acc, x, target, inertia);
    gsq[v] = (double)cvm::dot(acc.g, acc.g);
    cvm::V3<float> center{(float)(ball[4 * v] - bx), (float)(ball[4 * v + 1] - by), (float)(ball[4 * v + 2] - bz)};
    bool tr = false;
    cvm::V3<float> accepted = x;
    float scale = 1;
    int a = 0;
    for (; a < 10; ++a) {
      cvm::V3<float> cand = cvm::truncate_ball<float>(x + delta * scale, center, (float)ball[4 * v + 3], tr);
      cvm::V3<float> dc = cand - target;
      float e = change(cand) + 0.5f * inertia * cvm::dot(cand - x, dc + dx);
      if (cvm::finite3(cand) && std::isfinite(e) && e <= slack) {
        accepted = cand;
        break;
      }
      scale *= 0.5f;
    }
    if (attempts) attempts[v] = a;
    upd.push_back({v, cvm::V3<double>{bx + (double)accepted.x, by + (double)accepted.y, bz + (double)accepted.z}});
    if (tr) (*clipped)++;
  }
  for (auto& [v, p] : upd) {
    pos[4 * v] = p.x;
    pos[4 * v + 1] = p.y;
    pos[4 * v + 2] = p.z;
  }
}
}  // namespace

extern "C" {
void cvh_sweep_color_f64(uint32_t count, const uint32_t* order, double* pos, const double* tgt,
                         const double* ball, const uint32_t* term_off, const cvm::TermRec64* terms,
                         const uint32_t* coff, const cvm::TermRec64* crec, double cr, double ck, double* gsq,
                         uint64_t* clipped, uint32_t* attempts) {
  sweep_color<double>(count, order, pos, tgt, ball, term_off, terms, coff, crec, cr, ck, gsq, clipped,
                      attempts);
}
void cvh_sweep_color_f32(uint32_t count, const uint32_t* order, double* pos, const double* tgt,
                         const double* ball, const uint32_t* term_off, const cvm::TermRec* terms,
                         const uint32_t* coff, const cvm::TermRec* crec, double cr, double ck, double* gsq,
                         uint64_t* clipped, uint32_t* attempts) {
  sweep_color_local(count, order, pos, tgt, ball, term_off, terms, coff, crec, (float)cr, (float)ck, gsq, clipped,
                    attempts);
}
}
