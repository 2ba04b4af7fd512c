// ORACLE — test infrastructure only.  extern "C" surface of the f64
// restatement, loaded by the tests and the bench's CPU-baseline leg through
// ctypes.  It shares the plain data structs of include/clothvbd.h so a test
// hands the same arrays to the oracle and to the product library.
#include <cstdio>

#include "../include/clothvbd.h"
#include "oracle_core.hpp"

using namespace cvo;

namespace {

int status_of(const std::string& msg) {
  if (msg.size() < 2 || msg[1] != '|') return CV_ERR_INVALID_ARGUMENT;
  switch (msg[0]) {
    case 'T': return CV_ERR_TOPOLOGY;
    case 'D': return CV_ERR_DEGENERATE_FACE;
    case 'P': return CV_ERR_INVALID_POSITIONS;
    case 'F': return CV_ERR_FIXED_CHANGED;
    case 'O': return CV_ERR_INVALID_OGC_SETTINGS;
    case 'S': return CV_ERR_INVALID_SETTINGS;
    case 'A': return CV_ERR_INVALID_ARGUMENT;
    case 'Z': return CV_ERR_ZERO_SEPARATION;
    case 'N': return CV_ERR_NONFINITE_RESULT;
    case 'R': return CV_ERR_TRUST_REGION;
    case 'I': return CV_ERR_INITIAL_INTERSECTION;
    case 'U': return CV_ERR_INITIAL_TOUCHING;
    case 'Q': return CV_ERR_QUERY_RADIUS;
    case 'V': return CV_ERR_INVALID_STATE;
    case 'G': return CV_ERR_INVALID_GRAB;
    case 'M': return CV_ERR_DIMENSIONS;
    case 'K': return CV_ERR_PIN_MISMATCH;
    case 'W': return CV_ERR_INVALID_AVATAR;
    case 'L': return CV_ERR_NONFINITE_RESULT;
  }
  return CV_ERR_INVALID_ARGUMENT;
}
void put_err(const std::string& msg, char* err, size_t n) {
  if (!err || n == 0) return;
  std::string text = (msg.size() > 2 && msg[1] == '|') ? msg.substr(2) : msg;
  std::snprintf(err, n, "%s", text.c_str());
}
template <class F>
int guarded(char* err, size_t n, F&& f) {
  try {
    f();
    return CV_OK;
  } catch (const std::exception& e) {
    put_err(e.what(), err, n);
    return status_of(e.what());
  }
}

Material material_of(const cv_material& m) {
  Material r;
  r.warp = m.warp;
  r.weft = m.weft;
  r.shear = m.shear;
  r.compression_ratio = m.compression_ratio;
  r.compression_transition = m.compression_transition;
  r.bend_warp = m.bend_warp;
  r.bend_weft = m.bend_weft;
  r.density = m.density;
  r.thickness = m.thickness;
  r.grain_degrees = m.grain_degrees;
  return r;
}

Garment garment_of(const cv_garment* g) {
  Garment o;
  o.n = (int)g->n_vertices;
  o.masses.assign(g->masses, g->masses + g->n_vertices);
  o.pins.resize(o.n);
  for (int i = 0; i < o.n; ++i)
    if (g->pinned && g->pinned[i])
      o.pins[i] = V3{g->pin_targets[3 * i], g->pin_targets[3 * i + 1], g->pin_targets[3 * i + 2]};
  for (uint32_t i = 0; i < g->n_materials; ++i) o.materials.push_back(material_of(g->materials[i]));
  for (uint32_t i = 0; i < g->n_triangles; ++i) {
    Triangle t;
    for (int k = 0; k < 3; ++k) t.v[k] = (int)g->tri_vertices[3 * i + k];
    for (int k = 0; k < 4; ++k) t.inv_dm[k] = g->tri_inv_dm[4 * i + k];
    t.area = g->tri_area[i];
    t.material = (int)g->tri_material[i];
    o.triangles.push_back(t);
  }
  for (uint32_t i = 0; i < g->n_hinges; ++i) {
    Hinge h;
    for (int k = 0; k < 4; ++k) h.v[k] = (int)g->hinge_vertices[4 * i + k];
    h.weight = g->hinge_weight[i];
    h.across[0] = g->hinge_across[2 * i];
    h.across[1] = g->hinge_across[2 * i + 1];
    h.material = (int)g->hinge_material[i];
    o.hinges.push_back(h);
  }
  for (uint32_t s = 0; s < g->n_stitches; ++s) {
    StitchPointT p;
    for (uint32_t k = g->stitch_offsets[s]; k < g->stitch_offsets[s + 1]; ++k)
      p.terms.push_back({(int)g->stitch_vertices[k], g->stitch_weights[k]});
    o.stitch_points.push_back(p);
    o.stitch_stiffness.push_back(g->stitch_stiffness[s]);
  }
  for (uint32_t s = 0; s < g->n_probes; ++s) {
    StitchPointT p;
    for (uint32_t k = g->probe_offsets[s]; k < g->probe_offsets[s + 1]; ++k)
      p.terms.push_back({(int)g->probe_vertices[k], g->probe_weights[k]});
    o.probes.push_back(p);
  }
  return o;
}
std::vector<V3> v3s(const double* p, size_t n) {
  std::vector<V3> r(n);
  for (size_t i = 0; i < n; ++i) r[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  return r;
}
State state_of(const cv_state* s) {
  State o;
  o.x = v3s(s->x, s->n);
  o.v = v3s(s->v, s->n);
  o.time = s->time;
  return o;
}
VbdSettings settings_of(const cv_settings* s) {
  VbdSettings o;
  o.dt = s->dt;
  o.iterations = (int)s->iterations;
  o.damping = s->damping;
  o.gravity = {s->gravity[0], s->gravity[1], s->gravity[2]};
  o.contact.radius = s->contact.radius;
  o.contact.stiffness = s->contact.stiffness;
  o.contact.relaxation = s->contact.relaxation;
  o.contact.refresh_fraction = s->contact.refresh_fraction;
  o.residual_tolerance = s->residual_tolerance;
  o.residual_interval = s->residual_interval;
  o.fast_coloring = (s->flags & CV_FLAG_FAST_COLORING) != 0;
  return o;
}
OgcSettings ogc_of(const cv_ogc_settings* s) {
  OgcSettings o;
  o.radius = s->radius;
  o.stiffness = s->stiffness;
  o.relaxation = s->relaxation;
  o.refresh_fraction = s->refresh_fraction;
  return o;
}
GrabT grab_of(const cv_grab* g) {
  GrabT o;
  for (int k = 0; k < 3; ++k) {
    o.vertices[k] = (int)g->vertices[k];
    o.weights[k] = g->weights[k];
  }
  o.target = {g->target[0], g->target[1], g->target[2]};
  o.stiffness = g->stiffness;
  return o;
}
void fill_metrics(const VbdMetrics& m, cv_metrics* out) {
  if (!out) return;
  *out = cv_metrics{};
  out->time_s = m.time_s;
  out->dt_s = m.dt_s;
  out->step_ms = m.step_ms;
  out->detection_ms = m.detection_ms;
  out->solve_ms = m.solve_ms;
  out->gpu_compute_ms = std::numeric_limits<double>::quiet_NaN();
  out->refit_ms = m.profile.refit_ms;
  out->candidates_ms = m.profile.candidates_ms;
  out->narrow_ms = m.profile.narrow_ms;
  out->candidate_pairs = m.profile.candidate_pairs;
  out->aabb_rejections = m.profile.aabb_rejections;
  out->iterations = (uint32_t)m.iterations;
  out->contacts = m.contacts;
  out->contact_terms = m.contact_terms;
  out->detection_updates = m.detection_updates;
  out->broad_phase_rebuilds = m.broad_phase_rebuilds;
  out->truncated_vertices = m.truncated_vertices;
  out->minimum_distance_m = m.minimum_distance_m;
  out->max_seam_gap_m = m.max_seam_gap_m;
  out->extension_p95_percent = m.extension_p95;
  out->extension_max_percent = m.extension_max;
  out->residual_rms_n = m.residual_rms_n;
  out->true_residual_rms_n = m.true_residual_rms_n;
  out->colors = m.colors;
  out->iterations_run = (uint32_t)m.iterations_run;
}

}  // namespace

// Exported contact-set view for tests (matches ContactSet, src/ogc/world.rs:74-87).
struct cvo_contacts {
  uint64_t capacity;          /* in: rows available in kind_ids/counts/eps */
  uint64_t n_contacts;        /* out */
  uint32_t* kind_ids;         /* [5*cap] kind, v0..v3 */
  uint32_t* counts;           /* [cap] */
  double* eps;                /* [cap] */
  double* radii;              /* [moving] or NULL */
  uint64_t active_contacts, vertex_face_candidates, edge_edge_candidates, candidate_pairs,
      aabb_rejections;
  double minimum_distance;
  int32_t broad_phase_rebuilt;
  uint32_t* incident_offsets; /* [moving+1] or NULL */
  uint32_t* incident;         /* [4*cap] or NULL */
};

static void export_contacts(const ContactSet& c, cvo_contacts* out) {
  if (!out) return;
  out->n_contacts = c.contacts.size();
  uint64_t m = std::min<uint64_t>(out->capacity, c.contacts.size());
  for (uint64_t i = 0; i < m; ++i) {
    if (out->kind_ids) {
      out->kind_ids[5 * i] = (uint32_t)c.contacts[i].kind;
      for (int k = 0; k < 4; ++k) out->kind_ids[5 * i + 1 + k] = (uint32_t)c.contacts[i].vertices[k];
    }
    if (out->counts) out->counts[i] = (uint32_t)c.contacts[i].count;
    if (out->eps) out->eps[i] = c.contacts[i].mollifier_epsilon;
  }
  if (out->radii)
    for (size_t i = 0; i < c.radii.size(); ++i) out->radii[i] = c.radii[i];
  out->active_contacts = c.active_contacts;
  out->vertex_face_candidates = c.vertex_face_candidates;
  out->edge_edge_candidates = c.edge_edge_candidates;
  out->candidate_pairs = c.profile.candidate_pairs;
  out->aabb_rejections = c.profile.aabb_rejections;
  out->minimum_distance = c.minimum_distance;
  out->broad_phase_rebuilt = c.broad_phase_rebuilt ? 1 : 0;
  if (out->incident_offsets) {
    uint32_t off = 0;
    for (size_t v = 0; v < c.incident.size(); ++v) {
      out->incident_offsets[v] = off;
      for (int ci : c.incident[v]) {
        if (out->incident && off < 4 * out->capacity) out->incident[off] = (uint32_t)ci;
        off++;
      }
    }
    out->incident_offsets[c.incident.size()] = off;
  }
}

extern "C" {

// ------------------------------------------------------------- solver
struct cvo_solver {
  VbdSolver s;
};

int cvo_solver_create(const cv_garment* g, const cv_surface* avatar, const cv_state* st,
                      cvo_solver** out, char* err, size_t errn) {
  return guarded(err, errn, [&] {
    Garment gg = garment_of(g);
    State s = state_of(st);
    std::vector<V3> av;
    std::vector<std::array<int, 3>> at;
    if (avatar) {
      av = v3s(avatar->vertices, avatar->n_vertices);
      for (uint32_t i = 0; i < avatar->n_triangles; ++i)
        at.push_back({(int)avatar->triangles[3 * i], (int)avatar->triangles[3 * i + 1],
                      (int)avatar->triangles[3 * i + 2]});
    }
    auto* h = new cvo_solver{VbdSolver::create(gg, avatar ? &av : nullptr, avatar ? &at : nullptr, s)};
    *out = h;
  });
}
void cvo_solver_destroy(cvo_solver* s) { delete s; }
void cvo_solver_set_threads(cvo_solver* s, int threads) { s->s.threads = std::max(1, threads); }

int cvo_solver_step(cvo_solver* s, cv_state* st, const cv_settings* settings, const cv_grab* grab,
                    cv_metrics* metrics, char* err, size_t errn) {
  return guarded(err, errn, [&] {
    State state = state_of(st);
    VbdSettings vs = settings_of(settings);
    GrabT g;
    if (grab) g = grab_of(grab);
    VbdMetrics m = s->s.step(state, vs, grab ? &g : nullptr);
    if ((settings->flags & CV_FLAG_TRUE_RESIDUAL) && !(vs.residual_tolerance > 0.0)) {
      // Post-sweep residual of the committed positions, against this step's
      // target/inertia and final contact set.
      int n = s->s.garment.n;
      State old = state_of(st);
      double dt = vs.dt, damp = 1.0 + vs.damping * dt;
      std::vector<V3> target(n);
      std::vector<double> inertia(n);
      for (int i = 0; i < n; ++i) {
        target[i] = old.x[i] + (old.v[i] * dt + vs.gravity * dt * dt) / damp;
        inertia[i] = s->s.garment.masses[i] * damp / (dt * dt);
      }
      m.true_residual_rms_n = s->s.true_residual(s->s.last_x, target, inertia, s->s.last_contacts,
                                                 vs, grab ? &g : nullptr);
    }
    for (uint32_t i = 0; i < st->n; ++i) {
      st->x[3 * i] = state.x[i].x;
      st->x[3 * i + 1] = state.x[i].y;
      st->x[3 * i + 2] = state.x[i].z;
      st->v[3 * i] = state.v[i].x;
      st->v[3 * i + 1] = state.v[i].y;
      st->v[3 * i + 2] = state.v[i].z;
    }
    st->time = state.time;
    fill_metrics(m, metrics);
  });
}

int cvo_solver_record_prediction(cvo_solver* s, const double* old_velocity, const cv_state* st,
                                 const cv_settings* settings) {
  return guarded(nullptr, 0, [&] {
    State state = state_of(st);
    s->s.prediction.record(v3s(old_velocity, st->n), state, settings_of(settings));
  });
}
void cvo_solver_clear_prediction(cvo_solver* s) { s->s.prediction = Prediction{}; }
// The next step starts from `guess` [3n] instead of Prediction::guess.
void cvo_solver_set_next_guess(cvo_solver* s, const double* guess) {
  s->s.next_guess = v3s(guess, s->s.garment.n);
}

// Last step's detection, colouring and gradients (test introspection).
void cvo_solver_last_contacts(const cvo_solver* s, cvo_contacts* out) {
  export_contacts(s->s.last_contacts, out);
}
uint32_t cvo_solver_last_colors(const cvo_solver* s, uint32_t* color_of_vertex, uint32_t n) {
  for (size_t c = 0; c < s->s.last_colors.size(); ++c)
    for (int v : s->s.last_colors[c])
      if ((uint32_t)v < n) color_of_vertex[v] = (uint32_t)c;
  return (uint32_t)s->s.last_colors.size();
}
// Colourings of a conflict graph, contact set and grab: 0 the reference's
// greedy in index order (src/vbd/coloring.rs:5-43), 1 the same greedy in
// smallest-last order of the static graph, 2 the fast colouring (fast_colors).
std::vector<std::vector<int>> color_mode(const std::vector<std::vector<int>>& graph, const ContactSet& c,
                                         const GrabT* g, int mode) {
  if (mode == 0) return colors(graph, c, g);
  std::vector<int> sl = smallest_last(graph);
  return mode == 1 ? colors(graph, c, g, &sl) : fast_colors(graph, c, g, sl);
}
// Colouring of an arbitrary contact set at positions x: runs detect on x with
// rq, then the colouring of `mode` (color_mode) — used to compare the
// product's colourings with the oracle's.
int cvo_solver_detect_and_color(cvo_solver* s, const double* x_moving, double rq,
                                const cv_ogc_settings* settings, const cv_grab* grab,
                                uint32_t* color_of_vertex, uint32_t* n_colors, cvo_contacts* out,
                                char* err, size_t errn, int mode) {
  return guarded(err, errn, [&] {
    auto x = s->s.world.positions(v3s(x_moving, s->s.garment.n));
    ContactSet c = s->s.world.detect(x, rq, ogc_of(settings));
    GrabT g;
    if (grab) g = grab_of(grab);
    auto cols = color_mode(s->s.graph, c, grab ? &g : nullptr, mode);
    for (size_t ci = 0; ci < cols.size(); ++ci)
      for (int v : cols[ci]) color_of_vertex[v] = (uint32_t)ci;
    *n_colors = (uint32_t)cols.size();
    export_contacts(c, out);
    s->s.last_contacts = std::move(c);  // local_gradient may then evaluate at this set
  });
}

// Local objective at vertex v (the Jet<3> evaluation of src/vbd.rs:345-356 plus
// inertia), using the last step's contact set — mirrors check_gradients.
int cvo_solver_local_gradient(cvo_solver* s, const double* x_moving, const double* target,
                              const double* inertia, const cv_settings* settings,
                              const cv_grab* grab, int use_last_contacts, double* out4,
                              char* err, size_t errn) {
  return guarded(err, errn, [&] {
    int n = s->s.garment.n;
    auto x = s->s.world.positions(v3s(x_moving, n));
    ContactSet empty;
    empty.incident.assign(n, {});
    bool have = use_last_contacts && s->s.last_contacts.incident.size() == (size_t)n;
    const ContactSet& c = have ? s->s.last_contacts : empty;
    VbdSettings vs = settings_of(settings);
    GrabT g;
    if (grab) g = grab_of(grab);
    for (int v = 0; v < n; ++v) {
      Jet<3> j = s->s.local<Jet<3>>(
          v, {Jet<3>::variable(x[v].x, 0), Jet<3>::variable(x[v].y, 1), Jet<3>::variable(x[v].z, 2)},
          x, c, vs, grab ? &g : nullptr);
      V3 t{target[3 * v], target[3 * v + 1], target[3 * v + 2]};
      V3 gg = V3{j.g[0], j.g[1], j.g[2]} + (x[v] - t) * inertia[v];
      out4[4 * v] = gg.x;
      out4[4 * v + 1] = gg.y;
      out4[4 * v + 2] = gg.z;
      out4[4 * v + 3] = j.v + 0.5 * inertia[v] * length2(x[v] - t);
    }
  });
}

// ------------------------------------------------------------- energies
// Membrane element: energy, gradient [9], Hessian [81] (Jet<9>, src/energy.rs:68-73).
double cvo_membrane(const double* x9, const double* inv_dm, double area, const cv_material* m,
                    double* g9, double* h81) {
  Triangle t;
  t.v = {0, 1, 2};
  for (int k = 0; k < 4; ++k) t.inv_dm[k] = inv_dm[k];
  t.area = area;
  t.material = 0;
  std::array<A3<Jet<9>>, 3> p;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) p[a][k] = Jet<9>::variable(x9[3 * a + k], 3 * a + k);
  Jet<9> j = membrane<Jet<9>>(p, t, material_of(*m));
  if (g9)
    for (int i = 0; i < 9; ++i) g9[i] = j.g[i];
  if (h81)
    for (int i = 0; i < 9; ++i)
      for (int k = 0; k < 9; ++k) h81[9 * i + k] = j.h[i][k];
  return j.v;
}
double cvo_membrane_value(const double* x9, const double* inv_dm, double area,
                          const cv_material* m) {
  Triangle t;
  t.v = {0, 1, 2};
  for (int k = 0; k < 4; ++k) t.inv_dm[k] = inv_dm[k];
  t.area = area;
  t.material = 0;
  std::array<A3<double>, 3> p;
  for (int a = 0; a < 3; ++a) p[a] = {x9[3 * a], x9[3 * a + 1], x9[3 * a + 2]};
  return membrane<double>(p, t, material_of(*m));
}
// Bending element (Jet<12>, src/energy.rs:74-79).
double cvo_bending(const double* x12, double weight, const double* across, const cv_material* m,
                   double* g12, double* h144) {
  Hinge h;
  h.v = {0, 1, 2, 3};
  h.weight = weight;
  h.across[0] = across[0];
  h.across[1] = across[1];
  h.material = 0;
  std::array<A3<Jet<12>>, 4> p;
  for (int a = 0; a < 4; ++a)
    for (int k = 0; k < 3; ++k) p[a][k] = Jet<12>::variable(x12[3 * a + k], 3 * a + k);
  Jet<12> j = bending<Jet<12>>(p, h, material_of(*m));
  if (g12)
    for (int i = 0; i < 12; ++i) g12[i] = j.g[i];
  if (h144)
    for (int i = 0; i < 12; ++i)
      for (int k = 0; k < 12; ++k) h144[12 * i + k] = j.h[i][k];
  return j.v;
}
double cvo_bending_value(const double* x12, double weight, const double* across,
                         const cv_material* m) {
  Hinge h;
  h.v = {0, 1, 2, 3};
  h.weight = weight;
  h.across[0] = across[0];
  h.across[1] = across[1];
  h.material = 0;
  std::array<A3<double>, 4> p;
  for (int a = 0; a < 4; ++a) p[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
  return bending<double>(p, h, material_of(*m));
}
// activation with first/second derivative (Jet<1>, src/ogc/potential.rs:8-22).
double cvo_activation(double d, double r, double k, double* g, double* h) {
  Jet<1> j = activation<Jet<1>>(Jet<1>::variable(d, 0), r, k);
  if (g) *g = j.g[0];
  if (h) *h = j.h[0][0];
  return j.v;
}
double cvo_activation_value(double d, double r, double k) { return activation<double>(d, r, k); }
// Contact::block (src/ogc/potential.rs:69-85) on 4 points x12.
double cvo_contact_block(int kind, const double* x12, int count, double eps, int vertex, double r,
                         double k, double* g3, double* h9) {
  Contact c{kind, {0, 1, 2, 3}, count, eps};
  std::array<A3<Jet<3>>, 4> p;
  for (int a = 0; a < 4; ++a)
    for (int q = 0; q < 3; ++q)
      p[a][q] = a == vertex ? Jet<3>::variable(x12[3 * a + q], q) : Jet<3>::c(x12[3 * a + q]);
  Jet<3> j = c.energy<Jet<3>>(p, r, k);
  for (int i = 0; i < 3; ++i) g3[i] = j.g[i];
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) h9[3 * col + row] = j.h[row][col];
  return j.v;
}
double cvo_contact_value(int kind, const double* x12, int count, double eps, double r, double k) {
  Contact c{kind, {0, 1, 2, 3}, count, eps};
  std::array<A3<double>, 4> p;
  for (int a = 0; a < 4; ++a) p[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
  return c.energy<double>(p, r, k);
}

// ------------------------------------------------------------- geometry
// closest_triangle on a single triangle (0,1,2) with point p; feature kind
// 0 vertex / 1 edge / 2 face, index = vertex id, topology edge id or 0.
void cvo_closest_triangle(const double* p3, const double* tri9, double* out3, int* kind,
                          int* index) {
  std::vector<V3> x = {{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]},
                       {tri9[6], tri9[7], tri9[8]}};
  Topology top = Topology::build(3, {{0, 1, 2}});
  auto [q, f] = closest_triangle(V3{p3[0], p3[1], p3[2]}, 0, x, top);
  out3[0] = q.x;
  out3[1] = q.y;
  out3[2] = q.z;
  *kind = f.kind;
  *index = f.index;
}
void cvo_closest_segments(const double* a, const double* b, const double* c, const double* d,
                          double* s, double* t) {
  auto r = closest_segments(V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}, V3{c[0], c[1], c[2]},
                            V3{d[0], d[1], d[2]});
  *s = r.first;
  *t = r.second;
}
double cvo_bounds_distance_squared(const double* a6, const double* b6) {
  Bounds a{{a6[0], a6[1], a6[2]}, {a6[3], a6[4], a6[5]}};
  Bounds b{{b6[0], b6[1], b6[2]}, {b6[3], b6[4], b6[5]}};
  return a.distance_squared(b);
}
int cvo_segment_triangle(const double* a, const double* b, const double* tri9) {
  return segment_triangle(V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]},
                          {V3{tri9[0], tri9[1], tri9[2]}, V3{tri9[3], tri9[4], tri9[5]},
                           V3{tri9[6], tri9[7], tri9[8]}})
             ? 1
             : 0;
}

// ------------------------------------------------------------- world
struct cvo_world {
  OgcWorld w;
};
int cvo_world_create(const double* x, uint32_t nx, const uint32_t* tris, uint32_t ntri,
                     uint32_t moving, const uint32_t* excluded, uint32_t nexcl, cvo_world** out,
                     char* err, size_t errn) {
  return guarded(err, errn, [&] {
    std::vector<std::array<int, 3>> t(ntri);
    for (uint32_t i = 0; i < ntri; ++i)
      t[i] = {(int)tris[3 * i], (int)tris[3 * i + 1], (int)tris[3 * i + 2]};
    std::vector<std::pair<int, int>> ex(nexcl);
    for (uint32_t i = 0; i < nexcl; ++i) ex[i] = {(int)excluded[2 * i], (int)excluded[2 * i + 1]};
    *out = new cvo_world{OgcWorld::create(v3s(x, nx), t, (int)moving, ex)};
  });
}
cvo_world* cvo_world_clone(const cvo_world* w) { return new cvo_world{w->w}; }
void cvo_world_destroy(cvo_world* w) { delete w; }
int cvo_world_validate_start(cvo_world* w, const double* x, uint32_t nx, char* err, size_t errn) {
  return guarded(err, errn, [&] { w->w.validate_start(v3s(x, nx)); });
}
int cvo_world_detect(cvo_world* w, const double* x, uint32_t nx, double rq,
                     const cv_ogc_settings* s, int prune, cvo_contacts* out, double* centers,
                     char* err, size_t errn) {
  return guarded(err, errn, [&] {
    ContactSet c = w->w.detect(v3s(x, nx), rq, ogc_of(s), prune != 0);
    export_contacts(c, out);
    if (centers)
      for (size_t i = 0; i < c.centers.size(); ++i) {
        centers[3 * i] = c.centers[i].x;
        centers[3 * i + 1] = c.centers[i].y;
        centers[3 * i + 2] = c.centers[i].z;
      }
  });
}
// Canonical candidate set of the last detection after the padded current-AABB
// test: VF pairs (v,t) and EE pairs (e,f) in the operand order the narrow
// phase uses, sorted.  Returns the counts through nvf/nee.
int cvo_world_pruned_candidates(cvo_world* w, const double* x, uint32_t nx, double rq,
                                uint32_t* vf, uint64_t vf_cap, uint64_t* nvf, uint32_t* ee,
                                uint64_t ee_cap, uint64_t* nee) {
  auto xs = v3s(x, nx);
  auto& W = w->w;
  if (!W.candidates) return CV_ERR_INVALID_ARGUMENT;
  double scale = rq;
  for (auto& p : xs) scale = fmax_(scale, max_element(vabs(p)));
  double pr = rq + 64.0 * EPS64 * scale;
  double pr2 = pr * pr;
  std::vector<std::array<int, 2>> a, b;
  for (auto& p : W.candidates->vf)
    if (!(W.points.primitive_bounds[p[0]].distance_squared(W.triangles.primitive_bounds[p[1]]) >
          pr2))
      a.push_back(p);
  for (auto& p : W.candidates->ee)
    if (!(W.edges.primitive_bounds[p[0]].distance_squared(W.edges.primitive_bounds[p[1]]) > pr2))
      b.push_back(p);
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  *nvf = a.size();
  *nee = b.size();
  for (size_t i = 0; i < a.size() && i < vf_cap; ++i) {
    vf[2 * i] = (uint32_t)a[i][0];
    vf[2 * i + 1] = (uint32_t)a[i][1];
  }
  for (size_t i = 0; i < b.size() && i < ee_cap; ++i) {
    ee[2 * i] = (uint32_t)b[i][0];
    ee[2 * i + 1] = (uint32_t)b[i][1];
  }
  return CV_OK;
}
uint32_t cvo_world_edges(const cvo_world* w, uint32_t* out, uint32_t cap) {
  auto& e = w->w.topology.edges;
  for (size_t i = 0; i < e.size() && i < cap; ++i) {
    out[2 * i] = (uint32_t)e[i][0];
    out[2 * i + 1] = (uint32_t)e[i][1];
  }
  return (uint32_t)e.size();
}
int cvo_world_edge_feasible(const cvo_world* w, const double* p3, int e, const double* x,
                            uint32_t nx) {
  return edge_feasible(V3{p3[0], p3[1], p3[2]}, e, v3s(x, nx), w->w.topology) ? 1 : 0;
}
double cvo_contacts_value(const cvo_world* w, const double* x, uint32_t nx, const uint32_t* kind_ids,
                          const uint32_t* counts, const double* eps, uint64_t n, double r,
                          double k) {
  (void)w;
  auto xs = v3s(x, nx);
  double s = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    Contact c{(int)kind_ids[5 * i],
              {(int)kind_ids[5 * i + 1], (int)kind_ids[5 * i + 2], (int)kind_ids[5 * i + 3],
               (int)kind_ids[5 * i + 4]},
              (int)counts[i],
              eps[i]};
    s += c.value(xs, r, k);
  }
  return s;
}

// ------------------------------------------------------------- BVH
struct cvo_bvh {
  Bvh b;
};
static std::vector<Bounds> bounds_of(const double* b6, uint32_t n) {
  std::vector<Bounds> r(n);
  for (uint32_t i = 0; i < n; ++i)
    r[i] = {{b6[6 * i], b6[6 * i + 1], b6[6 * i + 2]}, {b6[6 * i + 3], b6[6 * i + 4], b6[6 * i + 5]}};
  return r;
}
cvo_bvh* cvo_bvh_create(const double* b6, const uint8_t* dynamic, uint32_t n) {
  std::vector<bool> d(n);
  for (uint32_t i = 0; i < n; ++i) d[i] = dynamic[i] != 0;
  return new cvo_bvh{Bvh::build(bounds_of(b6, n), d)};
}
void cvo_bvh_destroy(cvo_bvh* b) { delete b; }
void cvo_bvh_refit_dynamic(cvo_bvh* b, const double* b6, uint32_t n) {
  auto bs = bounds_of(b6, n);
  b->b.refit_dynamic([&](int i) { return bs[i]; });
}
uint64_t cvo_bvh_pairs_with(const cvo_bvh* a, const cvo_bvh* b, double radius, uint32_t* out,
                            uint64_t cap) {
  uint64_t k = 0;
  a->b.pairs(b->b, radius, false, [&](int i, int j) {
    if (k < cap) {
      out[2 * k] = (uint32_t)i;
      out[2 * k + 1] = (uint32_t)j;
    }
    k++;
  });
  return k;
}
uint64_t cvo_bvh_self_pairs(const cvo_bvh* a, double radius, uint32_t* out, uint64_t cap) {
  uint64_t k = 0;
  a->b.pairs(a->b, radius, true, [&](int i, int j) {
    if (k < cap) {
      out[2 * k] = (uint32_t)i;
      out[2 * k + 1] = (uint32_t)j;
    }
    k++;
  });
  return k;
}
uint64_t cvo_bvh_query(const cvo_bvh* a, const double* q6, int dynamic_only, uint32_t* out,
                       uint64_t cap) {
  Bounds q{{q6[0], q6[1], q6[2]}, {q6[3], q6[4], q6[5]}};
  uint64_t k = 0;
  a->b.query(q, dynamic_only != 0, [&](int i) {
    if (k < cap) out[k] = (uint32_t)i;
    k++;
  });
  return k;
}

// ------------------------------------------------------------- Newton
struct cvo_newton_settings {
  double dt;
  double gravity[3];
  double damping;
  uint32_t newton_iterations, pcg_iterations;
  double absolute_tolerance, relative_tolerance;
};
struct cvo_newton_metrics {
  uint32_t newton_iterations, pcg_iterations, line_search_backtracks, converged;
  double residual_rms_n, tolerance_n, objective_initial_j, objective_final_j, max_regularization;
};
struct cvo_frozen {
  uint64_t n_contacts;
  const uint32_t* kind_ids; /* [5n] */
  const uint32_t* counts;   /* [n]  */
  const double* eps;        /* [n]  */
  double radius, stiffness;
  uint32_t n_fixed;
  const double* fixed;      /* [3*n_fixed] avatar positions appended after the garment */
};
static SolverSettings newton_of(const cvo_newton_settings* s) {
  SolverSettings o;
  o.dt = s->dt;
  o.gravity = {s->gravity[0], s->gravity[1], s->gravity[2]};
  o.damping = s->damping;
  o.newton_iterations = (int)s->newton_iterations;
  o.pcg_iterations = (int)s->pcg_iterations;
  o.absolute_tolerance = s->absolute_tolerance;
  o.relative_tolerance = s->relative_tolerance;
  return o;
}
static FrozenContacts frozen_of(const cvo_frozen* f) {
  FrozenContacts o;
  for (uint64_t i = 0; i < f->n_contacts; ++i)
    o.contacts.push_back({(int)f->kind_ids[5 * i],
                          {(int)f->kind_ids[5 * i + 1], (int)f->kind_ids[5 * i + 2],
                           (int)f->kind_ids[5 * i + 3], (int)f->kind_ids[5 * i + 4]},
                          (int)f->counts[i],
                          f->eps[i]});
  o.radius = f->radius;
  o.stiffness = f->stiffness;
  o.fixed = v3s(f->fixed, f->n_fixed);
  return o;
}
int cvo_newton_step(const cv_garment* g, cv_state* st, const cvo_newton_settings* s,
                    const double* sphere /* cx,cy,cz,radius,stiffness or NULL */,
                    const cvo_frozen* frozen, cvo_newton_metrics* m, char* err, size_t errn) {
  return guarded(err, errn, [&] {
    Garment gg = garment_of(g);
    State state = state_of(st);
    Sphere sp;
    if (sphere) {
      sp.center = {sphere[0], sphere[1], sphere[2]};
      sp.radius = sphere[3];
      sp.stiffness = sphere[4];
    }
    FrozenContacts fz;
    if (frozen) fz = frozen_of(frozen);
    NewtonMetrics r = newton_step(gg, state, newton_of(s), sphere ? &sp : nullptr,
                                  frozen ? &fz : nullptr);
    for (uint32_t i = 0; i < st->n; ++i)
      for (int k = 0; k < 3; ++k) {
        st->x[3 * i + k] = state.x[i][k];
        st->v[3 * i + k] = state.v[i][k];
      }
    st->time = state.time;
    if (m) {
      m->newton_iterations = (uint32_t)r.newton_iterations;
      m->pcg_iterations = (uint32_t)r.pcg_iterations;
      m->line_search_backtracks = (uint32_t)r.line_search_backtracks;
      m->converged = r.converged ? 1 : 0;
      m->residual_rms_n = r.residual_rms_n;
      m->tolerance_n = r.tolerance_n;
      m->objective_initial_j = r.objective_initial_j;
      m->objective_final_j = r.objective_final_j;
      m->max_regularization = r.max_regularization;
    }
  });
}
double cvo_newton_objective(const cv_garment* g, const cv_state* old, const double* x,
                            const cvo_newton_settings* s, const double* sphere,
                            const cvo_frozen* frozen) {
  Garment gg = garment_of(g);
  State o = state_of(old);
  Sphere sp;
  if (sphere) {
    sp.center = {sphere[0], sphere[1], sphere[2]};
    sp.radius = sphere[3];
    sp.stiffness = sphere[4];
  }
  FrozenContacts fz;
  if (frozen) fz = frozen_of(frozen);
  return newton_objective(gg, o, v3s(x, old->n), newton_of(s), sphere ? &sp : nullptr,
                          frozen ? &fz : nullptr);
}

// Diagnostics (src/garment/diagnostics.rs:7-28, template.rs:536-538).
void cvo_extension_percent(const cv_garment* g, const double* x, double* out2) {
  Garment gg = garment_of(g);
  auto e = extension_percent(gg, v3s(x, gg.n));
  out2[0] = e[0];
  out2[1] = e[1];
}
double cvo_max_seam_gap(const cv_garment* g, const double* x) {
  Garment gg = garment_of(g);
  return gg.max_seam_gap(v3s(x, gg.n));
}

}  // extern "C"

// ------------------------------------------------------------ prediction
extern "C" {
struct cvo_prediction {
  Prediction p;
};
cvo_prediction* cvo_prediction_create() { return new cvo_prediction(); }
void cvo_prediction_destroy(cvo_prediction* p) { delete p; }
void cvo_prediction_guess(const cvo_prediction* p, const cv_state* st, const cv_settings* s, double* out) {
  State state = state_of(st);
  auto g = p->p.guess(state, settings_of(s));
  for (size_t i = 0; i < g.size(); ++i) {
    out[3 * i] = g[i].x;
    out[3 * i + 1] = g[i].y;
    out[3 * i + 2] = g[i].z;
  }
}
void cvo_prediction_record(cvo_prediction* p, const double* old_v, const cv_state* st, const cv_settings* s) {
  p->p.record(v3s(old_v, st->n), state_of(st), settings_of(s));
}
}

// Colouring of `mode` (color_mode; 0 is the reference's greedy,
// src/vbd/coloring.rs:5-43) on an explicit conflict graph (CSR) and contact
// rows (kind, ids[4], count), with an optional grab.
extern "C" uint32_t cvo_colors(uint32_t n, const uint32_t* g_off, const uint32_t* g_nb, const uint32_t* kind_ids,
                               const uint32_t* counts, uint64_t m, const cv_grab* grab, uint32_t* color_of,
                               int mode) {
  std::vector<std::vector<int>> graph(n);
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t k = g_off[v]; k < g_off[v + 1]; ++k) graph[v].push_back((int)g_nb[k]);
  ContactSet c;
  c.incident.assign(n, {});
  for (uint64_t i = 0; i < m; ++i) {
    Contact ct{(int)kind_ids[5 * i],
               {(int)kind_ids[5 * i + 1], (int)kind_ids[5 * i + 2], (int)kind_ids[5 * i + 3],
                (int)kind_ids[5 * i + 4]},
               (int)counts[i],
               0.0};
    c.contacts.push_back(ct);
    for (int k = 0; k < ct.count; ++k)
      if (ct.vertices[k] < (int)n) c.incident[ct.vertices[k]].push_back((int)i);
  }
  GrabT g;
  if (grab) g = grab_of(grab);
  auto cols = color_mode(graph, c, grab ? &g : nullptr, mode);
  for (size_t k = 0; k < cols.size(); ++k)
    for (int v : cols[k]) color_of[v] = (uint32_t)k;
  return (uint32_t)cols.size();
}
