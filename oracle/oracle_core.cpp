// ORACLE — test infrastructure only (see oracle_core.hpp).  Restates
// src/ogc/world.rs detection, src/vbd.rs stepping, src/garment/diagnostics.rs
// and src/solver.rs Newton in f64.
#include "oracle_core.hpp"

#include <chrono>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cvo {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// src/ogc/world.rs:242-307
void OgcWorld::validate_start(const std::vector<V3>& x) {
  refit(x);
  const auto& tris = topology.triangles;
  for (size_t i = 0; i < tris.size(); ++i) {
    auto t = tris[i];
    bool t_fixed = t[0] >= moving && t[1] >= moving && t[2] >= moving;
    int hit = -1;
    triangles.query(Bounds::of({x[t[0]], x[t[1]], x[t[2]]}).inflate(1e-10), t_fixed, [&](int j) {
      auto q = tris[j];
      bool q_fixed = q[0] >= moving && q[1] >= moving && q[2] >= moving;
      if (j <= (int)i || (t_fixed && q_fixed)) return;
      for (int a : t)
        for (int b : q)
          if (excluded.count({std::min(a, b), std::max(a, b)})) return;
      auto contains = [](const std::array<int, 3>& s, int v) {
        return s[0] == v || s[1] == v || s[2] == v;
      };
      for (int k = 0; k < 3; ++k) {
        bool h1 = !contains(q, t[k]) && !contains(q, t[(k + 1) % 3]) &&
                  segment_triangle(x[t[k]], x[t[(k + 1) % 3]], {x[q[0]], x[q[1]], x[q[2]]});
        bool h2 = !contains(t, q[k]) && !contains(t, q[(k + 1) % 3]) &&
                  segment_triangle(x[q[k]], x[q[(k + 1) % 3]], {x[t[0]], x[t[1]], x[t[2]]});
        if (h1 || h2) hit = j;
      }
    });
    if (hit >= 0)
      throw std::runtime_error("I|OGC initial intersection between faces " + std::to_string(i) +
                               " and " + std::to_string(hit) +
                               "; reset/rearrange before simulating");
  }
  OgcSettings s;
  s.radius = 1e-5;
  ContactSet c = detect(x, 1e-5, s);
  if (c.minimum_distance <= 1e-10)
    throw std::runtime_error(
        "U|OGC initial touching/coplanar features: separation must be positive");
}

// src/ogc/world.rs:316-557 (one narrow-phase lane; the reference merges its
// lanes with a min and a canonical sort, so the lane count never shows).
ContactSet OgcWorld::detect(const std::vector<V3>& x, double rq, const OgcSettings& s,
                            bool prune) {
  s.validate();
  if (!std::isfinite(rq) || rq < s.radius)
    throw std::runtime_error("Q|OGC query radius must cover the contact radius");
  double t0 = now_ms();
  refit(x);
  double refit_ms = now_ms() - t0;
  double t1 = now_ms();
  bool rebuilt = !candidates || !candidates->covers(x, moving, rq);
  if (rebuilt) {
    Candidates c;
    c.centers.assign(x.begin(), x.begin() + moving);
    c.radius = fmax_(std::ceil(rq / s.radius) * s.radius, rq);
    c.skin = s.radius;
    double search = c.radius + 2.0 * c.skin + 1e-12;
    const auto& top = topology;
    points.pairs(triangles, search, false, [&](int v, int t) {
      auto& tr = top.triangles[t];
      if (!(adjacent(v, tr[0]) || adjacent(v, tr[1]) || adjacent(v, tr[2]))) c.vf.push_back({v, t});
    });
    edges.pairs(edges, search, true, [&](int e, int f) {
      auto ids = top.edges[e];
      auto js = top.edges[f];
      bool adj = false;
      for (int i : ids)
        for (int j : js) adj = adj || adjacent(i, j);
      if (!adj) {
        if (ids[0] >= moving && ids[1] >= moving)
          c.ee.push_back({f, e});
        else
          c.ee.push_back({e, f});
      }
    });
    candidates = std::move(c);
  }
  double candidates_ms = now_ms() - t1;
  double t2 = now_ms();
  double scale = rq;
  for (auto& p : x) scale = fmax_(scale, max_element(vabs(p)));
  double pr = rq + 64.0 * EPS64 * scale;
  double padded_radius2 = pr * pr;
  const auto& top = topology;
  std::vector<double> distance(x.size(), rq);
  std::vector<Contact> found;
  uint64_t vf_count = 0, ee_count = 0, rejected = 0;
  auto add = [&](int kind, std::array<int, 4> ids, int count, double eps) {
    if (kind == PointPoint && ids[0] > ids[1]) std::swap(ids[0], ids[1]);
    if (kind == PointEdge && ids[1] > ids[2]) std::swap(ids[1], ids[2]);
    if (kind == PointFace) std::sort(ids.begin() + 1, ids.end());
    if (kind == EdgeEdge) {
      std::sort(ids.begin(), ids.begin() + 2);
      std::sort(ids.begin() + 2, ids.end());
      if (ids[0] > ids[2]) {
        std::swap(ids[0], ids[2]);
        std::swap(ids[1], ids[3]);
      }
    }
    found.push_back({kind, ids, count, eps});
  };
  for (auto& pair : candidates->vf) {
    int v = pair[0], t = pair[1];
    if (prune && points.primitive_bounds[v].distance_squared(triangles.primitive_bounds[t]) >
                     padded_radius2) {
      rejected++;
      continue;
    }
    auto ids = top.triangles[t];
    auto [p, f] = closest_triangle(x[v], t, x, top);
    double d = cvo::distance(x[v], p);
    if (d >= rq) continue;
    vf_count++;
    distance[v] = fmin_(distance[v], d);
    for (int i : ids) distance[i] = fmin_(distance[i], d);
    if (f.kind == FVertex) {
      if (vertex_feasible(x[v], f.index, x, top)) add(PointPoint, {v, f.index, 0, 0}, 2, 0.0);
    } else if (f.kind == FEdge) {
      if (edge_feasible(x[v], f.index, x, top))
        add(PointEdge, {v, top.edges[f.index][0], top.edges[f.index][1], 0}, 3, 0.0);
    } else {
      add(PointFace, {v, ids[0], ids[1], ids[2]}, 4, 0.0);
    }
  }
  for (auto& pair : candidates->ee) {
    int e = pair[0], f = pair[1];
    if (prune &&
        edges.primitive_bounds[e].distance_squared(edges.primitive_bounds[f]) > padded_radius2) {
      rejected++;
      continue;
    }
    auto ids = top.edges[e];
    auto js = top.edges[f];
    V3 a = x[ids[0]], b = x[ids[1]], c = x[js[0]], d = x[js[1]];
    auto [sv, tv] = closest_segments(a, b, c, d);
    V3 p = a + (b - a) * sv;
    V3 q = c + (d - c) * tv;
    double dist = cvo::distance(p, q);
    if (dist >= rq) continue;
    ee_count++;
    for (int v : {ids[0], ids[1], js[0], js[1]}) distance[v] = fmin_(distance[v], dist);
    int end_a = sv <= 0.0 ? ids[0] : (sv >= 1.0 ? ids[1] : -1);
    int end_b = tv <= 0.0 ? js[0] : (tv >= 1.0 ? js[1] : -1);
    if ((end_a >= 0 && !vertex_feasible(q, end_a, x, top)) ||
        (end_b >= 0 && !vertex_feasible(p, end_b, x, top)))
      continue;
    if (end_a >= 0 && end_b >= 0)
      add(PointPoint, {end_a, end_b, 0, 0}, 2, 0.0);
    else if (end_a >= 0)
      add(PointEdge, {end_a, js[0], js[1], 0}, 3, 0.0);
    else if (end_b >= 0)
      add(PointEdge, {end_b, ids[0], ids[1], 0}, 3, 0.0);
    else
      add(EdgeEdge, {ids[0], ids[1], js[0], js[1]}, 4,
          1e-3 * rest_edge_length2[e] * rest_edge_length2[f]);
  }
  std::sort(found.begin(), found.end(),
            [](const Contact& a, const Contact& b) { return a.key() < b.key(); });
  std::vector<Contact> contacts;
  for (auto& c : found)
    if (contacts.empty() || contacts.back().key() != c.key()) contacts.push_back(c);
  ContactSet out;
  out.active_contacts = 0;
  for (auto& c : contacts)
    if (c.value(x, s.radius, s.stiffness) > 0.0) out.active_contacts++;
  out.incident.assign(moving, {});
  for (size_t ci = 0; ci < contacts.size(); ++ci)
    for (int k = 0; k < contacts[ci].count; ++k)
      if (contacts[ci].vertices[k] < moving) out.incident[contacts[ci].vertices[k]].push_back((int)ci);
  double md = rq;
  for (int i = 0; i < moving; ++i) md = fmin_(md, distance[i]);
  out.profile.refit_ms = refit_ms;
  out.profile.candidates_ms = candidates_ms;
  out.profile.narrow_ms = now_ms() - t2;
  out.profile.candidate_pairs = candidates->vf.size() + candidates->ee.size();
  out.profile.aabb_rejections = rejected;
  out.contacts = std::move(contacts);
  out.centers.assign(x.begin(), x.begin() + moving);
  out.radii.resize(moving);
  for (int i = 0; i < moving; ++i) out.radii[i] = s.relaxation * distance[i];
  out.minimum_distance = md;
  out.vertex_face_candidates = vf_count;
  out.edge_edge_candidates = ee_count;
  out.broad_phase_rebuilt = rebuilt;
  return out;
}

// src/garment/diagnostics.rs:7-28 with energy::stretches (src/energy.rs:80-114)
std::array<double, 2> extension_percent(const Garment& g, const std::vector<V3>& x) {
  std::vector<std::pair<double, double>> values;
  values.reserve(g.triangles.size());
  for (auto& t : g.triangles) {
    const Material& m = g.materials[t.material];
    V3 ds0 = x[t.v[1]] - x[t.v[0]], ds1 = x[t.v[2]] - x[t.v[0]];
    V3 f0 = ds0 * t.inv_dm[0] + ds1 * t.inv_dm[1];
    V3 f1 = ds0 * t.inv_dm[2] + ds1 * t.inv_dm[3];
    double s, c;
    grain_sin_cos(m.grain_degrees, s, c);
    V3 w = f0 * c + f1 * s;
    V3 v = (-f0) * s + f1 * c;
    double s0 = length(w), s1 = length(v);
    values.push_back({100.0 * fmax_(fmax_(s0, s1) - 1.0, 0.0), t.area});
  }
  std::sort(values.begin(), values.end(),
            [](auto& a, auto& b) { return a.first < b.first; });
  double total = 0.0;
  for (auto& v : values) total += v.second;
  double area = 0.0, p95 = 0.0;
  for (auto& v : values) {
    area += v.second;
    if (area >= 0.95 * total) {
      p95 = v.first;
      break;
    }
  }
  return {p95, values.empty() ? 0.0 : values.back().first};
}

// src/vbd.rs:101-192
VbdSolver VbdSolver::create(const Garment& g, const std::vector<V3>* avatar_v,
                            const std::vector<std::array<int, 3>>* avatar_t, const State& s) {
  VbdSolver sv;
  int n = g.n;
  if ((int)s.x.size() != n || (int)s.v.size() != n)
    throw std::runtime_error("M|VBD state dimensions do not match garment");
  for (int i = 0; i < n; ++i)
    if (g.pins[i] && distance(s.x[i], *g.pins[i]) > 1e-10)
      throw std::runtime_error(
          "K|VBD pinned positions must match their constraints before initialization");
  std::vector<V3> x = s.x;
  std::vector<std::array<int, 3>> faces;
  for (auto& t : g.triangles) faces.push_back(t.v);
  if (avatar_v) {
    // SurfaceMesh::validate (src/avatar.rs:78-97)
    if (avatar_v->empty() || !avatar_t || avatar_t->empty())
      throw std::runtime_error("W|Empty/non-finite avatar");
    for (auto& p : *avatar_v)
      if (!finite(p)) throw std::runtime_error("W|Empty/non-finite avatar");
    for (auto& t : *avatar_t) {
      for (int k = 0; k < 3; ++k)
        if (t[k] < 0 || t[k] >= (int)avatar_v->size())
          throw std::runtime_error("W|Invalid avatar index");
      V3 a = (*avatar_v)[t[0]], b = (*avatar_v)[t[1]], c = (*avatar_v)[t[2]];
      if (length2(cross(b - a, c - a)) < 1e-24)
        throw std::runtime_error("W|Degenerate avatar triangle");
    }
    x.insert(x.end(), avatar_v->begin(), avatar_v->end());
    for (auto& t : *avatar_t) faces.push_back({t[0] + n, t[1] + n, t[2] + n});
  }
  std::set<std::pair<int, int>> excluded;
  for (size_t si = 0; si < g.stitch_points.size(); ++si) {
    auto& terms = g.stitch_points[si].terms;
    for (auto& [a, wa] : terms)
      for (auto& [b, wb] : terms)
        if (wa * wb < 0.0) excluded.insert({std::min(a, b), std::max(a, b)});
    sv.stitches.push_back({terms, g.stitch_stiffness[si]});
  }
  std::vector<std::pair<int, int>> ex(excluded.begin(), excluded.end());
  sv.world = OgcWorld::create(x, faces, n, ex);
  sv.world.validate_start(x);
  sv.tri_of.assign(n, {});
  sv.hinge_of.assign(n, {});
  sv.stitch_of.assign(n, {});
  std::vector<std::set<int>> graph(n);
  auto connect = [&](const int* ids, int k) {
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b)
        if (ids[a] != ids[b]) graph[ids[a]].insert(ids[b]);
  };
  for (size_t i = 0; i < g.triangles.size(); ++i) {
    for (int v : g.triangles[i].v) sv.tri_of[v].push_back((int)i);
    connect(g.triangles[i].v.data(), 3);
  }
  for (size_t i = 0; i < g.hinges.size(); ++i) {
    for (int v : g.hinges[i].v) sv.hinge_of[v].push_back((int)i);
    connect(g.hinges[i].v.data(), 4);
  }
  for (size_t i = 0; i < sv.stitches.size(); ++i) {
    std::vector<int> ids;
    for (auto& t : sv.stitches[i].terms) ids.push_back(t.first);
    for (int v : ids) sv.stitch_of[v].push_back((int)i);
    connect(ids.data(), (int)ids.size());
  }
  sv.graph.resize(n);
  for (int i = 0; i < n; ++i) sv.graph[i].assign(graph[i].begin(), graph[i].end());
  sv.garment = g;
  return sv;
}

// src/vbd.rs:196-248
template <class S>
S VbdSolver::local(int v, const A3<S>& p, const std::vector<V3>& x, const ContactSet& contact,
                   const VbdSettings& st, const GrabT* grab) const {
  using T = Sc<S>;
  auto get = [&](int i) -> A3<S> {
    if (i == v) return p;
    return {T::c(x[i].x), T::c(x[i].y), T::c(x[i].z)};
  };
  S value = T::c(0.0);
  for (int i : tri_of[v]) {
    auto& t = garment.triangles[i];
    value = value + membrane<S>({get(t.v[0]), get(t.v[1]), get(t.v[2])}, t,
                                garment.materials[t.material]);
  }
  for (int i : hinge_of[v]) {
    auto& h = garment.hinges[i];
    value = value + bending<S>({get(h.v[0]), get(h.v[1]), get(h.v[2]), get(h.v[3])}, h,
                               garment.materials[h.material]);
  }
  for (int i : stitch_of[v]) {
    auto& s = stitches[i];
    A3<S> d = {T::c(0.0), T::c(0.0), T::c(0.0)};
    for (auto& [j, w] : s.terms) {
      A3<S> q = get(j);
      for (int k = 0; k < 3; ++k) d[k] = d[k] + T::c(w) * q[k];
    }
    value = value + T::c(0.5 * s.stiffness) * sdot(d, d);
  }
  for (int i : contact.incident[v]) {
    auto& c = contact.contacts[i];
    value = value + c.energy<S>({get(c.vertices[0]), get(c.vertices[1]), get(c.vertices[2]),
                                 get(c.vertices[3])},
                                st.contact.radius, st.contact.stiffness);
  }
  if (grab && grab->contains(v)) {
    A3<S> d = {T::c(-grab->target.x), T::c(-grab->target.y), T::c(-grab->target.z)};
    for (int j = 0; j < 3; ++j) {
      A3<S> q = get(grab->vertices[j]);
      for (int k = 0; k < 3; ++k) d[k] = d[k] + T::c(grab->weights[j]) * q[k];
    }
    value = value + T::c(0.5 * grab->stiffness) * sdot(d, d);
  }
  return value;
}
template double VbdSolver::local<double>(int, const A3<double>&, const std::vector<V3>&,
                                         const ContactSet&, const VbdSettings&,
                                         const GrabT*) const;
template Jet<3> VbdSolver::local<Jet<3>>(int, const A3<Jet<3>>&, const std::vector<V3>&,
                                         const ContactSet&, const VbdSettings&,
                                         const GrabT*) const;

double VbdSolver::true_residual(const std::vector<V3>& x, const std::vector<V3>& target,
                                const std::vector<double>& inertia, const ContactSet& c,
                                const VbdSettings& st, const GrabT* grab) const {
  int n = garment.n;
  double sum = 0.0;
  int free = 0;
  for (int v = 0; v < n; ++v) {
    if (garment.pins[v]) continue;
    free++;
    Jet<3> j = local<Jet<3>>(
        v, {Jet<3>::variable(x[v].x, 0), Jet<3>::variable(x[v].y, 1), Jet<3>::variable(x[v].z, 2)},
        x, c, st, grab);
    V3 g = V3{j.g[0], j.g[1], j.g[2]} + (x[v] - target[v]) * inertia[v];
    sum += length2(g);
  }
  return std::sqrt(sum / (3.0 * std::max(free, 1)));
}

// src/vbd.rs:250-422
VbdMetrics VbdSolver::step(State& state, const VbdSettings& st, const GrabT* grab) {
  st.validate();
  int n = garment.n;
  bool bad = (int)state.x.size() != n || (int)state.v.size() != n || !std::isfinite(state.time);
  if (!bad)
    for (int i = 0; i < n; ++i) bad = bad || !finite(state.x[i]) || !finite(state.v[i]);
  if (bad) throw std::runtime_error("V|Invalid VBD state");
  if (grab) {
    bool g_bad = !finite(grab->target) || !std::isfinite(grab->stiffness) || grab->stiffness <= 0.0;
    double wsum = 0.0;
    for (int k = 0; k < 3; ++k) {
      g_bad = g_bad || grab->vertices[k] < 0 || grab->vertices[k] >= n;
      g_bad = g_bad || !std::isfinite(grab->weights[k]) || grab->weights[k] < 0.0;
      wsum += grab->weights[k];
    }
    if (g_bad || std::fabs(wsum - 1.0) > 1e-8) throw std::runtime_error("G|Invalid mouse attachment");
  }
  double start = now_ms();
  VbdMetrics m;
  m.dt_s = st.dt;
  m.iterations = st.iterations;
  double dt = st.dt;
  std::vector<V3> old_velocity = state.v;
  std::vector<V3> guess = next_guess.size() == (size_t)n ? next_guess : prediction.guess(state, st);
  next_guess.clear();
  double damp = 1.0 + st.damping * dt;
  std::vector<V3> target(n);
  std::vector<double> inertia(n);
  for (int i = 0; i < n; ++i)
    target[i] = state.x[i] + (state.v[i] * dt + st.gravity * dt * dt) / damp;
  for (int i = 0; i < n; ++i) inertia[i] = garment.masses[i] * damp / (dt * dt);
  double motion = 0.0;
  for (int i = 0; i < n; ++i) motion = fmax_(motion, distance(target[i], state.x[i]));
  double rq = std::clamp(st.contact.radius + motion, 2.0 * st.contact.radius,
                         4.0 * st.contact.radius);
  std::vector<V3> x = world.positions(state.x);
  double td = now_ms();
  ContactSet contacts = world.detect(x, rq, st.contact);
  if (contacts.minimum_distance <= 1e-12)
    throw std::runtime_error("Z|OGC zero separation; step rejected, reset the arrangement");
  m.detection_ms += now_ms() - td;
  m.detection_updates += 1;
  m.broad_phase_rebuilds += contacts.broad_phase_rebuilt ? 1 : 0;
  auto addp = [&](const DetectionProfile& p) {
    m.profile.refit_ms += p.refit_ms;
    m.profile.candidates_ms += p.candidates_ms;
    m.profile.narrow_ms += p.narrow_ms;
    m.profile.candidate_pairs += p.candidate_pairs;
    m.profile.aabb_rejections += p.aabb_rejections;
  };
  addp(contacts.profile);
  std::vector<int> sl;
  if (st.fast_coloring) sl = smallest_last(graph);
  auto step_colors = [&] {
    return st.fast_coloring ? fast_colors(graph, contacts, grab, sl) : colors(graph, contacts, grab);
  };
  auto cols = step_colors();
  uint64_t clipped = 0;
  for (int i = 0; i < n; ++i) {
    if (!garment.pins[i]) {
      auto [p, b] = contacts.truncate(i, guess[i]);
      x[i] = p;
      clipped += b ? 1 : 0;
    }
  }
  last_gsq.assign(n, 0.0);
  struct Upd {
    V3 p;
    bool c;
    double r;
  };
  std::vector<Upd> upd;
  for (int it = 0; it < st.iterations; ++it) {
    if (clipped > 0 && (double)clipped >= st.contact.refresh_fraction * n) {
      double t = now_ms();
      contacts = world.detect(x, rq, st.contact);
      cols = step_colors();
      m.detection_ms += now_ms() - t;
      m.detection_updates += 1;
      m.broad_phase_rebuilds += contacts.broad_phase_rebuilt ? 1 : 0;
      addp(contacts.profile);
    }
    clipped = 0;
    double residual = 0.0;
    for (auto& color : cols) {
      upd.assign(color.size(), {});
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
      for (long ci = 0; ci < (long)color.size(); ++ci) {
        int v = color[ci];
        if (garment.pins[v]) {
          upd[ci] = {x[v], false, 0.0};
          continue;
        }
        Jet<3> jet = local<Jet<3>>(v,
                                   {Jet<3>::variable(x[v].x, 0), Jet<3>::variable(x[v].y, 1),
                                    Jet<3>::variable(x[v].z, 2)},
                                   x, contacts, st, grab);
        V3 g = V3{jet.g[0], jet.g[1], jet.g[2]} + (x[v] - target[v]) * inertia[v];
        M3 h;
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) h.c[c][r] = jet.h[r][c] + (r == c ? inertia[v] : 0.0);
        double lower = INF;
        for (int i = 0; i < 3; ++i) {
          double off = 0.0;
          for (int j = 0; j < 3; ++j)
            if (j != i) off += std::fabs(h.c[i][j]);
          lower = fmin_(lower, h.c[i][i] - off);
        }
        if (lower < 1e-8) {
          double shift = 1e-8 - lower + 0.01 * inertia[v];
          for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) h.c[c][r] = h.c[c][r] + (r == c ? shift : 0.0);
        }
        V3 delta = -m3_mul(m3_inverse(h), g);
        double old = jet.v + 0.5 * inertia[v] * length2(x[v] - target[v]);
        V3 accepted = x[v];
        bool truncated = false;
        double scale = 1.0;
        for (int attempt = 0; attempt < 10; ++attempt) {
          V3 cand = x[v] + delta * scale;
          auto [cc, bound] = contacts.truncate(v, cand);
          truncated = truncated || bound;
          double e = local<double>(v, {cc.x, cc.y, cc.z}, x, contacts, st, grab) +
                     0.5 * inertia[v] * length2(cc - target[v]);
          if (finite(cc) && std::isfinite(e) && e <= old + 1e-14) {
            accepted = cc;
            break;
          }
          scale *= 0.5;
        }
        upd[ci] = {accepted, truncated, length2(g)};
      }
      for (size_t ci = 0; ci < color.size(); ++ci) {
        int v = color[ci];
        x[v] = upd[ci].p;
        clipped += upd[ci].c ? 1 : 0;
        residual += upd[ci].r;
        last_gsq[v] = upd[ci].r;
      }
    }
    m.truncated_vertices += clipped;
    m.residual_rms_n = std::sqrt(residual / (3.0 * n));
    m.iterations_run = it + 1;
    const int every = (int)std::max<uint32_t>(st.residual_interval, 1);
    if (st.residual_tolerance > 0.0 && ((it + 1) % every == 0 || it + 1 == st.iterations)) {
      // Extension: true residual after the sweep, against the current contacts.
      m.true_residual_rms_n = true_residual(x, target, inertia, contacts, st, grab);
      if (m.true_residual_rms_n <= st.residual_tolerance) break;
    }
  }
  for (auto& p : x)
    if (!finite(p)) throw std::runtime_error("N|Non-finite VBD result; state unchanged");
  for (int i = 0; i < n; ++i)
    if (distance(x[i], contacts.centers[i]) > contacts.radii[i] + 1e-10)
      throw std::runtime_error("R|OGC trust-region invariant failed");
  last_x = x;
  last_contacts = contacts;
  last_colors = cols;
  for (int i = 0; i < n; ++i) {
    state.v[i] = (x[i] - state.x[i]) / dt;
    state.x[i] = x[i];
  }
  state.time += dt;
  prediction.record(old_velocity, state, st);
  m.time_s = state.time;
  m.contacts = contacts.active_contacts;
  m.contact_terms = contacts.contacts.size();
  m.minimum_distance_m = contacts.minimum_distance;
  m.max_seam_gap_m = garment.max_seam_gap(state.x);
  auto ext = extension_percent(garment, state.x);
  m.extension_p95 = ext[0];
  m.extension_max = ext[1];
  m.colors = (uint32_t)cols.size();
  m.step_ms = now_ms() - start;
  m.solve_ms = m.step_ms - m.detection_ms;
  return m;
}

// ----------------------------------------------------------------- Newton
// src/solver.rs:142-231 (System, pcg, rms) and :233-647 (ReferenceSolver).
namespace {
struct Element {  // src/energy.rs:48-67
  std::vector<int> dofs;
  double energy;
  std::vector<double> gradient, hessian;
};
template <int N>
Element from_jet(const int* vertices, int nv, const Jet<N>& j) {
  Element e;
  for (int a = 0; a < nv; ++a)
    for (int k = 0; k < 3; ++k) e.dofs.push_back(3 * vertices[a] + k);
  e.energy = j.v;
  e.gradient.assign(j.g, j.g + N);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k) e.hessian.push_back(j.h[i][k]);
  return e;
}
struct System {
  std::vector<Element> elements;
  std::vector<double> gradient, mass_diagonal;
  std::vector<bool> fixed;
  double energy = 0;
  std::vector<double> diagonal(double shift) const {
    std::vector<double> d(mass_diagonal.size());
    for (size_t i = 0; i < d.size(); ++i) d[i] = mass_diagonal[i] + shift;
    for (auto& e : elements) {
      size_t n = e.dofs.size();
      for (size_t i = 0; i < n; ++i) d[e.dofs[i]] += e.hessian[i * n + i];
    }
    for (size_t i = 0; i < d.size(); ++i) d[i] = fixed[i] ? 1.0 : fmax_(d[i], 1e-12);
    return d;
  }
  std::vector<double> apply(const std::vector<double>& x, double shift) const {
    std::vector<double> out(x.size());
    for (size_t i = 0; i < x.size(); ++i) out[i] = x[i] * (mass_diagonal[i] + shift);
    for (auto& e : elements) {
      size_t n = e.dofs.size();
      for (size_t i = 0; i < n; ++i) {
        int gi = e.dofs[i];
        if (fixed[gi]) continue;
        for (size_t j = 0; j < n; ++j) {
          int gj = e.dofs[j];
          if (!fixed[gj]) out[gi] += e.hessian[i * n + j] * x[gj];
        }
      }
    }
    for (size_t i = 0; i < out.size(); ++i)
      if (fixed[i]) out[i] = x[i];
    return out;
  }
  static double vdot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  std::pair<std::optional<std::vector<double>>, int> pcg(double shift, int limit) const {
    size_t n = gradient.size();
    auto diag = diagonal(shift);
    std::vector<double> x(n, 0.0), r(n), z(n), p;
    for (size_t i = 0; i < n; ++i) r[i] = -gradient[i];
    for (size_t i = 0; i < n; ++i) z[i] = r[i] / diag[i];
    p = z;
    double rz = vdot(r, z);
    double tol = 1e-5 * std::sqrt(vdot(r, r));
    for (int it = 0; it < limit; ++it) {
      auto ap = apply(p, shift);
      double curv = vdot(p, ap);
      if (!std::isfinite(curv) || curv <= 0.0) return {std::nullopt, it + 1};
      double alpha = rz / curv;
      for (size_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * ap[i];
      }
      if (std::sqrt(vdot(r, r)) <= tol) return {x, it + 1};
      for (size_t i = 0; i < n; ++i) z[i] = r[i] / diag[i];
      double next = vdot(r, z);
      double beta = next / rz;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      rz = next;
    }
    return {x, limit};
  }
};
double rms(const std::vector<double>& g, const std::vector<bool>& fixed) {
  size_t free = 0;
  for (bool f : fixed) free += f ? 0 : 1;
  return std::sqrt(System::vdot(g, g) / (double)std::max<size_t>(free, 1));
}
// src/contact.rs:69-78
template <class S>
S sphere_energy(const Sphere& sp, const A3<S>& x, double thickness, double area) {
  using T = Sc<S>;
  A3<S> p = {x[0] - T::c(sp.center.x), x[1] - T::c(sp.center.y), x[2] - T::c(sp.center.z)};
  S d = T::sqrt(sdot(p, p));
  S gap = d - T::c(sp.radius + thickness * 0.5);
  if (T::value(gap) >= 0.0) return T::c(0.0);
  return T::c(0.5 * sp.stiffness * area) * gap * gap;
}
double sphere_penetration(const Sphere& sp, V3 x, double thickness) {
  return fmax_(sp.radius + thickness * 0.5 - distance(x, sp.center), 0.0);
}
// Per-vertex thickness/area inputs: the single-material Problem uses
// materials[0]; garments use vertex_material (src/solver.rs:105-108).  The
// build's garments carry one material per triangle, so the thickness of the
// first incident triangle's material is used; every config shares thickness.
bool geometry_valid(const Garment& g, const std::vector<V3>& x, const Sphere* sp) {  // :235-242
  for (auto& p : x)
    if (!finite(p) || (sp && !(distance2(p, sp->center) > 1e-20))) return false;
  for (auto& t : g.triangles) {
    V3 a = x[t.v[0]], b = x[t.v[1]], c = x[t.v[2]];
    if (!(length(cross(b - a, c - a)) > 2.0 * t.area * 1e-6)) return false;
  }
  return true;
}
double inertial(const Garment& g, const State& old, const std::vector<V3>& x,
                const SolverSettings& s) {  // :243-254
  double sum = 0.0;
  for (size_t i = 0; i < x.size(); ++i) {
    V3 y = old.x[i] + s.dt * old.v[i];
    sum += g.masses[i] * (0.5 * length2(x[i] - y) / (s.dt * s.dt) +
                          0.5 * s.damping / s.dt * length2(x[i] - old.x[i]) - dot(s.gravity, x[i]));
  }
  return sum;
}
std::vector<double> rest_area_of(const Garment& g) {
  // Lumped rest area = mass / density of the incident material (model.rs:212-214).
  std::vector<double> a(g.n, 0.0);
  for (auto& t : g.triangles)
    for (int v : t.v) a[v] += t.area / 3.0;
  return a;
}
std::vector<V3> frozen_positions(const std::vector<V3>& x, const FrozenContacts* fz) {
  std::vector<V3> all(x);
  if (fz) all.insert(all.end(), fz->fixed.begin(), fz->fixed.end());
  return all;
}
}  // namespace

double newton_objective(const Garment& g, const State& old, const std::vector<V3>& x,
                        const SolverSettings& s, const Sphere* sphere,
                        const FrozenContacts* frozen) {  // :255-296
  if (!geometry_valid(g, x, sphere)) return INF;
  double membrane_e = 0.0, bending_e = 0.0;
  for (auto& t : g.triangles)
    membrane_e += membrane<double>(
        {A3<double>{x[t.v[0]].x, x[t.v[0]].y, x[t.v[0]].z},
         A3<double>{x[t.v[1]].x, x[t.v[1]].y, x[t.v[1]].z},
         A3<double>{x[t.v[2]].x, x[t.v[2]].y, x[t.v[2]].z}},
        t, g.materials[t.material]);
  for (auto& h : g.hinges) {
    std::array<A3<double>, 4> p;
    for (int k = 0; k < 4; ++k) p[k] = {x[h.v[k]].x, x[h.v[k]].y, x[h.v[k]].z};
    bending_e += bending<double>(p, h, g.materials[h.material]);
  }
  double contact = 0.0;
  if (sphere) {
    auto area = rest_area_of(g);
    double th = g.materials.empty() ? 0.001 : g.materials[0].thickness;
    for (size_t i = 0; i < x.size(); ++i)
      contact += sphere_energy<double>(*sphere, {x[i].x, x[i].y, x[i].z}, th, area[i]);
  }
  double seam = 0.0;
  for (size_t si = 0; si < g.stitch_points.size(); ++si)
    seam += 0.5 * g.stitch_stiffness[si] * length2(g.stitch_points[si].gap(x));
  double ogc = 0.0;
  if (frozen) {
    auto all = frozen_positions(x, frozen);
    for (auto& c : frozen->contacts) ogc += c.value(all, frozen->radius, frozen->stiffness);
  }
  return inertial(g, old, x, s) + membrane_e + bending_e + contact + seam + ogc;
}

NewtonMetrics newton_step(const Garment& g, State& state, const SolverSettings& s,
                          const Sphere* sphere, const FrozenContacts* frozen) {  // :422-647
  if (!std::isfinite(s.dt) || !(s.dt >= 1e-6 && s.dt <= 0.1) || !finite(s.gravity) ||
      !std::isfinite(s.damping) || s.damping < 0.0 || s.newton_iterations == 0 ||
      s.newton_iterations > 200 || s.pcg_iterations == 0 || s.pcg_iterations > 5000 ||
      !std::isfinite(s.absolute_tolerance) || s.absolute_tolerance <= 0.0 ||
      !std::isfinite(s.relative_tolerance) ||
      !(s.relative_tolerance >= 0.0 && s.relative_tolerance < 1.0))
    throw std::runtime_error("S|Invalid solver settings");
  int n = g.n;
  bool bad = (int)state.x.size() != n || (int)state.v.size() != n || !std::isfinite(state.time);
  for (int i = 0; !bad && i < n; ++i) bad = !finite(state.v[i]);
  if (bad || !geometry_valid(g, state.x, sphere))
    throw std::runtime_error("V|Invalid state, collapsed triangle or vertex at sphere centre");
  std::vector<V3> x = state.x;
  for (int i = 0; i < n; ++i)
    if (g.pins[i]) x[i] = *g.pins[i];
  auto area = rest_area_of(g);
  double th = g.materials.empty() ? 0.001 : g.materials[0].thickness;
  NewtonMetrics stats;
  double tolerance = 0.0;
  auto assemble = [&](const std::vector<V3>& xx) {
    System sys;
    for (auto& t : g.triangles) {
      std::array<A3<Jet<9>>, 3> p;
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) p[a][k] = Jet<9>::variable(xx[t.v[a]][k], 3 * a + k);
      sys.elements.push_back(from_jet<9>(t.v.data(), 3, membrane<Jet<9>>(p, t, g.materials[t.material])));
    }
    for (auto& h : g.hinges) {
      std::array<A3<Jet<12>>, 4> p;
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) p[a][k] = Jet<12>::variable(xx[h.v[a]][k], 3 * a + k);
      sys.elements.push_back(from_jet<12>(h.v.data(), 4, bending<Jet<12>>(p, h, g.materials[h.material])));
    }
    if (sphere) {
      for (int i = 0; i < n; ++i) {
        if (sphere_penetration(*sphere, xx[i], th) == 0.0) continue;
        A3<Jet<3>> p = {Jet<3>::variable(xx[i].x, 0), Jet<3>::variable(xx[i].y, 1),
                        Jet<3>::variable(xx[i].z, 2)};
        sys.elements.push_back(from_jet<3>(&i, 1, sphere_energy<Jet<3>>(*sphere, p, th, area[i])));
      }
    }
    // Seam elements (src/garment/seam.rs:240-265).
    for (size_t si = 0; si < g.stitch_points.size(); ++si) {
      auto& sp = g.stitch_points[si];
      double coef = g.stitch_stiffness[si];
      V3 gap = sp.gap(xx);
      Element e;
      size_t m = sp.terms.size(), nd = 3 * m;
      e.gradient.assign(nd, 0.0);
      e.hessian.assign(nd * nd, 0.0);
      for (size_t i = 0; i < m; ++i) {
        double a = sp.terms[i].second;
        for (int k = 0; k < 3; ++k) e.gradient[3 * i + k] = coef * a * gap[k];
        for (size_t j = 0; j < m; ++j) {
          double b = sp.terms[j].second;
          for (int k = 0; k < 3; ++k) e.hessian[(3 * i + k) * nd + 3 * j + k] = coef * a * b;
        }
      }
      for (auto& t : sp.terms)
        for (int k = 0; k < 3; ++k) e.dofs.push_back(3 * t.first + k);
      e.energy = 0.5 * coef * length2(gap);
      sys.elements.push_back(std::move(e));
    }
    // Frozen OGC contacts (extension): one element per contact over its moving vertices.
    if (frozen) {
      auto all = frozen_positions(xx, frozen);
      for (auto& c : frozen->contacts) {
        Jet<12> j;
        std::array<A3<Jet<12>>, 4> p;
        for (int a = 0; a < 4; ++a)
          for (int k = 0; k < 3; ++k) {
            int vid = c.vertices[a];
            bool mov = a < c.count && vid < n;
            p[a][k] = mov ? Jet<12>::variable(all[vid][k], 3 * a + k) : Jet<12>::c(all[vid][k]);
          }
        j = c.energy<Jet<12>>(p, frozen->radius, frozen->stiffness);
        Element e;
        std::vector<int> keep;
        for (int a = 0; a < c.count; ++a)
          if (c.vertices[a] < n) keep.push_back(a);
        for (int a : keep)
          for (int k = 0; k < 3; ++k) e.dofs.push_back(3 * c.vertices[a] + k);
        e.energy = j.v;
        for (int a : keep)
          for (int k = 0; k < 3; ++k) e.gradient.push_back(j.g[3 * a + k]);
        for (int a : keep)
          for (int k = 0; k < 3; ++k)
            for (int b : keep)
              for (int l = 0; l < 3; ++l) e.hessian.push_back(j.h[3 * a + k][3 * b + l]);
        sys.elements.push_back(std::move(e));
      }
    }
    size_t nd = (size_t)n * 3;
    sys.gradient.assign(nd, 0.0);
    sys.mass_diagonal.assign(nd, 0.0);
    sys.fixed.assign(nd, false);
    for (int i = 0; i < n; ++i) {
      V3 delta = g.masses[i] * ((xx[i] - state.x[i] - s.dt * state.v[i]) / (s.dt * s.dt) +
                                s.damping / s.dt * (xx[i] - state.x[i]) - s.gravity);
      for (int k = 0; k < 3; ++k) {
        sys.gradient[3 * i + k] = delta[k];
        sys.mass_diagonal[3 * i + k] = g.masses[i] * (1.0 / (s.dt * s.dt) + s.damping / s.dt);
        sys.fixed[3 * i + k] = (bool)g.pins[i];
      }
    }
    for (auto& e : sys.elements)
      for (size_t i = 0; i < e.dofs.size(); ++i) sys.gradient[e.dofs[i]] += e.gradient[i];
    for (size_t i = 0; i < nd; ++i)
      if (sys.fixed[i]) sys.gradient[i] = 0.0;
    double esum = 0.0;
    for (auto& e : sys.elements) esum += e.energy;
    sys.energy = inertial(g, state, xx, s) + esum;
    return sys;
  };
  for (int iteration = 0; iteration <= s.newton_iterations; ++iteration) {
    System sys = assemble(x);
    double residual = rms(sys.gradient, sys.fixed);
    if (!std::isfinite(residual) || !std::isfinite(sys.energy))
      throw std::runtime_error("N|Non-finite FEM assembly; step rolled back");
    if (iteration == 0) {
      stats.objective_initial_j = sys.energy;
      tolerance = fmax_(s.absolute_tolerance, s.relative_tolerance * residual);
    }
    stats.residual_rms_n = residual;
    stats.tolerance_n = tolerance;
    stats.objective_final_j = sys.energy;
    stats.newton_iterations = iteration;
    if (residual <= tolerance) {
      stats.converged = true;
      break;
    }
    if (iteration == s.newton_iterations) break;
    double scale = 1.0;
    for (double d : sys.diagonal(0.0)) scale = fmax_(scale, d);
    std::optional<std::vector<V3>> accepted;
    double shift = 0.0;
    for (int rep = 0; rep < 12; ++rep) {
      auto [solve, its] = sys.pcg(shift, s.pcg_iterations);
      stats.pcg_iterations += its;
      if (solve) {
        double slope = System::vdot(sys.gradient, *solve);
        if (std::isfinite(slope) && slope < 0.0) {
          double alpha = 1.0;
          for (int k = 0; k < 24; ++k) {
            std::vector<V3> trial(n);
            for (int i = 0; i < n; ++i)
              trial[i] = g.pins[i] ? *g.pins[i]
                                   : x[i] + alpha * V3{(*solve)[3 * i], (*solve)[3 * i + 1],
                                                       (*solve)[3 * i + 2]};
            double obj = newton_objective(g, state, trial, s, sphere, frozen);
            if (std::isfinite(obj) && obj <= sys.energy + 1e-4 * alpha * slope) {
              accepted = trial;
              break;
            }
            alpha *= 0.5;
            stats.line_search_backtracks++;
          }
        }
      }
      if (accepted) break;
      shift = shift == 0.0 ? scale * 1e-6 : shift * 10.0;
    }
    stats.max_regularization = fmax_(stats.max_regularization, shift);
    if (!accepted)
      throw std::runtime_error("L|Newton line search failed; state unchanged. Reduce time step.");
    x = *accepted;
  }
  for (int i = 0; i < n; ++i) state.v[i] = g.pins[i] ? V3{0, 0, 0} : (x[i] - state.x[i]) / s.dt;
  state.x = x;
  state.time += s.dt;
  return stats;
}

}  // namespace cvo
