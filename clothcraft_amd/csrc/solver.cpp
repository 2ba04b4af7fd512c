// solver.cpp — cv_solver construction (VbdSolver::new, src/vbd.rs:101-192),
// step-argument validation and greedy colouring.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>

#include "ogc_math.hpp"

[[noreturn]] void fail(cv_status st, const std::string& msg) { throw CvError{st, msg}; }

namespace {

using ogm::D3;

D3 pt(const std::vector<double>& x, uint32_t i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

void validate_material(const cv_material& m) {  // src/model.rs:40-63
  const double pos[] = {m.warp, m.weft, m.shear, m.density, m.thickness, m.compression_transition};
  const char* names[] = {"warp", "weft", "shear", "density", "thickness", "compression_transition"};
  for (int i = 0; i < 6; ++i)
    if (!std::isfinite(pos[i]) || pos[i] <= 0.0)
      fail(CV_ERR_INVALID_MATERIAL, std::string(names[i]) + " must be finite and positive");
  if (!std::isfinite(m.grain_degrees) || !std::isfinite(m.compression_ratio) ||
      !(m.compression_ratio >= 0.0 && m.compression_ratio <= 1.0) || !std::isfinite(m.bend_warp) ||
      m.bend_warp < 0.0 || !std::isfinite(m.bend_weft) || m.bend_weft < 0.0)
    fail(CV_ERR_INVALID_MATERIAL, "Invalid bending, compression or grain parameter");
}

// Topology::new (src/ogc/topology.rs:13-52).
void build_topology(OgcTopology& t, uint32_t n_total, const std::vector<uint32_t>& faces) {
  t.n_total = n_total;
  t.faces = faces;
  size_t F = faces.size() / 3;
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> map;
  std::vector<std::set<uint32_t>> nb(n_total);
  std::vector<std::vector<uint32_t>> opp;
  t.tri_edges.resize(3 * F);
  for (size_t f = 0; f < F; ++f) {
    const uint32_t* tr = &faces[3 * f];
    for (int k = 0; k < 3; ++k)
      if (tr[k] >= n_total) fail(CV_ERR_TOPOLOGY, "OGC: invalid triangle indices");
    if (tr[0] == tr[1] || tr[1] == tr[2] || tr[0] == tr[2])
      fail(CV_ERR_TOPOLOGY, "OGC: invalid triangle indices");
    for (int k = 0; k < 3; ++k) {
      uint32_t a = tr[k], b = tr[(k + 1) % 3], c = tr[(k + 2) % 3];
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = map.find(key);
      uint32_t e;
      if (it == map.end()) {
        e = (uint32_t)(t.edges.size() / 2);
        t.edges.push_back(key.first);
        t.edges.push_back(key.second);
        opp.push_back({});
        map.emplace(key, e);
      } else {
        e = it->second;
      }
      opp[e].push_back(c);
      nb[a].insert(b);
      nb[b].insert(a);
      t.tri_edges[3 * f + k] = e;
    }
  }
  t.opposites.assign(opp.size() * 2, -1);
  for (size_t e = 0; e < opp.size(); ++e) {
    if (opp[e].size() > 2)
      fail(CV_ERR_TOPOLOGY, "OGC requires manifold triangle edges (at most two incident faces)");
    for (size_t k = 0; k < opp[e].size(); ++k) t.opposites[2 * e + k] = (int32_t)opp[e][k];
  }
  t.nb_off.assign(n_total + 1, 0);
  for (uint32_t i = 0; i < n_total; ++i) t.nb_off[i + 1] = t.nb_off[i] + (uint32_t)nb[i].size();
  t.nb.reserve(t.nb_off[n_total]);
  for (uint32_t i = 0; i < n_total; ++i) t.nb.insert(t.nb.end(), nb[i].begin(), nb[i].end());
}

// A uniform grid over boxes for the one-time start-state checks.
struct HostGrid {
  double h = 1;
  D3 lo{0, 0, 0};
  std::vector<std::pair<uint64_t, uint32_t>> cells;  // (cell key, primitive), sorted
  static uint64_t key(int64_t i, int64_t j, int64_t k) {
    return ((uint64_t)(i & 0x1FFFFF) << 42) | ((uint64_t)(j & 0x1FFFFF) << 21) | (uint64_t)(k & 0x1FFFFF);
  }
  void range(const ogm::Box& b, int64_t r[6]) const {
    r[0] = (int64_t)std::floor((b.lo.x - lo.x) / h);
    r[1] = (int64_t)std::floor((b.lo.y - lo.y) / h);
    r[2] = (int64_t)std::floor((b.lo.z - lo.z) / h);
    r[3] = (int64_t)std::floor((b.hi.x - lo.x) / h);
    r[4] = (int64_t)std::floor((b.hi.y - lo.y) / h);
    r[5] = (int64_t)std::floor((b.hi.z - lo.z) / h);
  }
  void build(const std::vector<ogm::Box>& boxes) {
    cells.clear();
    for (uint32_t p = 0; p < boxes.size(); ++p) {
      int64_t r[6];
      range(boxes[p], r);
      for (int64_t i = r[0]; i <= r[3]; ++i)
        for (int64_t j = r[1]; j <= r[4]; ++j)
          for (int64_t k = r[2]; k <= r[5]; ++k) cells.push_back({key(i, j, k), p});
    }
    std::sort(cells.begin(), cells.end());
  }
  template <class F>
  void visit(const ogm::Box& b, F&& f) const {
    int64_t r[6];
    range(b, r);
    for (int64_t i = r[0]; i <= r[3]; ++i)
      for (int64_t j = r[1]; j <= r[4]; ++j)
        for (int64_t k = r[2]; k <= r[5]; ++k) {
          uint64_t kk = key(i, j, k);
          auto it = std::lower_bound(cells.begin(), cells.end(), std::make_pair(kk, (uint32_t)0));
          for (; it != cells.end() && it->first == kk; ++it) f(it->second);
        }
  }
};

ogm::Box inflate(ogm::Box b, double r) {
  return {{b.lo.x - r, b.lo.y - r, b.lo.z - r}, {b.hi.x + r, b.hi.y + r, b.hi.z + r}};
}
bool overlaps(const ogm::Box& a, const ogm::Box& b) {
  return a.lo.x <= b.hi.x && a.lo.y <= b.hi.y && a.lo.z <= b.hi.z && a.hi.x >= b.lo.x &&
         a.hi.y >= b.lo.y && a.hi.z >= b.lo.z;
}

bool excluded_pair(const OgcTopology& t, uint32_t a, uint32_t b) {
  return std::binary_search(t.ex.begin() + t.ex_off[a], t.ex.begin() + t.ex_off[a + 1], b);
}
bool adjacent(const OgcTopology& t, uint32_t a, uint32_t b) { return a == b || excluded_pair(t, a, b); }

// OgcWorld::validate_start (src/ogc/world.rs:242-307) on the initial positions.
void validate_start(const OgcTopology& t, const std::vector<double>& x) {
  size_t F = t.faces.size() / 3, E = t.edges.size() / 2;
  uint32_t mv = t.moving;
  std::vector<ogm::Box> fb(F);
  double ext = 0;
  for (size_t f = 0; f < F; ++f) {
    fb[f] = ogm::box_of(pt(x, t.faces[3 * f]), pt(x, t.faces[3 * f + 1]), pt(x, t.faces[3 * f + 2]));
    ext = std::max({ext, fb[f].hi.x - fb[f].lo.x, fb[f].hi.y - fb[f].lo.y, fb[f].hi.z - fb[f].lo.z});
  }
  auto fixed_face = [&](size_t f) {
    return t.faces[3 * f] >= mv && t.faces[3 * f + 1] >= mv && t.faces[3 * f + 2] >= mv;
  };
  HostGrid grid;
  D3 lo{INFINITY, INFINITY, INFINITY};
  for (auto& b : fb) lo = {std::min(lo.x, b.lo.x), std::min(lo.y, b.lo.y), std::min(lo.z, b.lo.z)};
  grid.lo = {lo.x - 1.0, lo.y - 1.0, lo.z - 1.0};
  // Cell size: the median face extent keeps cell occupancy small for mixed sizes.
  std::vector<double> exts(F);
  for (size_t f = 0; f < F; ++f)
    exts[f] = std::max({fb[f].hi.x - fb[f].lo.x, fb[f].hi.y - fb[f].lo.y, fb[f].hi.z - fb[f].lo.z});
  std::nth_element(exts.begin(), exts.begin() + F / 2, exts.end());
  grid.h = std::max(exts[F / 2], 1e-6) * 2.0;
  grid.build(fb);
  std::vector<int64_t> stamp(F, -1);
  for (size_t i = 0; i < F; ++i) {
    const uint32_t* tr = &t.faces[3 * i];
    bool tf = fixed_face(i);
    ogm::Box q = inflate(fb[i], 1e-10);
    int64_t hit = -1;
    grid.visit(q, [&](uint32_t j) {
      if (stamp[j] == (int64_t)i) return;
      stamp[j] = (int64_t)i;
      if (j <= i || !overlaps(fb[j], q)) return;
      const uint32_t* qq = &t.faces[3 * j];
      if (tf && fixed_face(j)) return;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          if (excluded_pair(t, std::min(tr[a], qq[b]), std::max(tr[a], qq[b]))) return;
      auto contains = [](const uint32_t* s, uint32_t v) { return s[0] == v || s[1] == v || s[2] == v; };
      for (int k = 0; k < 3; ++k) {
        bool h1 = !contains(qq, tr[k]) && !contains(qq, tr[(k + 1) % 3]) &&
                  ogm::segment_triangle(pt(x, tr[k]), pt(x, tr[(k + 1) % 3]), pt(x, qq[0]),
                                        pt(x, qq[1]), pt(x, qq[2]));
        bool h2 = !contains(tr, qq[k]) && !contains(tr, qq[(k + 1) % 3]) &&
                  ogm::segment_triangle(pt(x, qq[k]), pt(x, qq[(k + 1) % 3]), pt(x, tr[0]),
                                        pt(x, tr[1]), pt(x, tr[2]));
        if (h1 || h2) hit = j;
      }
    });
    if (hit >= 0)
      fail(CV_ERR_INITIAL_INTERSECTION, "OGC initial intersection between faces " + std::to_string(i) +
                                            " and " + std::to_string(hit) +
                                            "; reset/rearrange before simulating");
  }
  // Minimum VF/EE separation at query radius 1e-5 (the detect call of :293-305).
  const double rq = 1e-5;
  double scale = rq;
  for (double v : x) scale = std::max(scale, std::fabs(v));
  double pr = rq + 64.0 * 2.220446049250313e-16 * scale;
  double pr2 = pr * pr;
  double md = rq;
  std::vector<ogm::Box> ib(F);
  for (size_t f = 0; f < F; ++f) ib[f] = inflate(fb[f], pr);
  grid.build(ib);
  for (uint32_t v = 0; v < t.n_total; ++v) {
    D3 p = pt(x, v);
    ogm::Box pb = ogm::box_of(p);
    grid.visit(pb, [&](uint32_t f) {
      const uint32_t* tr = &t.faces[3 * f];
      if (v >= mv && fixed_face(f)) return;
      if (adjacent(t, v, tr[0]) || adjacent(t, v, tr[1]) || adjacent(t, v, tr[2])) return;
      if (ogm::box_dist2(pb, fb[f]) > pr2) return;
      D3 q;
      int feat, loc;
      ogm::closest_triangle(p, pt(x, tr[0]), pt(x, tr[1]), pt(x, tr[2]), q, feat, loc);
      double d = ogm::dist(p, q);
      if (d < rq) md = std::min(md, d);
    });
  }
  std::vector<ogm::Box> eb(E), eib(E);
  for (size_t e = 0; e < E; ++e) {
    eb[e] = ogm::box_of(pt(x, t.edges[2 * e]), pt(x, t.edges[2 * e + 1]));
    eib[e] = inflate(eb[e], pr);
  }
  grid.build(eib);
  std::vector<int64_t> estamp(E, -1);
  for (uint32_t e = 0; e < E; ++e) {
    uint32_t a = t.edges[2 * e], b = t.edges[2 * e + 1];
    grid.visit(eb[e], [&](uint32_t f) {
      if (f <= e || estamp[f] == (int64_t)e) return;
      estamp[f] = (int64_t)e;
      uint32_t jc = t.edges[2 * f], jd = t.edges[2 * f + 1];
      if (!t.edge_moving[e] && !t.edge_moving[f]) return;
      if (adjacent(t, a, jc) || adjacent(t, a, jd) || adjacent(t, b, jc) || adjacent(t, b, jd)) return;
      if (ogm::box_dist2(eb[e], eb[f]) > pr2) return;
      uint32_t e0 = e, f0 = f;
      if (!t.edge_moving[e0]) std::swap(e0, f0);  // moving edge first
      D3 A = pt(x, t.edges[2 * e0]), B = pt(x, t.edges[2 * e0 + 1]);
      D3 C = pt(x, t.edges[2 * f0]), Dd = pt(x, t.edges[2 * f0 + 1]);
      double s, tt;
      ogm::closest_segments(A, B, C, Dd, s, tt);
      D3 P = ogm::add(A, ogm::scale(ogm::sub(B, A), s));
      D3 Q = ogm::add(C, ogm::scale(ogm::sub(Dd, C), tt));
      double d = ogm::dist(P, Q);
      if (d < rq) md = std::min(md, d);
    });
  }
  if (md <= 1e-10)
    fail(CV_ERR_INITIAL_TOUCHING, "OGC initial touching/coplanar features: separation must be positive");
}

// Smallest-last vertex order of a graph (CSR, symmetric): repeatedly remove
// a vertex of least remaining degree (ties: smallest index), then reverse the
// removal sequence.  Greedy colouring in this order needs at most one colour
// more than the graph's degeneracy, and fewer colours than index order on
// irregular meshes.
std::vector<uint32_t> smallest_last_order(uint32_t n, const std::vector<uint32_t>& off,
                                          const std::vector<uint32_t>& adj) {
  std::vector<uint32_t> deg(n);
  std::set<std::pair<uint32_t, uint32_t>> queue;
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = off[v + 1] - off[v];
    queue.insert({deg[v], v});
  }
  std::vector<char> removed(n, 0);
  std::vector<uint32_t> seq;
  seq.reserve(n);
  while (!queue.empty()) {
    uint32_t v = queue.begin()->second;
    queue.erase(queue.begin());
    removed[v] = 1;
    seq.push_back(v);
    for (uint32_t k = off[v]; k < off[v + 1]; ++k) {
      uint32_t u = adj[k];
      if (u >= n || removed[u]) continue;
      queue.erase({deg[u], u});
      queue.insert({--deg[u], u});
    }
  }
  std::reverse(seq.begin(), seq.end());
  return seq;
}

}  // namespace

// VbdSolver::new (src/vbd.rs:101-192).
void build_solver(cv_solver& s, const cv_garment* g, const cv_surface* avatar, const cv_state* st) {
  if (!g || !st) fail(CV_ERR_INVALID_ARGUMENT, "null garment or state");
  uint32_t n = g->n_vertices;
  if (st->n != n || !st->x || !st->v)
    fail(CV_ERR_DIMENSIONS, "VBD state dimensions do not match garment");
  s.n = n;
  s.masses.assign(g->masses, g->masses + n);
  s.pinned.assign(n, 0);
  s.pin_targets.assign(3 * (size_t)n, 0.0);
  for (uint32_t i = 0; i < n; ++i)
    if (g->pinned && g->pinned[i]) {
      s.pinned[i] = 1;
      for (int k = 0; k < 3; ++k) s.pin_targets[3 * i + k] = g->pin_targets[3 * i + k];
    }
  s.n_free = 0;
  for (uint32_t i = 0; i < n; ++i) s.n_free += s.pinned[i] ? 0u : 1u;
  s.materials.assign(g->materials, g->materials + g->n_materials);
  for (auto& m : s.materials) validate_material(m);
  s.tri_v.assign(g->tri_vertices, g->tri_vertices + 3 * (size_t)g->n_triangles);
  s.tri_inv_dm.assign(g->tri_inv_dm, g->tri_inv_dm + 4 * (size_t)g->n_triangles);
  s.tri_area.assign(g->tri_area, g->tri_area + g->n_triangles);
  s.tri_mat.assign(g->tri_material, g->tri_material + g->n_triangles);
  s.hinge_v.assign(g->hinge_vertices, g->hinge_vertices + 4 * (size_t)g->n_hinges);
  s.hinge_weight.assign(g->hinge_weight, g->hinge_weight + g->n_hinges);
  s.hinge_across.assign(g->hinge_across, g->hinge_across + 2 * (size_t)g->n_hinges);
  s.hinge_mat.assign(g->hinge_material, g->hinge_material + g->n_hinges);
  for (uint32_t v : s.tri_v)
    if (v >= n) fail(CV_ERR_TOPOLOGY, "OGC: invalid triangle indices");
  for (uint32_t m : s.tri_mat)
    if (m >= s.materials.size()) fail(CV_ERR_INVALID_MATERIAL, "Invalid material binding or mesh topology");
  for (uint32_t v : s.hinge_v)
    if (v >= n) fail(CV_ERR_TOPOLOGY, "Invalid material binding or mesh topology");
  for (uint32_t m : s.hinge_mat)
    if (m >= s.materials.size()) fail(CV_ERR_INVALID_MATERIAL, "Invalid material binding or mesh topology");
  s.st_off.assign(1, 0);
  if (g->n_stitches) {
    s.st_off.assign(g->stitch_offsets, g->stitch_offsets + g->n_stitches + 1);
    s.st_v.assign(g->stitch_vertices, g->stitch_vertices + s.st_off.back());
    s.st_w.assign(g->stitch_weights, g->stitch_weights + s.st_off.back());
    s.st_k.assign(g->stitch_stiffness, g->stitch_stiffness + g->n_stitches);
  }
  for (uint32_t i = 0; i + 1 < s.st_off.size(); ++i) {
    if (s.st_off[i + 1] - s.st_off[i] > 4)
      fail(CV_ERR_CAPACITY, "stitch samples couple at most 4 vertices");
    for (uint32_t k = s.st_off[i]; k < s.st_off[i + 1]; ++k)
      if (s.st_v[k] >= n || !std::isfinite(s.st_w[k])) fail(CV_ERR_INVALID_SEAM, "Invalid compiled seam");
  }
  s.pr_off.assign(1, 0);
  if (g->n_probes) {
    s.pr_off.assign(g->probe_offsets, g->probe_offsets + g->n_probes + 1);
    s.pr_v.assign(g->probe_vertices, g->probe_vertices + s.pr_off.back());
    s.pr_w.assign(g->probe_weights, g->probe_weights + s.pr_off.back());
  }
  // Pins must match the state (src/vbd.rs:110-120).
  for (uint32_t i = 0; i < n; ++i)
    if (s.pinned[i]) {
      double dx = st->x[3 * i] - s.pin_targets[3 * i], dy = st->x[3 * i + 1] - s.pin_targets[3 * i + 1],
             dz = st->x[3 * i + 2] - s.pin_targets[3 * i + 2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) > 1e-10)
        fail(CV_ERR_PIN_MISMATCH, "VBD pinned positions must match their constraints before initialization");
    }
  std::vector<double> x(st->x, st->x + 3 * (size_t)n);
  std::vector<uint32_t> faces(s.tri_v);
  uint32_t n_fixed = 0;
  if (avatar) {  // SurfaceMesh::validate (src/avatar.rs:78-97)
    if (avatar->n_vertices == 0 || avatar->n_triangles == 0) fail(CV_ERR_INVALID_AVATAR, "Empty/non-finite avatar");
    for (size_t i = 0; i < 3 * (size_t)avatar->n_vertices; ++i)
      if (!std::isfinite(avatar->vertices[i])) fail(CV_ERR_INVALID_AVATAR, "Empty/non-finite avatar");
    for (size_t f = 0; f < avatar->n_triangles; ++f) {
      const uint32_t* t = &avatar->triangles[3 * f];
      for (int k = 0; k < 3; ++k)
        if (t[k] >= avatar->n_vertices) fail(CV_ERR_INVALID_AVATAR, "Invalid avatar index");
      D3 a{avatar->vertices[3 * t[0]], avatar->vertices[3 * t[0] + 1], avatar->vertices[3 * t[0] + 2]};
      D3 b{avatar->vertices[3 * t[1]], avatar->vertices[3 * t[1] + 1], avatar->vertices[3 * t[1] + 2]};
      D3 c{avatar->vertices[3 * t[2]], avatar->vertices[3 * t[2] + 1], avatar->vertices[3 * t[2] + 2]};
      if (ogm::len2(ogm::cross(ogm::sub(b, a), ogm::sub(c, a))) < 1e-24)
        fail(CV_ERR_INVALID_AVATAR, "Degenerate avatar triangle");
    }
    n_fixed = avatar->n_vertices;
    s.avatar.assign(avatar->vertices, avatar->vertices + 3 * (size_t)n_fixed);
    x.insert(x.end(), s.avatar.begin(), s.avatar.end());
    for (size_t i = 0; i < 3 * (size_t)avatar->n_triangles; ++i) faces.push_back(avatar->triangles[i] + n);
  }
  for (double v : x)
    if (!std::isfinite(v)) fail(CV_ERR_INVALID_POSITIONS, "Invalid OGC positions");
  // Seam-excluded vertex pairs: opposite-signed stitch weights (src/vbd.rs:128-138).
  std::set<std::pair<uint32_t, uint32_t>> excluded;
  for (uint32_t i = 0; i + 1 < s.st_off.size(); ++i)
    for (uint32_t a = s.st_off[i]; a < s.st_off[i + 1]; ++a)
      for (uint32_t b = s.st_off[i]; b < s.st_off[i + 1]; ++b)
        if (s.st_w[a] * s.st_w[b] < 0.0)
          excluded.insert({std::min(s.st_v[a], s.st_v[b]), std::max(s.st_v[a], s.st_v[b])});
  OgcTopology& t = s.top;
  t.moving = n;
  build_topology(t, n + n_fixed, faces);
  size_t F = faces.size() / 3, E = t.edges.size() / 2;
  for (size_t f = 0; f < F; ++f) {
    D3 a = pt(x, faces[3 * f]), b = pt(x, faces[3 * f + 1]), c = pt(x, faces[3 * f + 2]);
    if (ogm::len2(ogm::cross(ogm::sub(b, a), ogm::sub(c, a))) < 1e-24)
      fail(CV_ERR_DEGENERATE_FACE, "OGC degenerate initial face");
  }
  t.face_moving.resize(F);
  t.moving_faces = 0;
  for (size_t f = 0; f < F; ++f) {
    t.face_moving[f] = faces[3 * f] < n || faces[3 * f + 1] < n || faces[3 * f + 2] < n;
    t.moving_faces += t.face_moving[f];
  }
  t.edge_moving.resize(E);
  t.rest_len2.resize(E);
  for (size_t e = 0; e < E; ++e) {
    t.edge_moving[e] = t.edges[2 * e] < n || t.edges[2 * e + 1] < n;
    D3 d = ogm::sub(pt(x, t.edges[2 * e]), pt(x, t.edges[2 * e + 1]));
    t.rest_len2[e] = ogm::len2(d);
  }
  std::vector<std::vector<uint32_t>> exl(n + n_fixed);
  for (auto& [a, b] : excluded) {
    exl[a].push_back(b);
    exl[b].push_back(a);
  }
  t.ex_off.assign(n + n_fixed + 1, 0);
  for (uint32_t i = 0; i < n + n_fixed; ++i) {
    std::sort(exl[i].begin(), exl[i].end());
    t.ex_off[i + 1] = t.ex_off[i] + (uint32_t)exl[i].size();
    t.ex.insert(t.ex.end(), exl[i].begin(), exl[i].end());
  }
  validate_start(t, x);
  {
    std::vector<double> ext;
    for (size_t f = 0; f < F; ++f) {
      if (!t.face_moving[f]) continue;
      ogm::Box b = ogm::box_of(pt(x, faces[3 * f]), pt(x, faces[3 * f + 1]), pt(x, faces[3 * f + 2]));
      ext.push_back(std::max({b.hi.x - b.lo.x, b.hi.y - b.lo.y, b.hi.z - b.lo.z}));
    }
    if (!ext.empty()) {
      std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
      s.face_extent_median = ext[ext.size() / 2];
    }
  }

  // Conflict graph (src/vbd.rs:147-178) and per-vertex term records.
  std::vector<std::set<uint32_t>> graph(n);
  auto connect = [&](const uint32_t* ids, int k) {
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b)
        if (ids[a] != ids[b]) graph[ids[a]].insert(ids[b]);
  };
  std::vector<std::vector<uint32_t>> tri_of(n), hinge_of(n), st_of(n);
  size_t T = s.tri_area.size(), H = s.hinge_weight.size();
  for (size_t i = 0; i < T; ++i) {
    for (int k = 0; k < 3; ++k) tri_of[s.tri_v[3 * i + k]].push_back((uint32_t)i);
    connect(&s.tri_v[3 * i], 3);
  }
  for (size_t i = 0; i < H; ++i) {
    for (int k = 0; k < 4; ++k) hinge_of[s.hinge_v[4 * i + k]].push_back((uint32_t)i);
    connect(&s.hinge_v[4 * i], 4);
  }
  for (uint32_t i = 0; i + 1 < s.st_off.size(); ++i) {
    for (uint32_t k = s.st_off[i]; k < s.st_off[i + 1]; ++k) st_of[s.st_v[k]].push_back(i);
    connect(&s.st_v[s.st_off[i]], (int)(s.st_off[i + 1] - s.st_off[i]));
  }
  s.graph_off.assign(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) s.graph_off[i + 1] = s.graph_off[i] + (uint32_t)graph[i].size();
  s.graph.clear();
  for (uint32_t i = 0; i < n; ++i) s.graph.insert(s.graph.end(), graph[i].begin(), graph[i].end());
  s.sl_order = smallest_last_order(n, s.graph_off, s.graph);

  s.term_off.assign(n + 1, 0);
  s.terms64.clear();
  for (uint32_t v = 0; v < n; ++v) {
    for (uint32_t ti : tri_of[v]) {
      const uint32_t* tv = &s.tri_v[3 * ti];
      const cv_material& m = s.materials[s.tri_mat[ti]];
      double sn = std::sin(m.grain_degrees * (M_PI / 180.0)), cs = std::cos(m.grain_degrees * (M_PI / 180.0));
      const double* inv = &s.tri_inv_dm[4 * ti];
      // u = x_axis cos + y_axis sin, v = -x_axis sin + y_axis cos (src/energy.rs:21-23)
      double ux = inv[0] * cs + inv[2] * sn, uy = inv[1] * cs + inv[3] * sn;
      double vx = -inv[0] * sn + inv[2] * cs, vy = -inv[1] * sn + inv[3] * cs;
      double cw[3] = {-ux - uy, ux, uy}, cu[3] = {-vx - vy, vx, vy};
      int slot = tv[0] == v ? 0 : (tv[1] == v ? 1 : 2);
      int o1 = (slot + 1) % 3, o2 = (slot + 2) % 3;
      cvm::TermRec64 r{};
      r.head = cvm::rec_head(cvm::TERM_MEMBRANE, slot, 3, 0);
      r.ids[0] = tv[o1];
      r.ids[1] = tv[o2];
      r.ids[2] = r.ids[3] = 0;
      r.f[0] = cw[o1];
      r.f[1] = cw[o2];
      r.f[2] = cu[o1];
      r.f[3] = cu[o2];
      r.f[4] = s.tri_area[ti];
      r.f[5] = m.warp;
      r.f[6] = m.weft;
      r.f[7] = m.shear;
      r.f[8] = m.compression_ratio;
      r.f[9] = m.compression_transition;
      s.terms64.push_back(r);
    }
    for (uint32_t hi : hinge_of[v]) {
      const uint32_t* hv = &s.hinge_v[4 * hi];
      const cv_material& m = s.materials[s.hinge_mat[hi]];
      double sn = std::sin(m.grain_degrees * (M_PI / 180.0)), cs = std::cos(m.grain_degrees * (M_PI / 180.0));
      double ax = s.hinge_across[2 * hi], ay = s.hinge_across[2 * hi + 1];
      double aw = ax * cs + ay * sn, af = -ax * sn + ay * cs;
      int slot = hv[0] == v ? 0 : (hv[1] == v ? 1 : (hv[2] == v ? 2 : 3));
      cvm::TermRec64 r{};
      r.head = cvm::rec_head(cvm::TERM_HINGE, slot, 4, 0);
      for (int k = 0; k < 4; ++k) r.ids[k] = hv[k];
      r.f[0] = (s.hinge_weight[hi] * (m.bend_warp * aw * aw + m.bend_weft * af * af));
      s.terms64.push_back(r);
    }
    for (uint32_t si : st_of[v]) {
      cvm::TermRec64 r{};
      int cnt = (int)(s.st_off[si + 1] - s.st_off[si]);
      int slot = 0;
      for (int k = 0; k < cnt; ++k) {
        r.ids[k] = s.st_v[s.st_off[si] + k];
        r.f[k] = s.st_w[s.st_off[si] + k];
        if (r.ids[k] == v) slot = k;
      }
      r.f[4] = s.st_k[si];
      r.head = cvm::rec_head(cvm::TERM_STITCH, slot, cnt, 0);
      s.terms64.push_back(r);
    }
    s.term_off[v + 1] = (uint32_t)s.terms64.size();
  }
  s.terms.resize(s.terms64.size());
  for (size_t i = 0; i < s.terms64.size(); ++i) {
    s.terms[i].head = s.terms64[i].head;
    for (int k = 0; k < 4; ++k) s.terms[i].ids[k] = s.terms64[i].ids[k];
    for (int k = 0; k < 11; ++k) s.terms[i].f[k] = (float)s.terms64[i].f[k];
  }
  static uint64_t counter = 0;
  s.generation = ++counter;
}

void validate_settings(const cv_settings* s) {  // src/vbd.rs:37-51, src/ogc/world.rs:56-71
  if (!s) fail(CV_ERR_INVALID_ARGUMENT, "null settings");
  const cv_ogc_settings& c = s->contact;
  if (!std::isfinite(c.radius) || c.radius <= 0.0 || !std::isfinite(c.stiffness) || c.stiffness <= 0.0 ||
      !std::isfinite(c.relaxation) || c.relaxation <= 0.0 || c.relaxation >= 0.5 ||
      !std::isfinite(c.refresh_fraction) || !(c.refresh_fraction >= 0.0 && c.refresh_fraction <= 1.0))
    fail(CV_ERR_INVALID_OGC_SETTINGS, "Invalid OGC radius, stiffness or trust-region parameters");
  if (!std::isfinite(s->dt) || !(s->dt >= 1e-6 && s->dt <= 0.025) || s->iterations == 0 ||
      s->iterations > 200 || !std::isfinite(s->damping) || s->damping < 0.0 ||
      !std::isfinite(s->gravity[0]) || !std::isfinite(s->gravity[1]) || !std::isfinite(s->gravity[2]))
    fail(CV_ERR_INVALID_SETTINGS, "Invalid VBD settings");
  if (!std::isfinite(s->residual_tolerance) || s->residual_tolerance < 0.0)
    fail(CV_ERR_INVALID_SETTINGS, "Invalid VBD settings");
  if (s->flags &
      ~(uint32_t)(CV_FLAG_DIAGNOSTICS_OFF | CV_FLAG_TRUE_RESIDUAL | CV_FLAG_FAST_COLORING | CV_FLAG_WARM_START))
    fail(CV_ERR_INVALID_SETTINGS, "Invalid VBD settings");
}

void validate_state(const cv_solver& s, const cv_state* st) {  // src/vbd.rs:257-264
  if (!st || !st->x || !st->v || st->n != s.n || !std::isfinite(st->time))
    fail(CV_ERR_INVALID_STATE, "Invalid VBD state");
  for (size_t i = 0; i < 3 * (size_t)s.n; ++i)
    if (!std::isfinite(st->x[i]) || !std::isfinite(st->v[i])) fail(CV_ERR_INVALID_STATE, "Invalid VBD state");
}

void validate_grab(const cv_solver& s, const cv_grab* g) {  // src/vbd.rs:265-274
  if (!g) return;
  bool bad = !std::isfinite(g->target[0]) || !std::isfinite(g->target[1]) || !std::isfinite(g->target[2]) ||
             !std::isfinite(g->stiffness) || g->stiffness <= 0.0;
  double sum = 0.0;
  for (int k = 0; k < 3; ++k) {
    bad = bad || g->vertices[k] >= s.n || !std::isfinite(g->weights[k]) || g->weights[k] < 0.0;
    sum += g->weights[k];
  }
  if (bad || std::fabs(sum - 1.0) > 1e-8) fail(CV_ERR_INVALID_GRAB, "Invalid mouse attachment");
}

bool prediction_valid(const cv_solver& s, const cv_state* st, const cv_settings* set) {
  // src/vbd/prediction.rs:16-22: bitwise equality with the last committed state.
  if (!s.has_last || s.last_time != st->time) return false;
  for (int k = 0; k < 3; ++k)
    if (s.last_gravity[k] != set->gravity[k]) return false;
  if (s.last_damping != set->damping) return false;
  for (size_t i = 0; i < 3 * (size_t)s.n; ++i)
    if (s.last_x[i] != st->x[i] || s.last_v[i] != st->v[i]) return false;
  return true;
}

namespace {
// Rust's `{:?}` of an f64 prints the shortest decimal that reads back to the
// same value; %.17g prints one that reads back exactly.
std::string fmt_dvec3(double x, double y, double z) {
  char buf[160];
  std::snprintf(buf, sizeof buf, "DVec3(%.17g, %.17g, %.17g)", x, y, z);
  return buf;
}
std::string fmt_g(double v) {
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}
}  // namespace

double check_probe(const cv_solver& s, const double* x, const double* t, const uint32_t* coff,
                   const cvm::TermRec64* crec, const double* origin, const cv_settings* set, const cv_grab* grab,
                   const double* probe) {
  using V = cvm::V3<double>;
  auto P = [&](int, uint32_t i) { return V{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; };
  const double cr = set->contact.radius, ck = set->contact.stiffness;
  double max_error = 0.0;
  for (uint32_t v = 0; v < s.n; ++v) {
    const V xv = P(0, v);
    cvm::Deriv<double> acc = cvm::deriv_zero<double>();
    for (uint32_t k = s.term_off[v]; k < s.term_off[v + 1]; ++k)
      cvm::deriv_add(acc, cvm::term_deriv<double>(s.terms64[k], xv, cr, ck, P));
    for (uint32_t k = coff[v]; k < coff[v + 1]; ++k) cvm::deriv_add(acc, cvm::term_deriv<double>(crec[k], xv, cr, ck, P));
    if (grab) {  // src/vbd.rs:237-246
      int slot = -1;
      for (int q = 0; q < 3; ++q)
        if (grab->vertices[q] == v) slot = q;
      if (slot >= 0) {
        V d = xv - V{grab->target[0] - origin[0], grab->target[1] - origin[1], grab->target[2] - origin[2]};
        for (int q = 0; q < 3; ++q)
          if (q != slot) d = d + (P(0, grab->vertices[q]) - xv) * grab->weights[q];
        cvm::deriv_add(acc, cvm::stitch_d<double>(d, grab->weights[slot], grab->stiffness));
      }
    }
    const V tv{t[4 * v], t[4 * v + 1], t[4 * v + 2]};
    const double inertia = t[4 * v + 3];
    const V g = acc.g + (xv - tv) * inertia;
    const V actual{probe[4 * v], probe[4 * v + 1], probe[4 * v + 2]};
    const V diff = actual - g;
    const double error = std::sqrt(cvm::dot(diff, diff));
    if (!std::isfinite(error) || error > 0.003 + 0.003 * std::sqrt(cvm::dot(g, g)))
      throw CvError{CV_ERR_GRADIENT_CHECK, "GPU gradient mismatch at " + std::to_string(v) + ": " +
                                               fmt_dvec3(actual.x, actual.y, actual.z) + " vs " +
                                               fmt_dvec3(g.x, g.y, g.z) + ", error " + fmt_g(error)};
    max_error = std::max(max_error, error);
    const double energy = acc.e + 0.5 * inertia * cvm::dot(xv - tv, xv - tv);
    if (std::fabs(energy - probe[4 * v + 3]) > 1e-5 + 0.001 * std::fabs(energy))
      throw CvError{CV_ERR_GRADIENT_CHECK, "GPU energy mismatch at " + std::to_string(v)};
  }
  return max_error;
}

uint32_t greedy_colors(const cv_solver& s, const uint32_t* cnb_off, const uint32_t* cnb,
                       const cv_grab* grab, std::vector<uint32_t>& color_of,
                       std::vector<uint32_t>& order, std::vector<uint32_t>& color_off, bool smallest_last) {
  uint32_t n = s.n;
  const uint32_t UNSET = 0xFFFFFFFFu;
  color_of.assign(n, UNSET);
  std::vector<uint32_t> seen(n + 2, UNSET);
  uint32_t ncolors = 0;
  std::vector<uint32_t> count;
  auto in_grab = [&](uint32_t v) {
    return grab && (grab->vertices[0] == v || grab->vertices[1] == v || grab->vertices[2] == v);
  };
  for (uint32_t idx = 0; idx < n; ++idx) {
    const uint32_t v = smallest_last ? s.sl_order[idx] : idx;
    auto mark = [&](uint32_t i) {
      if (i < n && color_of[i] != UNSET) seen[color_of[i]] = v;
    };
    for (uint32_t k = s.graph_off[v]; k < s.graph_off[v + 1]; ++k) mark(s.graph[k]);
    if (cnb_off)
      for (uint32_t k = cnb_off[v]; k < cnb_off[v + 1]; ++k) mark(cnb[k]);
    if (in_grab(v))
      for (int k = 0; k < 3; ++k) mark(grab->vertices[k]);
    uint32_t c = 0;
    while (seen[c] == v) ++c;
    if (c == ncolors) {
      ncolors++;
      count.push_back(0);
    }
    color_of[v] = c;
    count[c]++;
  }
  color_off.assign(ncolors + 1, 0);
  for (uint32_t c = 0; c < ncolors; ++c) color_off[c + 1] = color_off[c] + count[c];
  order.assign(n, 0);
  std::vector<uint32_t> cur(color_off.begin(), color_off.end() - 1);
  for (uint32_t v = 0; v < n; ++v) order[cur[color_of[v]]++] = v;
  return ncolors;
}


void ContactNb::build(uint32_t n, const uint32_t* kind_ids, const uint32_t* counts, uint64_t m) {
  const uint32_t nb = n / 64 + 1;
  off.assign(nb + 1, 0);
  for (uint64_t c = 0; c < m; ++c)
    for (uint32_t q = 0; q < counts[c]; ++q) {
      const uint32_t v = kind_ids[5 * c + 1 + q];
      if (v < n) off[(v >> 6) + 1]++;
    }
  for (uint32_t b = 0; b < nb; ++b) off[b + 1] += off[b];
  pairs.resize(off[nb]);
  std::vector<uint32_t> cur(off.begin(), off.end() - 1);
  for (uint64_t c = 0; c < m; ++c)
    for (uint32_t q = 0; q < counts[c]; ++q) {
      const uint32_t v = kind_ids[5 * c + 1 + q];
      if (v < n) pairs[cur[v >> 6]++] = (uint64_t)v << 32 | (uint32_t)c;
    }
}

uint32_t update_colors(const cv_solver& s, ColoringState& cs, const uint32_t* kind_ids, const uint32_t* counts,
                       uint64_t m, const cv_grab* grab, bool smallest_last, std::vector<uint32_t>& color_off) {
  const uint32_t n = s.n;
  uint32_t key[4] = {grab ? 1u : 0u, grab ? grab->vertices[0] : 0u, grab ? grab->vertices[1] : 0u,
                     grab ? grab->vertices[2] : 0u};
  cs.changed.clear();
  auto finish = [&] {
    cs.prev.resize(6 * m);
    for (uint64_t c = 0; c < m; ++c) {
      std::memcpy(&cs.prev[6 * c], kind_ids + 5 * c, 5 * sizeof(uint32_t));
      cs.prev[6 * c + 5] = counts[c];
    }
    cs.valid = true;
    cs.smallest_last = smallest_last;
    cs.generation = s.generation;
    std::memcpy(cs.grab_key, key, sizeof key);
    uint32_t ncolors = 0;
    for (uint32_t c = (uint32_t)cs.count.size(); c > 0; --c)
      if (cs.count[c - 1]) {
        ncolors = c;
        break;
      }
    color_off.assign(ncolors + 1, 0);
    for (uint32_t c = 0; c < ncolors; ++c) color_off[c + 1] = color_off[c] + cs.count[c];
    return ncolors;
  };
  if (!cs.valid || cs.generation != s.generation || cs.smallest_last != smallest_last) {
    // From scratch, on the dense per-vertex contact neighbour lists.
    std::vector<uint32_t> off(n + 1, 0), nb, order, coff;
    for (uint64_t c = 0; c < m; ++c)
      for (uint32_t q = 0; q < counts[c]; ++q) {
        const uint32_t v = kind_ids[5 * c + 1 + q];
        if (v < n) off[v + 1] += counts[c];
      }
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    nb.resize(off[n]);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint64_t c = 0; c < m; ++c)
      for (uint32_t q = 0; q < counts[c]; ++q) {
        const uint32_t v = kind_ids[5 * c + 1 + q];
        if (v < n)
          for (uint32_t r = 0; r < counts[c]; ++r) nb[cur[v]++] = kind_ids[5 * c + 1 + r];
      }
    greedy_colors(s, off.data(), nb.data(), grab, cs.color_of, order, coff, smallest_last);
    cs.pos.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) cs.pos[smallest_last ? s.sl_order[i] : i] = i;
    cs.count.assign(coff.size() - 1, 0);
    for (size_t c = 0; c + 1 < coff.size(); ++c) cs.count[c] = coff[c + 1] - coff[c];
    cs.dirty.assign(n / 64 + 1, 0);
    cs.seen.assign(n + 2, 0);
    cs.stamp = 0;
    cs.full = true;
    return finish();
  }
  cs.full = false;
  const ContactNb& nx = cs.nb;
  cs.nb.build(n, kind_ids, counts, m);
  // Vertices to recolour, as a bitmap over colouring positions scanned upward
  // (the order a from-scratch pass reaches them): a recoloured vertex only
  // marks later positions, which the same scan reaches.
  std::vector<uint64_t>& bits = cs.dirty;
  auto enqueue = [&](uint32_t v) {
    if (v >= n) return;
    const uint32_t p = cs.pos[v];
    bits[p >> 6] |= 1ull << (p & 63);
  };
  // Contacts in only one of the old and new lists: their vertices' contact
  // neighbourhoods may have changed.
  const uint64_t mo = cs.prev.size() / 6;
  auto cmp = [&](uint64_t i, uint64_t j) {  // old row i against new row j
    const uint32_t* a = &cs.prev[6 * i];
    const uint32_t* b = kind_ids + 5 * j;
    for (int k = 0; k < 5; ++k)
      if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
    return a[5] == counts[j] ? 0 : (a[5] < counts[j] ? -1 : 1);
  };
  auto mark_old = [&](uint64_t i) {
    for (uint32_t q = 0; q < cs.prev[6 * i + 5]; ++q) enqueue(cs.prev[6 * i + 1 + q]);
  };
  auto mark_new = [&](uint64_t j) {
    for (uint32_t q = 0; q < counts[j]; ++q) enqueue(kind_ids[5 * j + 1 + q]);
  };
  for (uint64_t i = 0, j = 0; i < mo || j < m;) {
    if (j == m) {
      mark_old(i++);
    } else if (i == mo) {
      mark_new(j++);
    } else {
      const int c = cmp(i, j);
      if (c < 0)
        mark_old(i++);
      else if (c > 0)
        mark_new(j++);
      else
        ++i, ++j;
    }
  }
  if (std::memcmp(key, cs.grab_key, sizeof key) != 0)
    for (int k = 1; k < 4; ++k) {
      if (cs.grab_key[0] && cs.grab_key[k] < n) enqueue(cs.grab_key[k]);
      if (key[0] && key[k] < n) enqueue(key[k]);
    }
  auto in_grab = [&](uint32_t v) {
    return grab && (grab->vertices[0] == v || grab->vertices[1] == v || grab->vertices[2] == v);
  };
  for (size_t w = 0; w < bits.size(); ++w) {
    while (bits[w]) {
      const uint32_t pv = (uint32_t)(w * 64 + __builtin_ctzll(bits[w]));
      bits[w] &= bits[w] - 1;
      const uint32_t v = smallest_last ? s.sl_order[pv] : pv;
      // Colour from the neighbours that come earlier in the order (the ones a
      // from-scratch pass has coloured by now).
      const uint64_t st = ++cs.stamp;
      auto mark = [&](uint32_t i) {
        if (i < n && cs.pos[i] < pv) cs.seen[cs.color_of[i]] = st;
      };
      for (uint32_t k = s.graph_off[v]; k < s.graph_off[v + 1]; ++k) mark(s.graph[k]);
      nx.for_rows(v, [&](uint32_t r) {
        for (uint32_t q = 0; q < counts[r]; ++q) mark(kind_ids[5 * r + 1 + q]);
      });
      if (in_grab(v))
        for (int k = 0; k < 3; ++k) mark(grab->vertices[k]);
      uint32_t c = 0;
      while (cs.seen[c] == st) ++c;
      if (c == cs.color_of[v]) continue;
      cs.count[cs.color_of[v]]--;
      if (c >= cs.count.size()) cs.count.resize(c + 1, 0);
      cs.count[c]++;
      cs.color_of[v] = c;
      cs.changed.push_back(v);
      // Later neighbours may now choose differently.
      auto touch = [&](uint32_t i) {
        if (i < n && cs.pos[i] > pv) enqueue(i);
      };
      for (uint32_t k = s.graph_off[v]; k < s.graph_off[v + 1]; ++k) touch(s.graph[k]);
      nx.for_rows(v, [&](uint32_t r) {
        for (uint32_t q = 0; q < counts[r]; ++q) touch(kind_ids[5 * r + 1 + q]);
      });
      if (in_grab(v))
        for (int k = 0; k < 3; ++k) touch(grab->vertices[k]);
    }
  }
  return finish();
}
