// ogc_device.hip — Offset Geometric Contact detection on the device
// (OgcWorld::detect, src/ogc/world.rs:308-420, moved off the host).
//
// f64 throughout and compiled with -ffp-contract=off: the broad phase tests
// the reference's own predicate `Bounds::distance_squared <= padded_radius^2`
// (src/ogc/world.rs:352-358, 462-472, 499-509) on f64 boxes, so the set of
// candidate pairs equals the f64 restatement's bit for bit, and the narrow
// phase (closest features, feasible regions, canonical ids) rounds exactly as
// the reference does.
//
// Broad phase: a uniform hash grid.  Every triangle (edge) is registered in
// the cells its box inflated by the padded radius covers; entries are bucketed
// by a hash of the cell key with a counting sort (bucket ranks from one
// atomic per distinct bucket and wavefront, one scan, one scatter), and each
// bucket entry keeps its cell key so hash collisions are filtered exactly.  A
// point looks only in its own cell; an edge scans the cells of its own box
// and keeps a pair only in the cell holding the lower corner of the box
// intersection, so each pair is examined once.  The queries are registered
// in the same counting sort (points and moving-edge cells in tables of their
// own), so the queries of one bucket are contiguous: the broad pass (k_pairs)
// holds a chunk of 64 bucket entries in a wavefront's lanes, one per lane, and
// steps through the bucket's queries, each query's data read once from LDS
// by all lanes.  Candidates go to per-wavefront LDS staging and then to the
// narrow pass's list.  Pairs are emitted with atomics and then put in the
// reference's canonical (kind, ids) order by a two-key radix sort, so results
// are deterministic.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gpu.hpp"
#include "ogc_math.hpp"

using ogm::D3;

namespace {

constexpr int TPB = 256;
constexpr long long kMaxCells = 1ll << 22;  // grid cells one primitive may cover

__device__ __forceinline__ D3 ld3(const double* x, uint32_t i) {
  return D3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
}
__device__ __forceinline__ ogm::Box ldbox(const double* b, uint32_t i) {
  return ogm::Box{{b[6 * i], b[6 * i + 1], b[6 * i + 2]}, {b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]}};
}
__device__ __forceinline__ void stbox(double* b, uint32_t i, const ogm::Box& v) {
  b[6 * i] = v.lo.x;
  b[6 * i + 1] = v.lo.y;
  b[6 * i + 2] = v.lo.z;
  b[6 * i + 3] = v.hi.x;
  b[6 * i + 4] = v.hi.y;
  b[6 * i + 5] = v.hi.z;
}
// Grid cell of a coordinate, from the inverse cell size (a multiply instead of
// an f64 division).  Any monotone map is a valid grid: registration and
// queries use this one function.
__device__ __forceinline__ long long cell_of(double v, double inv_h) { return (long long)floor(v * inv_h); }
__device__ __forceinline__ unsigned long long cell_key(long long i, long long j, long long k) {
  return ((unsigned long long)(i & 0x1FFFFF) << 42) | ((unsigned long long)(j & 0x1FFFFF) << 21) |
         (unsigned long long)(k & 0x1FFFFF);
}
__device__ __forceinline__ ogm::Box inflate(const ogm::Box& b, double r) {
  return {{b.lo.x - r, b.lo.y - r, b.lo.z - r}, {b.hi.x + r, b.hi.y + r, b.hi.z + r}};
}

// Primitive boxes of faces and edges; degenerate moving faces flag an error
// (src/ogc/world.rs:225-233).
__global__ void k_boxes(const double* x, const uint32_t* faces, const uint8_t* face_moving,
                        uint32_t F, const uint32_t* edges, uint32_t E, double* fbox, double* ebox,
                        Scalars* sc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) {
    D3 a = ld3(x, faces[3 * i]), b = ld3(x, faces[3 * i + 1]), c = ld3(x, faces[3 * i + 2]);
    stbox(fbox, i, ogm::box_of(a, b, c));
    if (face_moving[i] && ogm::len2(ogm::cross(ogm::sub(b, a), ogm::sub(c, a))) < 1e-24)
      atomicOr(&sc->err_degenerate, 1u);
  } else if (i < F + E) {
    uint32_t e = i - F;
    stbox(ebox, e, ogm::box_of(ld3(x, edges[2 * e]), ld3(x, edges[2 * e + 1])));
  }
}

__device__ __forceinline__ void cell_range(const ogm::Box& b, double inv_h, long long r[6]) {
  r[0] = cell_of(b.lo.x, inv_h);
  r[1] = cell_of(b.lo.y, inv_h);
  r[2] = cell_of(b.lo.z, inv_h);
  r[3] = cell_of(b.hi.x, inv_h);
  r[4] = cell_of(b.hi.y, inv_h);
  r[5] = cell_of(b.hi.z, inv_h);
}

__device__ __forceinline__ uint32_t bucket_of(unsigned long long key, int log2t) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - log2t));
}
// The registered primitives, in index order: faces [0, F) and edges
// [F, F+E) with their boxes inflated by the padded radius (the grid's
// entries, tables 0 and 1), then the queries: points [F+E, F+E+P) in their
// own cell (table 2) and moving edges [F+E+P, F+2E+P) in the cells of their
// own box (table 3; a fixed edge registers no cell).
struct RegPrims {
  const double* fbox;
  const double* ebox;
  const double* x;
  const uint8_t* edge_moving;
  uint32_t F, E, P;
  double pr;
  __device__ __forceinline__ uint32_t count() const { return F + 2 * E + P; }
  // Box, table and primitive id of registered primitive i; false: no cells.
  __device__ __forceinline__ bool box(uint32_t i, ogm::Box& b, uint32_t& table, uint32_t& prim) const {
    if (i < F) {
      b = inflate(ldbox(fbox, i), pr);
      table = 0;
      prim = i;
    } else if (i < F + E) {
      prim = i - F;
      b = inflate(ldbox(ebox, prim), pr);
      table = 1;
    } else if (i < F + E + P) {
      prim = i - F - E;
      b = ogm::box_of(ld3(x, prim));
      table = 2;
    } else {
      prim = i - F - E - P;
      if (!edge_moving[prim]) return false;
      b = ldbox(ebox, prim);
      table = 3;
    }
    return true;
  }
};

// Cells covered by each registered primitive.
__global__ void k_reg_count(RegPrims rp, double inv_h, uint32_t* count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rp.count()) return;
  ogm::Box bx;
  uint32_t table, prim;
  if (!rp.box(i, bx, table, prim)) {
    count[i] = 0;
    return;
  }
  long long r[6];
  cell_range(bx, inv_h, r);
  long long c = (r[3] - r[0] + 1) * (r[4] - r[1] + 1) * (r[5] - r[2] + 1);
  count[i] = (uint32_t)(c > kMaxCells ? 1 : c);  // over the cap: one sentinel entry, see k_reg_fill
}
// Owner primitive of every entry (entries of primitive i are off[i]..off[i+1]).
// Entries past the registration capacity raise the capacity flag (the host
// grows the buffers and detects again); thread 0 records the entry count.
__global__ void k_reg_owner(const uint32_t* off, uint32_t N, uint32_t* owner, uint64_t cap, Scalars* sc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) sc->grid_entries = off[N];
  if (i >= N) return;
  if (off[N] > cap) {
    if (i == 0) atomicOr(&sc->err_overflow, 8u);
    return;
  }
  for (uint32_t o = off[i]; o < off[i + 1]; ++o) owner[o] = i;
}
// One thread per entry: its cell key, primitive id and rank in its bucket.
// Table t uses buckets [tT, (t+1)T).  The grid covers the capacity; the entry
// count M = off[N] is read here.
__global__ void k_reg_fill(RegPrims rp, double inv_h, const uint32_t* off, const uint32_t* owner, uint64_t cap,
                           int log2t, unsigned long long* keys, uint32_t* prims, uint32_t* ranks, uint32_t* bcount,
                           Scalars* sc) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t N = rp.count(), M = off[N];
  if (M > cap || j >= M) return;
  uint32_t i = owner[j];
  ogm::Box bx;
  uint32_t table = 0, prim = 0;
  rp.box(i, bx, table, prim);  // true: primitive i owns an entry
  long long r[6];
  cell_range(bx, inv_h, r);
  long long nx = r[3] - r[0] + 1, ny = r[4] - r[1] + 1, nz = r[5] - r[2] + 1;
  unsigned long long key;
  if (nx * ny * nz > kMaxCells) {  // over the per-primitive cap: report, never drop silently
    key = ~0ull;                    // sentinel key never matches a cell (cell keys use 63 bits)
    atomicOr(&sc->err_overflow, 2u);
  } else {
    // 32-bit index arithmetic: k < nx ny nz <= kMaxCells.
    const uint32_t k = j - off[i], uz = (uint32_t)nz, uy = (uint32_t)ny;
    const uint32_t kz = k / uz, z = k - kz * uz, x = kz / uy, y = kz - x * uy;
    key = cell_key(r[0] + x, r[1] + y, r[2] + z);
  }
  keys[j] = key;
  prims[j] = prim;
  // Lanes of a wavefront with the same bucket (neighbouring primitives cover
  // the same cells) take their ranks from one atomic: the lanes are grouped
  // by bucket first (a wave-uniform loop over the active lanes, no memory
  // traffic), then every group's leader adds its group's size in the same
  // instruction, so a wavefront waits for one atomic round trip.
  const uint32_t bk = (table << log2t) + bucket_of(key, log2t), lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1);
  uint32_t leader_of = lane, within = 0, group = 0;
  while (todo) {
    const uint32_t leader = __ffsll(todo) - 1;
    const uint32_t bl = __shfl(bk, leader, 64);
    const unsigned long long peers = __ballot(bk == bl);
    if (bk == bl) {
      leader_of = leader;
      within = (uint32_t)__popcll(peers & ((1ull << lane) - 1));
    }
    if (lane == leader) group = (uint32_t)__popcll(peers);
    todo &= ~peers;
  }
  uint32_t base = 0;
  if (lane == leader_of) base = atomicAdd(&bcount[bk], group);
  base = __shfl(base, leader_of, 64);
  ranks[j] = base + within;
}
// Bucketed entries carry the primitive's (unpadded) box, a query entry its
// own box (a point's: the point twice), so the broad pass reads contiguous
// memory and tests boxes without a dependent gather; each entry is written as
// one 64-byte line (four 16-byte stores).
__global__ void k_reg_scatter(RegPrims rp, const unsigned long long* keys, const uint32_t* prims,
                              const uint32_t* ranks, const uint32_t* off, uint64_t cap, int log2t,
                              const uint32_t* bstart, GridEntry* ent) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t N = rp.count(), M = off[N];
  if (M > cap || j >= M) return;
  // Tables in entry order: entries of faces, edges, points, moving edges.
  const uint32_t table = (j >= off[rp.F]) + (j >= off[rp.F + rp.E]) + (j >= off[rp.F + rp.E + rp.P]);
  const unsigned long long key = keys[j];
  uint32_t pos = bstart[(table << log2t) + bucket_of(key, log2t)] + ranks[j];
  uint32_t p = prims[j];
  double2* dst = (double2*)(ent + pos);
  dst[0] = make_double2(__longlong_as_double((long long)key), __longlong_as_double((long long)p));
  if (table == 2) {
    const D3 q = ld3(rp.x, p);
    dst[1] = make_double2(q.x, q.y);
    dst[2] = make_double2(q.z, q.x);
    dst[3] = make_double2(q.y, q.z);
  } else {
    const double2* src = (const double2*)(table == 0 ? rp.fbox + 6 * (size_t)p : rp.ebox + 6 * (size_t)p);
    dst[1] = src[0];
    dst[2] = src[1];
    dst[3] = src[2];
  }
}

struct DetectArgs {
  const double* x;
  uint32_t n, n_total, F, E;
  const uint32_t* faces;
  const uint32_t* edges;
  const uint32_t* tri_edges;
  const int32_t* opp;
  const uint32_t* nb_off;
  const uint32_t* nb;
  const uint32_t* ex_off;
  const uint32_t* ex;
  const uint8_t* face_moving;
  const uint8_t* edge_moving;
  const double* rest_len2;
  const double* fbox;
  const double* ebox;
  double rq, pr, pr2, inv_h;
  const GridEntry* ent;            // bucketed entries: cell key, primitive id and box
  const uint32_t* bstart;          // [4T+1] bucket starts: faces, edges, points, moving-edge cells
  int log2t;
  unsigned long long* dist_bits;
  uint32_t* raw_ids;
  uint32_t* raw_count;
  double* raw_eps;
  uint64_t cap;
  uint32_t* cand_vf;
  uint32_t* cand_ee;
  uint64_t cand_cap;
  Scalars* sc;
};

__device__ __forceinline__ bool is_excluded(const DetectArgs& a, uint32_t u, uint32_t w) {
  uint32_t lo = a.ex_off[u], hi = a.ex_off[u + 1];
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t v = a.ex[mid];
    if (v == w) return true;
    if (v < w)
      lo = mid + 1;
    else
      hi = mid;
  }
  return false;
}
// Per-vertex minimum distance as the bit pattern of a non-negative double
// (ordered like the value).  Values only decrease, so a plain read that is
// already <= d (even a stale one) proves the atomic would change nothing:
// most pairs skip it, which removes the same-address atomic contention of
// vertices that appear in hundreds of candidate pairs.
__device__ __forceinline__ void dmin(const DetectArgs& a, uint32_t v, double d) {
  unsigned long long bits = (unsigned long long)__double_as_longlong(d);
  if (a.dist_bits[v] <= bits) return;
  atomicMin(&a.dist_bits[v], bits);
}
// A contact found by one narrow-phase test (at most one per candidate pair),
// in the canonical id order; the narrow pass stages the found ones per
// wavefront and allocates their raw-list slots with one atomic per flush.
struct RawContact {
  uint32_t kind, ids[4];
  uint32_t count;
  double eps;
  bool valid;
};
__device__ __forceinline__ void set_contact(RawContact& out, int kind, uint32_t i0, uint32_t i1, uint32_t i2,
                                            uint32_t i3, int count, double eps) {
  uint32_t ids[4] = {i0, i1, i2, i3};
  ogm::canonical(kind, ids);
  out.kind = (uint32_t)kind;
  for (int k = 0; k < 4; ++k) out.ids[k] = ids[k];
  out.count = (uint32_t)count;
  out.eps = eps;
  out.valid = true;
}

struct PairCounts {
  unsigned long long tests = 0, vf_cand = 0, vf_near = 0, ee_cand = 0, ee_near = 0;
};

// Detection runs in two passes.  The broad pass takes one (query, grid
// entry) pair per thread and applies the cheap filters below; the pairs that
// survive (the candidates of src/ogc/candidates.rs) are appended to a list,
// and the narrow pass runs the closest-feature geometry over that list with
// every lane busy.  Per-vertex distance minima and the contact set do not
// depend on the order in which pairs are examined (contacts are sorted and
// deduplicated afterwards), so the split leaves results unchanged.

// The face data the vertex-face filter needs, and the edge data of the
// edge-edge filter: loaded when the filter first needs them (the BVH walk's
// leaves) or held per lane for a whole chunk of queries (the grid pass).
struct FaceRec {
  uint32_t v0, v1, v2;
  bool moving;
};
struct EdgeRec {
  uint32_t c, d;
  bool moving;
};
__device__ __forceinline__ FaceRec load_face(const DetectArgs& a, uint32_t t) {
  return {a.faces[3 * t], a.faces[3 * t + 1], a.faces[3 * t + 2], a.face_moving[t] != 0};
}
__device__ __forceinline__ EdgeRec load_edge(const DetectArgs& a, uint32_t f) {
  return {a.edges[2 * f], a.edges[2 * f + 1], a.edge_moving[f] != 0};
}

// Vertex-face filter (src/ogc/world.rs:461-470): padded box test, moving
// check, incidence and seam exclusion.  `face()` yields face t's FaceRec.
template <class FaceOf>
__device__ __forceinline__ bool vf_filter(const DetectArgs& a, uint32_t v, bool v_moving, bool v_ex,
                                          const ogm::Box& pb, uint32_t t, const ogm::Box& tb, PairCounts& c,
                                          FaceOf face) {
  ++c.tests;
  if (ogm::box_dist2(pb, tb) > a.pr2) return false;
  const FaceRec fr = face();
  if (!v_moving && !fr.moving) return false;
  const uint32_t t0 = fr.v0, t1 = fr.v1, t2 = fr.v2;
  if (v == t0 || v == t1 || v == t2) return false;
  if (v_ex && (is_excluded(a, v, t0) || is_excluded(a, v, t1) || is_excluded(a, v, t2))) return false;
  ++c.vf_cand;
  if (a.cand_vf) {
    unsigned long long s = atomicAdd(&a.sc->cand_export_vf, 1ull);
    if (s < a.cand_cap) {
      a.cand_vf[2 * s] = v;
      a.cand_vf[2 * s + 1] = t;
    }
  }
  return true;
}
// Vertex-face narrow phase (src/ogc/world.rs:471-497): closest point,
// distance, minima, feasible feature, contact.
__device__ __forceinline__ void vf_narrow(const DetectArgs& a, uint32_t v, uint32_t t, PairCounts& c,
                                          RawContact& out) {
  uint32_t t0 = a.faces[3 * t], t1 = a.faces[3 * t + 1], t2 = a.faces[3 * t + 2];
  D3 p = ld3(a.x, v);
  D3 q;
  int feat, loc;
  ogm::closest_triangle(p, ld3(a.x, t0), ld3(a.x, t1), ld3(a.x, t2), q, feat, loc);
  double d = ogm::dist(p, q);
  if (d >= a.rq) return;
  ++c.vf_near;
  dmin(a, v, d);
  dmin(a, t0, d);
  dmin(a, t1, d);
  dmin(a, t2, d);
  if (feat == ogm::F_VERTEX) {
    uint32_t w = loc == 0 ? t0 : (loc == 1 ? t1 : t2);
    if (ogm::vertex_feasible(p, (int)w, (const D3*)a.x, a.nb_off, a.nb)) set_contact(out, 0, v, w, 0, 0, 2, 0.0);
  } else if (feat == ogm::F_EDGE) {
    uint32_t e = a.tri_edges[3 * t + loc];
    uint32_t ea = a.edges[2 * e], eb = a.edges[2 * e + 1];
    if (ogm::edge_feasible(p, ld3(a.x, ea), ld3(a.x, eb), &a.opp[2 * e], (const D3*)a.x))
      set_contact(out, 1, v, ea, eb, 0, 3, 0.0);
  } else {
    set_contact(out, 2, v, t0, t1, t2, 4, 0.0);
  }
}

// Edge-edge filter (src/ogc/world.rs:498-512).  The moving edge e is the
// first operand; moving/moving pairs are taken once with e < f.  `edge()`
// yields edge f's EdgeRec; `in_cell()` is the grid's rule that a pair met in
// several cells is kept only in the cell holding the lower corner of
// box(e) ∩ inflated box(f) (a BVH walk meets each pair once: always true).
template <class EdgeOf, class InCell>
__device__ __forceinline__ bool ee_filter(const DetectArgs& a, uint32_t e, const ogm::Box& be, uint32_t ia,
                                          uint32_t ib, bool ia_ex, bool ib_ex, uint32_t f, const ogm::Box& bfr,
                                          PairCounts& c, EdgeOf edge, InCell in_cell) {
  ++c.tests;
  if (ogm::box_dist2(be, bfr) > a.pr2) return false;
  if (f == e) return false;
  const EdgeRec er = edge();
  if (er.moving && f < e) return false;
  if (!in_cell()) return false;
  const uint32_t jc = er.c, jd = er.d;
  if (ia == jc || ia == jd || ib == jc || ib == jd) return false;
  if ((ia_ex && (is_excluded(a, ia, jc) || is_excluded(a, ia, jd))) ||
      (ib_ex && (is_excluded(a, ib, jc) || is_excluded(a, ib, jd))))
    return false;
  ++c.ee_cand;
  if (a.cand_ee) {
    unsigned long long s = atomicAdd(&a.sc->cand_export_ee, 1ull);
    if (s < a.cand_cap) {
      a.cand_ee[2 * s] = e;
      a.cand_ee[2 * s + 1] = f;
    }
  }
  return true;
}
// Edge-edge narrow phase (src/ogc/world.rs:513-555).
__device__ __forceinline__ void ee_narrow(const DetectArgs& a, uint32_t e, uint32_t f, PairCounts& c,
                                          RawContact& out) {
  uint32_t ia = a.edges[2 * e], ib = a.edges[2 * e + 1];
  uint32_t jc = a.edges[2 * f], jd = a.edges[2 * f + 1];
  D3 A = ld3(a.x, ia), B = ld3(a.x, ib), C = ld3(a.x, jc), Dd = ld3(a.x, jd);
  double s, t;
  ogm::closest_segments(A, B, C, Dd, s, t);
  D3 P = ogm::add(A, ogm::scale(ogm::sub(B, A), s));
  D3 Q = ogm::add(C, ogm::scale(ogm::sub(Dd, C), t));
  double dist = ogm::dist(P, Q);
  if (dist >= a.rq) return;
  ++c.ee_near;
  dmin(a, ia, dist);
  dmin(a, ib, dist);
  dmin(a, jc, dist);
  dmin(a, jd, dist);
  int end_a = s <= 0.0 ? (int)ia : (s >= 1.0 ? (int)ib : -1);
  int end_b = t <= 0.0 ? (int)jc : (t >= 1.0 ? (int)jd : -1);
  if ((end_a >= 0 && !ogm::vertex_feasible(Q, end_a, (const D3*)a.x, a.nb_off, a.nb)) ||
      (end_b >= 0 && !ogm::vertex_feasible(P, end_b, (const D3*)a.x, a.nb_off, a.nb)))
    return;
  if (end_a >= 0 && end_b >= 0)
    set_contact(out, 0, end_a, end_b, 0, 0, 2, 0.0);
  else if (end_a >= 0)
    set_contact(out, 1, end_a, jc, jd, 0, 3, 0.0);
  else if (end_b >= 0)
    set_contact(out, 1, end_b, ia, ib, 0, 3, 0.0);
  else
    set_contact(out, 3, ia, ib, jc, jd, 4, 1e-3 * a.rest_len2[e] * a.rest_len2[f]);
}

// The pair statistics summed over the block and added with one atomic per
// counter and block: the counters are five addresses shared by the whole
// device, where atomics serialise.  Every thread of the block calls it once,
// at the end of the kernel.
__device__ __forceinline__ void add_counts(const DetectArgs& a, const PairCounts& c) {
  constexpr int NW = TPB / 64;
  __shared__ unsigned long long part[NW][5];
  unsigned long long v[5] = {c.tests, c.vf_cand, c.vf_near, c.ee_cand, c.ee_near};
#pragma unroll
  for (int k = 0; k < 5; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 5) {
    unsigned long long sum = 0;
    for (uint32_t i = 0; i < blockDim.x / 64; ++i) sum += part[i][threadIdx.x];
    unsigned long long* const dst[5] = {&a.sc->grid_tests, &a.sc->vf_cand, &a.sc->vf_count, &a.sc->ee_cand,
                                        &a.sc->ee_count};
    if (sum) atomicAdd(dst[threadIdx.x], sum);
  }
}

// Per-wavefront candidate staging of the broad pass (grid or LBVH walks): a
// lane with a candidate appends it to its wavefront's LDS buffer (positions
// from a ballot prefix count), and a wavefront whose buffer could overflow on
// its next step writes the buffer out with one global atomic: same-address
// global atomics per test step would serialise across the whole device.  The
// loops that stage are wavefront-uniform and need no block barrier.  Entries
// beyond `cap` are counted but not written: the capacity flag (32) makes the
// host grow the region and detect again.
template <int NB>
struct WaveStage {
  static constexpr uint32_t kCap = NB;
  uint2 buf[NB];
  uint32_t n;
};
template <class WS>
__device__ __forceinline__ void wave_stage(WS& ws, bool take, uint32_t q, uint32_t p) {
  const uint64_t mask = __ballot(take);
  if (!mask) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t base = ws.n;  // wavefront-uniform
  if (take) ws.buf[base + __popcll(mask & ((1ull << lane) - 1))] = make_uint2(q, p);
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) ws.n = base + (uint32_t)__popcll(mask);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
template <class WS>
__device__ __forceinline__ void wave_flush(WS& ws, unsigned long long* total, uint64_t cap, uint2* region,
                                           Scalars* sc) {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t n = ws.n, lane = threadIdx.x & 63;
  if (!n) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(total, (unsigned long long)n);
  base = __shfl(base, 0, 64);
  for (uint32_t k = lane; k < n; k += 64)
    if (base + k < cap) region[base + k] = ws.buf[k];
  if (lane == 0 && base + n > cap) atomicOr(&sc->err_overflow, 32u);
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) ws.n = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// LDS written by some lanes of a wavefront and read by others: a wavefront's
// LDS operations execute in order, the fences keep the compiler from moving
// them across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Broad-pass counters, each on its own 128-byte line: same-address atomics
// from the whole device serialise, so the grid pass keeps one candidate
// counter per kind and XCD (and one span counter per XCD), and its candidate
// regions are split into kSub parts, part x filled by the blocks on XCD x.
// The LBVH walks use part 0 of each kind as one region.
constexpr uint32_t kSub = 8;
constexpr uint32_t kCtrStride = 16;  // u64 per counter line
constexpr uint32_t kCtrCount = 3 * kSub;  // VF parts, EE parts, span counters
__device__ __forceinline__ unsigned long long* list_ctr(unsigned long long* ctr, uint32_t kind, uint32_t sub) {
  return ctr + (size_t)(kind * kSub + sub) * kCtrStride;
}

// Grid broad pass, bucket-major.  The query entries (tables 2 and 3 of the
// registration: points, then the cells of moving edges) are taken in spans of
// 64 consecutive entries per wavefront, the spans split evenly over the
// XCDs, each XCD's wavefronts taking them from its own counter.  A span's queries
// are staged in LDS (box, cell key, the bucket's entry range, ids, flags) and
// fall into runs of consecutive queries of one bucket, since the registration
// put each bucket's queries together.  For each run the wavefront loads the
// bucket's entries in chunks of 64, one per lane with its face or edge data,
// and tests every query of the run against its lanes' entries, the query's
// data read by all lanes at one LDS address: each entry is loaded once per
// run instead of once per (query, entry) pair.  The pairs examined, and so the
// candidates, are those of each query scanning its own bucket.
constexpr uint32_t kSpan = 64;
struct SpanQuery {
  double box[6];          // a point's box is the point twice
  unsigned long long key; // the query's cell
  uint32_t s, e;          // the bucket's entries
  uint32_t id;            // point or edge
  uint32_t fl;            // point: moving | excluded list << 1; edge: a excluded | b excluded << 1
  uint32_t qa, qb;        // an edge's vertices
  long long c[3];         // an edge: the cell of its box's lower corner
};
// A query as the test loop holds it in registers (read from LDS one query
// ahead of its use, so the read's latency overlaps the previous query's tests).
struct QueryRegs {
  ogm::Box box;
  unsigned long long key;
  uint32_t id, fl, qa, qb;
  long long c0, c1, c2;
};
template <bool EDGE>
__device__ __forceinline__ QueryRegs query_regs(const SpanQuery& q) {
  QueryRegs r;
  r.box = ogm::Box{{q.box[0], q.box[1], q.box[2]}, {q.box[3], q.box[4], q.box[5]}};
  r.key = q.key;
  r.id = q.id;
  r.fl = q.fl;
  if (EDGE) {
    r.qa = q.qa;
    r.qb = q.qb;
    r.c0 = q.c[0];
    r.c1 = q.c[1];
    r.c2 = q.c[2];
  } else {
    r.qa = r.qb = 0;
    r.c0 = r.c1 = r.c2 = 0;
  }
  return r;
}
// Candidate staging with the staged count in a register (wavefront-uniform):
// a step's candidates are placed by a ballot prefix count, with no LDS read
// on the test loop's path.
__device__ __forceinline__ void stage_take(uint2* buf, uint32_t& n, bool take, uint32_t q, uint32_t p) {
  const unsigned long long mask = __ballot(take);
  const uint32_t lane = threadIdx.x & 63;
  if (take) buf[n + (uint32_t)__popcll(mask & ((1ull << lane) - 1))] = make_uint2(q, p);
  n += (uint32_t)__popcll(mask);
}
__device__ __forceinline__ void stage_flush(const uint2* buf, uint32_t& n, unsigned long long* total, uint64_t cap,
                                            uint2* region, Scalars* sc) {
  if (!n) return;
  wave_sync();  // the staged entries are read by other lanes
  const uint32_t lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(total, (unsigned long long)n);
  base = __shfl(base, 0, 64);
  for (uint32_t k = lane; k < n; k += 64)
    if (base + k < cap) region[base + k] = buf[k];
  if (lane == 0 && base + n > cap) atomicOr(&sc->err_overflow, 32u);
  wave_sync();  // read before the buffer is staged again
  n = 0;
}
// The tests of one run's queries [k0, k1) against a chunk of the bucket's
// entries, one entry per lane (`ok`: the lane holds one).
template <bool EDGE>
__device__ __forceinline__ void run_chunk(const DetectArgs& a, const SpanQuery* qs, uint32_t k0, uint32_t k1,
                                          const GridEntry& en, bool ok, uint2* buf, uint32_t& n, uint32_t ncap,
                                          unsigned long long* total, uint64_t cap, uint2* region, PairCounts& cnt) {
  const ogm::Box eb{{en.box[0], en.box[1], en.box[2]}, {en.box[3], en.box[4], en.box[5]}};
  const uint32_t prim = en.prim;
  const unsigned long long ekey = ok ? en.key : ~0ull;  // a lane past the bucket matches no query
  FaceRec fr{};
  EdgeRec er{};
  long long f0 = 0, f1 = 0, f2 = 0;
  if (ok) {
    if (EDGE) {
      er = load_edge(a, prim);
      // Cells of the inflated entry box's lower corner: the lower corner of
      // box(e) ∩ inflated box(f) is the per-axis max of the two lower
      // corners, and the cell map is monotone, so its cell is the per-axis
      // max of the two corners' cells.
      f0 = cell_of(eb.lo.x - a.pr, a.inv_h);
      f1 = cell_of(eb.lo.y - a.pr, a.inv_h);
      f2 = cell_of(eb.lo.z - a.pr, a.inv_h);
    } else {
      fr = load_face(a, prim);
    }
  }
  QueryRegs cur = query_regs<EDGE>(qs[k0]);
  for (uint32_t k = k0; k < k1; ++k) {
    const QueryRegs nxt = query_regs<EDGE>(qs[min(k + 1, k1 - 1)]);
    bool take = false;
    if (ekey == cur.key) {
      if (EDGE)
        take = ee_filter(a, cur.id, cur.box, cur.qa, cur.qb, cur.fl & 1u, cur.fl & 2u, prim, eb, cnt,
                         [&] { return er; },
                         [&] { return cell_key(max(cur.c0, f0), max(cur.c1, f1), max(cur.c2, f2)) == cur.key; });
      else
        take = vf_filter(a, cur.id, cur.fl & 1u, cur.fl & 2u, cur.box, prim, eb, cnt, [&] { return fr; });
    }
    stage_take(buf, n, take, cur.id, prim);
    if (n > ncap - 64) stage_flush(buf, n, total, cap, region, a.sc);
    cur = nxt;
  }
}
__global__ __launch_bounds__(TPB) void k_pairs(DetectArgs a, uint64_t cap_vf, uint2* vf, uint64_t cap_ee,
                                               uint2* ee, unsigned long long* ctr) {
  constexpr uint32_t kOut = 256;  // staged candidates per wavefront (with the queries: 4 wavefronts per SIMD)
  struct Stage {
    SpanQuery q[kSpan];
    uint2 out[kOut];
  };
  __shared__ Stage stage[TPB / 64];
  Stage& st = stage[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63, x = blockIdx.x % kSub;  // blocks b with b % 8 == x run on XCD x
  const uint32_t T = 1u << a.log2t;
  const uint32_t P0 = a.bstart[2 * T], PE = a.bstart[3 * T], P1 = a.bstart[4 * T];
  const uint32_t NS = (P1 - P0 + kSpan - 1) / kSpan;
  const uint32_t span_lo = (uint32_t)((uint64_t)NS * x / kSub), span_hi = (uint32_t)((uint64_t)NS * (x + 1) / kSub);
  // This XCD's candidate parts and counters.
  const uint64_t sub_vf = cap_vf / kSub, sub_ee = cap_ee / kSub;
  vf += sub_vf * x;
  ee += sub_ee * x;
  cap_vf = sub_vf;
  cap_ee = sub_ee;
  unsigned long long* const n_vf = list_ctr(ctr, 0, x);
  unsigned long long* const n_ee = list_ctr(ctr, 1, x);
  unsigned int* const span_next = (unsigned int*)list_ctr(ctr, 2, x);
  PairCounts cnt;
  uint32_t n_out = 0;   // staged candidates (wavefront-uniform)
  bool out_ee = false;  // their kind
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(span_next, 1u);
    const uint32_t span = span_lo + (uint32_t)__shfl((int)t, 0, 64);
    if (span >= span_hi) break;
    const uint64_t p0 = P0 + (uint64_t)kSpan * span;
    wave_sync();  // the previous span's queries are no longer read
    const uint64_t p = p0 + lane;
    bool member = false;
    uint32_t s = 0;
    if (p < P1) {
      const GridEntry qe = a.ent[p];
      const bool edge = p >= PE;
      uint32_t e = 0;
      if (qe.key != ~0ull) {  // a sentinel (over the cell cap, flagged) queries nothing
        const uint32_t bk = (edge ? T : 0u) + bucket_of(qe.key, a.log2t);
        s = a.bstart[bk];
        e = a.bstart[bk + 1];
        member = e > s;
      }
      if (member) {
        SpanQuery& q = st.q[lane];
#pragma unroll
        for (int k = 0; k < 6; ++k) q.box[k] = qe.box[k];
        q.key = qe.key;
        q.s = s;
        q.e = e;
        q.id = qe.prim;
        if (!edge) {
          const uint32_t v = qe.prim;
          q.fl = (v < a.n ? 1u : 0u) | (a.ex_off[v] != a.ex_off[v + 1] ? 2u : 0u);
        } else {
          const uint32_t ia = a.edges[2 * qe.prim], ib = a.edges[2 * qe.prim + 1];
          q.qa = ia;
          q.qb = ib;
          q.fl = (a.ex_off[ia] != a.ex_off[ia + 1] ? 1u : 0u) | (a.ex_off[ib] != a.ex_off[ib + 1] ? 2u : 0u);
          q.c[0] = cell_of(qe.box[0], a.inv_h);
          q.c[1] = cell_of(qe.box[1], a.inv_h);
          q.c[2] = cell_of(qe.box[2], a.inv_h);
        }
      }
    }
    wave_sync();
    // Runs: a member starts one unless the lane before it is a member of the
    // same bucket (the entry range's start identifies a non-empty bucket).
    const uint32_t prev_s = (uint32_t)__shfl((int)s, (int)((lane + 63) & 63), 64);
    const bool prev_member = __shfl((int)member, (int)((lane + 63) & 63), 64) != 0;
    const bool head = member && (lane == 0 || !prev_member || prev_s != s);
    unsigned long long heads = __ballot(head);
    const unsigned long long members = __ballot(member);
    while (heads) {
      const uint32_t k0 = (uint32_t)__ffsll((long long)heads) - 1;
      heads &= heads - 1;
      const unsigned long long after = k0 == 63 ? 0ull : (~0ull << (k0 + 1));
      const unsigned long long stop = (heads | ~members) & after;
      const uint32_t k1 = stop ? (uint32_t)__ffsll((long long)stop) - 1 : kSpan;
      const bool edge_run = p0 + k0 >= PE;
      const uint32_t rs = st.q[k0].s, re = st.q[k0].e;
      if (edge_run != out_ee) {
        stage_flush(st.out, n_out, out_ee ? n_ee : n_vf, out_ee ? cap_ee : cap_vf, out_ee ? ee : vf, a.sc);
        out_ee = edge_run;
      }
      unsigned long long* total = edge_run ? n_ee : n_vf;
      const uint64_t cap = edge_run ? cap_ee : cap_vf;
      uint2* region = edge_run ? ee : vf;
      for (uint32_t c0 = rs; c0 < re; c0 += 64) {
        const uint32_t j = c0 + lane;
        const bool ok = j < re;
        GridEntry en{};
        if (ok) en = a.ent[j];
        if (edge_run)
          run_chunk<true>(a, st.q, k0, k1, en, ok, st.out, n_out, kOut, total, cap, region, cnt);
        else
          run_chunk<false>(a, st.q, k0, k1, en, ok, st.out, n_out, kOut, total, cap, region, cnt);
      }
    }
  }
  stage_flush(st.out, n_out, out_ee ? n_ee : n_vf, out_ee ? cap_ee : cap_vf, out_ee ? ee : vf, a.sc);
  add_counts(a, cnt);  // every lane reaches it
}
// Narrow pass over the candidate regions, each kind in `nsub` equal parts
// with their own counters (the counts are the device's; a part that
// overflowed is redone by the host).  Block 0 records the totals and the
// largest part for the host.
__global__ __launch_bounds__(TPB) void k_narrow_lists(DetectArgs a, const uint2* vf, uint64_t cap_vf, const uint2* ee,
                                                      uint64_t cap_ee, const unsigned long long* ctr, uint32_t nsub) {
  // Exclusive prefix of the parts' counts: VF parts, then EE parts.
  __shared__ unsigned long long pre[2 * kSub + 1];
  const uint64_t sub_vf = cap_vf / nsub, sub_ee = cap_ee / nsub;
  if (threadIdx.x == 0) {
    unsigned long long acc = 0, tot_v = 0, tot_e = 0, max_v = 0, max_e = 0;
    for (uint32_t k = 0; k < 2 * nsub; ++k) {
      const bool v = k < nsub;
      const unsigned long long c = ctr[(size_t)(v ? k : kSub + (k - nsub)) * kCtrStride];
      pre[k] = acc;
      acc += min(c, (unsigned long long)(v ? sub_vf : sub_ee));
      if (v) {
        tot_v += c;
        max_v = max(max_v, c);
      } else {
        tot_e += c;
        max_e = max(max_e, c);
      }
    }
    pre[2 * nsub] = acc;
    if (blockIdx.x == 0) {
      a.sc->vf_list_n = tot_v;
      a.sc->ee_list_n = tot_e;
      a.sc->vf_list_max = max_v;
      a.sc->ee_list_max = max_e;
    }
  }
  __syncthreads();
  // Found contacts staged per wavefront (count in a register), written to the
  // raw list with one slot allocation per flush: the raw-contact counter is
  // one address for the whole device.
  constexpr uint32_t kStage = 128;
  __shared__ RawContact staged[TPB / 64][kStage];
  RawContact* const buf = staged[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63;
  uint32_t n = 0;
  auto flush = [&] {
    if (!n) return;
    wave_sync();  // the staged records are read by other lanes
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.sc->raw_contacts, (unsigned long long)n);
    base = __shfl(base, 0, 64);
    for (uint32_t k = lane; k < n; k += 64) {
      const unsigned long long slot = base + k;
      if (slot >= a.cap) continue;
      const RawContact& rc = buf[k];
      a.raw_ids[5 * slot] = rc.kind;
      for (int q = 0; q < 4; ++q) a.raw_ids[5 * slot + 1 + q] = rc.ids[q];
      a.raw_count[slot] = rc.count;
      a.raw_eps[slot] = rc.eps;
    }
    if (lane == 0 && base + n > a.cap) atomicOr(&a.sc->err_overflow, 1u);
    wave_sync();  // read before the buffer is staged again
    n = 0;
  };
  PairCounts cnt;
  const uint64_t total = pre[2 * nsub], step = (uint64_t)gridDim.x * blockDim.x;
  // Wavefront-uniform trip count: every lane reaches the staging steps.
  for (uint64_t j0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); j0 < total; j0 += step) {
    const uint64_t j = j0 + lane;
    RawContact rc;
    rc.valid = false;
    if (j < total) {
      uint32_t k = 0;  // the part holding list position j
      while (pre[k + 1] <= j) ++k;
      const uint64_t r = j - pre[k];
      if (k < nsub) {
        const uint2 p = vf[sub_vf * k + r];
        vf_narrow(a, p.x, p.y, cnt, rc);
      } else {
        const uint2 p = ee[sub_ee * (k - nsub) + r];
        ee_narrow(a, p.x, p.y, cnt, rc);
      }
    }
    const unsigned long long mask = __ballot(rc.valid);
    if (rc.valid) buf[n + (uint32_t)__popcll(mask & ((1ull << lane) - 1))] = rc;
    n += (uint32_t)__popcll(mask);
    if (n > kStage - 64) flush();
  }
  flush();
  add_counts(a, cnt);
}

// ------------------------------------------------------- LBVH broad phase
// The alternative to the hash grid (cv_gpu_set_broad_phase): one linear BVH
// per primitive kind (faces, edges) over the same f64 boxes.  Leaves are
// ordered by 63-bit Morton codes of the box centres in the cube
// [-scale, scale]^3 of the padded radius, the internal nodes form a Karras
// radix tree (a node's range and split from the longest common prefixes of
// the sorted codes, ties broken by leaf index), and node boxes are fitted
// bottom up (the second child to arrive unions both).  A query walks its
// tree with a stack and descends only into nodes with
// box_dist2(query box, node box) <= pr^2: node boxes are exact f64 min/max
// unions, so a node's distance never exceeds a leaf's below it and the walk
// keeps every pair the reference's predicate keeps — candidate sets stay
// bit-exact.  Candidates are counted per query in a first walk, placed by an
// exclusive scan and written by a second walk (no atomics, deterministic
// order).  Nodes: internal 0..N-2, leaves N-1..2N-2 (root 0 in both cases).
struct Bvh {
  const double* box;     // [6 (2N - 1)]
  const int32_t* child;  // [2 (N - 1)]
  const uint32_t* prim;  // [N] primitive of each leaf, in leaf order
  uint32_t N;
};
constexpr int kBvhStack = 64;

__device__ __forceinline__ unsigned long long spread21(unsigned long long v) {
  v &= 0x1FFFFFull;
  v = (v | v << 32) & 0x1F00000000FFFFull;
  v = (v | v << 16) & 0x1F0000FF0000FFull;
  v = (v | v << 8) & 0x100F00F00F00F00Full;
  v = (v | v << 4) & 0x10C30C30C30C30C3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}
__global__ void k_morton(const double* box, uint32_t N, double lo, double scale21, unsigned long long* key,
                         uint32_t* idx) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* b = box + 6 * (size_t)i;
  unsigned long long q[3];
  for (int k = 0; k < 3; ++k) {
    double c = ((0.5 * (b[k] + b[3 + k])) - lo) * scale21;
    q[k] = (unsigned long long)fmin(fmax(c, 0.0), 2097151.0);
  }
  key[i] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
  idx[i] = i;
}
__device__ __forceinline__ int bvh_delta(const unsigned long long* k, int N, int i, int j) {
  if (j < 0 || j >= N) return -1;
  const unsigned long long a = k[i], b = k[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll((long long)(a ^ b));
}
__global__ void k_karras(const unsigned long long* k, int N, int32_t* child, int32_t* parent, int32_t* range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) parent[0] = -1;
  if (i >= N - 1) return;
  const int d = bvh_delta(k, N, i, i + 1) - bvh_delta(k, N, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = bvh_delta(k, N, i, i - d);
  int lmax = 2;
  while (bvh_delta(k, N, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (bvh_delta(k, N, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = bvh_delta(k, N, i, j);
  int s = 0;
  for (int div = 2;; div <<= 1) {
    const int t = (l + div - 1) / div;
    if (bvh_delta(k, N, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int left = min(i, j) == gamma ? N - 1 + gamma : gamma;
  const int right = max(i, j) == gamma + 1 ? N - 1 + gamma + 1 : gamma + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  parent[left] = i;
  parent[right] = i;
  range[2 * i] = min(i, j);  // the node's leaves (sorted order)
  range[2 * i + 1] = max(i, j);
}
// Refit: node boxes as unions of their children, in two launches that need
// only workgroup-scope ordering (device-scope fences per tree level would
// write back the L2 each time).  k_refit_chunks: one block per kRefitChunk
// consecutive sorted leaves writes the leaf boxes and refits every internal
// node whose leaf range lies inside its chunk, bottom up: the second of the
// two children to finish (arrival count in flag[], workgroup-scope atomics)
// forms the union.  A node whose parent's range leaves the chunk is listed
// in `top`.  k_refit_top: one block climbs from the listed nodes through the
// nodes that span chunks, the same way.  flag[] (N - 1 zeros) and *n_top
// (zero) are cleared before the first launch.
constexpr int kRefitChunk = 1024;
__device__ __forceinline__ void union_children(const int32_t* child, int par, double* box) {
  const double* a = box + 6 * (size_t)child[2 * par];
  const double* b = box + 6 * (size_t)child[2 * par + 1];
  double* o = box + 6 * (size_t)par;
  for (int k = 0; k < 3; ++k) {
    o[k] = fmin(a[k], b[k]);
    o[3 + k] = fmax(a[3 + k], b[3 + k]);
  }
}
__device__ __forceinline__ bool second_arrival(uint32_t* flag) {
  return __hip_atomic_fetch_add(flag, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
}
__global__ __launch_bounds__(kRefitChunk) void k_refit_chunks(const double* pbox, const uint32_t* prim,
                                                              const int32_t* child, const int32_t* parent,
                                                              const int32_t* range, uint32_t N, uint32_t* flag,
                                                              double* box, uint32_t* top, uint32_t* n_top) {
  const uint32_t t = blockIdx.x * kRefitChunk + threadIdx.x;
  if (t >= N) return;  // no block barriers below
  const int lo = (int)(blockIdx.x * kRefitChunk), hi = (int)min(N, (blockIdx.x + 1) * kRefitChunk) - 1;
  int cur = (int)(N - 1 + t);
  const double* src = pbox + 6 * (size_t)prim[t];
  for (int k = 0; k < 6; ++k) box[6 * (size_t)cur + k] = src[k];
  for (uint32_t up = 0; up < N; ++up) {  // a path has at most N - 1 internal nodes
    const int par = parent[cur];
    if (par < 0) return;  // the root (a tree within one chunk)
    if (range[2 * par] < lo || range[2 * par + 1] > hi) {
      top[atomicAdd(n_top, 1u)] = (uint32_t)cur;
      return;
    }
    if (!second_arrival(&flag[par])) return;
    union_children(child, par, box);
    cur = par;
  }
}
__global__ __launch_bounds__(kRefitChunk) void k_refit_top(const int32_t* child, const int32_t* parent, uint32_t N,
                                                           uint32_t* flag, double* box, const uint32_t* top,
                                                           const uint32_t* n_top) {
  const uint32_t m = *n_top;
  for (uint32_t k = threadIdx.x; k < m; k += kRefitChunk) {
    int cur = (int)top[k];
    for (uint32_t up = 0; up < N; ++up) {
      const int par = parent[cur];
      if (par < 0) break;
      if (!second_arrival(&flag[par])) break;
      union_children(child, par, box);
      cur = par;
    }
  }
}
// Morton codes of points (the VF walk's query order).
__global__ void k_morton_points(const double* x, uint32_t N, double lo, double scale21, unsigned long long* key,
                                uint32_t* idx) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  unsigned long long q[3];
  for (int k = 0; k < 3; ++k) q[k] = (unsigned long long)fmin(fmax((x[3 * (size_t)i + k] - lo) * scale21, 0.0), 2097151.0);
  key[i] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
  idx[i] = i;
}

// One LBVH walk per query, queries in Morton order (`order`), so the lanes
// of a wavefront descend the same subtrees.  A popped internal node is known
// to be within reach; its children are tested before they are pushed, and a
// popped leaf goes through the pair filter.  KIND 0: points against the face
// tree (vertex-face); KIND 1: moving edges against the edge tree (edge-edge).
// Every wavefront runs the loop until all its lanes are done.
template <int KIND>
__global__ __launch_bounds__(TPB) void k_bvh_walk(DetectArgs a, Bvh t, const uint32_t* order, uint32_t nq,
                                                  uint64_t cap, uint2* region, unsigned long long* ctr) {
  __shared__ WaveStage<512> stage[TPB / 64];
  WaveStage<512>& ws = stage[threadIdx.x >> 6];
  if ((threadIdx.x & 63) == 0) ws.n = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  PairCounts cnt;
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t q = 0;
  ogm::Box qb{};
  bool moving = false, ex = false;
  uint32_t ia = 0, ib = 0;
  bool ia_ex = false, ib_ex = false;
  int stack[kBvhStack];
  int sp = 0;
  if (qi < nq && t.N) {
    q = order[qi];
    if (KIND == 0) {
      qb = ogm::box_of(ld3(a.x, q));
      moving = q < a.n;
      ex = a.ex_off[q] != a.ex_off[q + 1];
      stack[sp++] = 0;
    } else if (a.edge_moving[q]) {
      qb = ldbox(a.ebox, q);
      ia = a.edges[2 * q];
      ib = a.edges[2 * q + 1];
      ia_ex = a.ex_off[ia] != a.ex_off[ia + 1];
      ib_ex = a.ex_off[ib] != a.ex_off[ib + 1];
      stack[sp++] = 0;
    }
  }
  const int leaf0 = (int)t.N - 1;
  // A walk pops each of the tree's 2N - 1 nodes at most once; more pops (a
  // malformed tree) or a full stack end it with the fallback flag (16).
  uint32_t pops = 0;
  while (__ballot(sp > 0)) {
    bool take = false;
    uint32_t prim = 0;
    if (sp > 0) {
      const int nd = stack[--sp];
      if (++pops > 2 * t.N) {
        atomicOr(&a.sc->err_overflow, 16u);
        sp = 0;
      } else if (nd >= leaf0) {
        prim = t.prim[nd - leaf0];
        const ogm::Box lb = ldbox(t.box, (uint32_t)nd);
        take = KIND == 0 ? vf_filter(a, q, moving, ex, qb, prim, lb, cnt, [&] { return load_face(a, prim); })
                         : ee_filter(a, q, qb, ia, ib, ia_ex, ib_ex, prim, lb, cnt,
                                     [&] { return load_edge(a, prim); }, [] { return true; });
      } else {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int ch = t.child[2 * nd + c];
          ++cnt.tests;
          if (ogm::box_dist2(qb, ldbox(t.box, (uint32_t)ch)) > a.pr2) continue;
          if (sp == kBvhStack) {
            atomicOr(&a.sc->err_overflow, 16u);
            sp = 0;
            break;
          }
          stack[sp++] = ch;
        }
      }
    }
    wave_stage(ws, take, q, prim);
    if (ws.n > ws.kCap - 64) wave_flush(ws, list_ctr(ctr, KIND, 0), cap, region, a.sc);
  }
  wave_flush(ws, list_ctr(ctr, KIND, 0), cap, region, a.sc);
  add_counts(a, cnt);  // every lane reaches it
}
__global__ void k_init_dist(unsigned long long* bits, uint32_t n, double rq) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bits[i] = (unsigned long long)__double_as_longlong(rq);
}

__global__ void k_contact_keys(const uint32_t* raw, uint64_t m, unsigned long long* hi,
                               unsigned long long* lo, uint32_t* perm) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t* r = &raw[5 * i];
  hi[i] = ((unsigned long long)r[0] << 62) | ((unsigned long long)r[1] << 31) | (unsigned long long)r[2];
  lo[i] = ((unsigned long long)r[3] << 32) | (unsigned long long)r[4];
  perm[i] = (uint32_t)i;
}
__global__ void k_gather_u64(const unsigned long long* src, const uint32_t* perm, uint64_t m,
                             unsigned long long* dst) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[perm[i]];
}
// Flag first occurrences of each (kind, ids) key in sorted order.
__global__ void k_unique_flags(const unsigned long long* hi_sorted, const unsigned long long* lo,
                               const uint32_t* perm, uint64_t m, uint32_t* flag) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  flag[i] = (i == 0 || hi_sorted[i] != hi_sorted[i - 1] || lo[perm[i]] != lo[perm[i - 1]]) ? 1u : 0u;
}
__global__ void k_unique_write(const uint32_t* raw, const uint32_t* raw_count, const double* raw_eps,
                               const uint32_t* perm, const uint32_t* flag, const uint32_t* pos,
                               uint64_t m, uint32_t* cids, uint32_t* ccount, double* ceps) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= m || !flag[i]) return;
  uint32_t src = perm[i], dst = pos[i];
  for (int k = 0; k < 5; ++k) cids[5 * dst + k] = raw[5 * src + k];
  ccount[dst] = raw_count[src];
  ceps[dst] = raw_eps[src];
}
__global__ void k_store_u64(unsigned long long* dst, unsigned long long v) { *dst = v; }
__global__ void k_store_count(const uint32_t* src, unsigned long long* dst) { *dst = *src; }

// Active contacts, incidence counts and incidence keys (v << 32 | contact).
// Over the unique contacts (count in sc->contacts; the grid covers the raw count).
__global__ void k_contact_post(const double* x, const uint32_t* cids, const uint32_t* ccount,
                               const double* ceps, uint32_t n, double r, double k, uint32_t* vcount,
                               unsigned long long* inc_key, Scalars* sc) {
  uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (c >= sc->contacts) return;
  int kind = (int)cids[5 * c];
  D3 p[4];
  for (int q = 0; q < 4; ++q) p[q] = ld3(x, cids[5 * c + 1 + q]);
  if (ogm::contact_value(kind, p, ceps[c], r, k) > 0.0) atomicAdd(&sc->active, 1ull);
  for (uint32_t q = 0; q < ccount[c]; ++q) {
    uint32_t v = cids[5 * c + 1 + q];
    if (v < n) {
      atomicAdd(&vcount[v], 1u);
      unsigned long long s = atomicAdd(&sc->incidences, 1ull);
      inc_key[s] = ((unsigned long long)v << 32) | (unsigned long long)c;
    }
  }
}
// Per-vertex contact term records in sorted incidence order.
template <class F>
__global__ void k_contact_records(const unsigned long long* inc_sorted, uint64_t m, const uint32_t* cids,
                                  const uint32_t* ccount, const double* ceps, cvm::TermRecT<F>* rec) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t v = (uint32_t)(inc_sorted[i] >> 32), c = (uint32_t)(inc_sorted[i] & 0xFFFFFFFFull);
  cvm::TermRecT<F> r;
  int cnt = (int)ccount[c], slot = 0;
  for (int q = 0; q < 4; ++q) {
    r.ids[q] = cids[5 * c + 1 + q];
    if (q < cnt && r.ids[q] == v) slot = q;
  }
  r.head = cvm::rec_head(cvm::TERM_CONTACT, slot, cnt, (int)cids[5 * c]);
  r.f[0] = (F)ceps[c];
  for (int q = 1; q < 11; ++q) r.f[q] = F(0);
  rec[i] = r;
}
// Radii and minimum distance over moving vertices (single block, fixed order).
// Radii relaxation * min(d, rq) per moving vertex and the minimum distance
// (order-independent: an atomic minimum on the bit patterns of non-negative
// doubles; sc->min_distance starts at rq).
__global__ void k_radii(const unsigned long long* bits, uint32_t n, double relaxation, double* radii,
                        Scalars* sc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long m = ~0ull;
  if (i < n) {
    unsigned long long b = bits[i];
    radii[i] = relaxation * __longlong_as_double((long long)b);
    m = b;
  }
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t = __shfl_xor(m, o, 64);
    m = t < m ? t : m;
  }
  if ((threadIdx.x & 63) == 0 && m != ~0ull) atomicMin((unsigned long long*)&sc->min_distance, m);
}

inline unsigned blocks(uint64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// Exclusive scan of uint32 counts in tiles of 4096 (1024 threads x 4): one
// launch up to 4096 elements, three (tiles, scan of tile sums, add) up to
// 4096^2.  rocPRIM's scan queries device properties on every call, host time
// that the detection's chain of short kernels would wait for.
constexpr uint32_t kScanTile = 4096;
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t count,
                                                     uint32_t* sums) {
  __shared__ uint32_t wsum[16];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = base + k < count ? in[base + k] : 0u;
  const uint32_t s = v[0] + v[1] + v[2] + v[3];
  const uint32_t incl = wave_incl_scan(s);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    uint32_t x = lane < 16 ? wsum[lane] : 0u;
    x = wave_incl_scan(x);
    if (lane < 16) wsum[lane] = x;
  }
  __syncthreads();
  uint32_t e = incl - s + (w ? wsum[w - 1] : 0u);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < count) out[base + k] = e;
    e += v[k];
  }
  if (sums && threadIdx.x == 1023) sums[blockIdx.x] = wsum[15];
}
__global__ __launch_bounds__(1024) void k_scan_add(uint32_t* out, uint32_t count, const uint32_t* sums) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4, add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < count) out[base + k] += add;
}

template <class F>
void with_temp(cv_gpu& g, F&& f) {
  size_t bytes = 0;
  CV_HIP(f(nullptr, bytes));
  g.b.temp.reserve(bytes + 256);
  bytes = g.b.temp.n;
  CV_HIP(f((void*)g.b.temp.p, bytes));
}

void scan_u32(cv_gpu& g, const uint32_t* in, uint32_t* out, size_t count) {
  hipStream_t st = g.stream;
  if (count == 0) return;
  if (count <= kScanTile) {
    k_scan_tiles<<<1, 1024, 0, st>>>(in, out, (uint32_t)count, nullptr);
  } else if (count <= (size_t)kScanTile * kScanTile) {
    uint32_t tiles = (uint32_t)((count + kScanTile - 1) / kScanTile);
    g.b.scan_sums.reserve(tiles);
    k_scan_tiles<<<tiles, 1024, 0, st>>>(in, out, (uint32_t)count, g.b.scan_sums.p);
    k_scan_tiles<<<1, 1024, 0, st>>>(g.b.scan_sums.p, g.b.scan_sums.p, tiles, nullptr);
    k_scan_add<<<tiles, 1024, 0, st>>>(out, (uint32_t)count, g.b.scan_sums.p);
  } else {
    with_temp(g, [&](void* t, size_t& bytes) {
      return rocprim::exclusive_scan(t, bytes, in, out, 0u, count, rocprim::plus<uint32_t>(), st);
    });
  }
  CV_HIP(hipGetLastError());
}

void read_scalars(cv_gpu& g) {
  CV_HIP(hipMemcpyAsync(g.h_scal, g.d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, g.stream));
  CV_HIP(hipStreamSynchronize(g.stream));
}

// Register face and edge boxes (inflated by pr) in the hash grid, and the
// queries (points in their own cell, moving edges in the cells of their box):
// entries bucketed by cell-key hash, table t in buckets [tT, (t+1)T) (faces,
// edges, points, moving edges).  Buffers are sized from the previous
// detection's entry count (with room to grow), so no read-back is needed
// here; an overflow is flagged on the device and the caller detects again
// with more room.  Returns log2(T).
int build_grid(cv_gpu& g, const RegPrims& rp, double h) {
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  const uint32_t N = rp.F + 2 * rp.E + rp.P;
  b.cell_count.reserve(N + 1);
  b.cell_off.reserve(N + 1);
  const double inv_h = 1.0 / h;
  k_reg_count<<<blocks(N), TPB, 0, st>>>(rp, inv_h, b.cell_count.p);
  CV_HIP(hipGetLastError());
  CV_HIP(hipMemsetAsync(b.cell_count.p + N, 0, sizeof(uint32_t), st));
  scan_u32(g, b.cell_count.p, b.cell_off.p, (size_t)N + 1);
  if (b.grid_cap == 0) {
    // First detection of this context: read the exact count once.
    uint32_t M0 = 0;
    CV_HIP(hipMemcpyAsync(&M0, b.cell_off.p + N, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CV_HIP(hipStreamSynchronize(st));
    b.grid_last = M0;
    b.grid_cap = b.grid_last + b.grid_last / 2 + 4096;
  }
  uint64_t M = std::max<uint64_t>(b.grid_last, 1), cap = b.grid_cap;
  if (cap > (1u << 30)) fail(CV_ERR_CAPACITY, "OGC grid registration exceeds 2^30 cells");
  int log2t = 10;
  while (log2t < 28 && (1ull << log2t) < M / 4) ++log2t;
  size_t T4 = (size_t)4 << log2t;
  b.cell_key.reserve(cap);
  b.cell_val.reserve(cap);
  b.cell_rank.reserve(cap);
  b.cell_entry.reserve(cap);
  b.cell_owner.reserve(cap);
  b.bcount.reserve(T4 + 1);
  b.bstart.reserve(T4 + 1);
  CV_HIP(hipMemsetAsync(b.bcount.p, 0, (T4 + 1) * sizeof(uint32_t), st));
  k_reg_owner<<<blocks(N), TPB, 0, st>>>(b.cell_off.p, N, b.cell_owner.p, cap, g.d_scal);
  k_reg_fill<<<blocks(cap), TPB, 0, st>>>(rp, inv_h, b.cell_off.p, b.cell_owner.p, cap, log2t, b.cell_key.p,
                                          b.cell_val.p, b.cell_rank.p, b.bcount.p, g.d_scal);
  CV_HIP(hipGetLastError());
  scan_u32(g, b.bcount.p, b.bstart.p, T4 + 1);
  k_reg_scatter<<<blocks(cap), TPB, 0, st>>>(rp, b.cell_key.p, b.cell_val.p, b.cell_rank.p, b.cell_off.p, cap, log2t,
                                             b.bstart.p, b.cell_entry.p);
  CV_HIP(hipGetLastError());
  return log2t;
}

// Garment diagnostics (src/garment/diagnostics.rs:7-28, src/garment/template.rs:536-538),
// compiled without FMA contraction so each triangle's value rounds as the
// host formula does.  Positive axial extension of every triangle in %, as a
// key whose order is the value's (non-negative doubles order like their bits).
__global__ void k_ext_values(const double* x, const uint32_t* tri, const double* inv, const double* cs,
                             const double* area, uint32_t T, unsigned long long* key, double* w) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  uint32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
  double co = cs[2 * t], sn = cs[2 * t + 1];
  double ww = 0.0, uu = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double d0 = x[3 * b + k] - x[3 * a + k], d1 = x[3 * c + k] - x[3 * a + k];
    double f0 = d0 * inv[4 * t] + d1 * inv[4 * t + 1], f1 = d0 * inv[4 * t + 2] + d1 * inv[4 * t + 3];
    double wk = f0 * co + f1 * sn, uk = -f0 * sn + f1 * co;
    ww += wk * wk;
    uu += uk * uk;
  }
  double v = 100.0 * fmax(fmax(sqrt(ww), sqrt(uu)) - 1.0, 0.0);
  key[t] = (unsigned long long)__double_as_longlong(v);
  w[t] = area[t];
}
// The first sorted triangle whose cumulative area reaches 95 % of the total,
// and the largest value.
__global__ void k_ext_pick(const unsigned long long* key, const double* prefix, uint32_t T, Scalars* sc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T) return;
  double thr = 0.95 * prefix[T - 1];
  if (prefix[i] >= thr && (i == 0 || prefix[i - 1] < thr)) sc->diag_p95 = __longlong_as_double((long long)key[i]);
  if (i == T - 1) sc->diag_max = __longlong_as_double((long long)key[i]);
}
__global__ void k_seam_gap(const double* x, const uint32_t* off, const uint32_t* v, const double* w, uint32_t P,
                           Scalars* sc) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double g[3] = {0.0, 0.0, 0.0};
  for (uint32_t k = off[p]; k < off[p + 1]; ++k)
    for (int c = 0; c < 3; ++c) g[c] += x[3 * v[k] + c] * w[k];
  double d = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  atomicMax(&sc->diag_gap_bits, (unsigned long long)__double_as_longlong(d));
}

}  // namespace

void device_scan_u32(cv_gpu& g, const uint32_t* in, uint32_t* out, size_t count) { scan_u32(g, in, out, count); }
namespace {
__global__ void k_iota(uint32_t* a, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
}  // namespace
void device_color_order(cv_gpu& g, uint32_t n, uint32_t ncolors) {
  if (n == 0) return;
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  if (b.color_iota.n < n) {
    b.color_iota.reserve(n);
    k_iota<<<blocks((uint32_t)b.color_iota.n), TPB, 0, st>>>(b.color_iota.p, (uint32_t)b.color_iota.n);
    CV_HIP(hipGetLastError());
  }
  b.color_keys.reserve(n);
  b.order.reserve(n);
  int bits = 1;
  while (bits < 32 && (1u << bits) < ncolors) ++bits;
  with_temp(g, [&](void* t, size_t& bytes) {  // stable: vertex order within a colour
    return rocprim::radix_sort_pairs(t, bytes, b.color_dev.p, b.color_keys.p, b.color_iota.p, b.order.p, (size_t)n,
                                     0, bits, st);
  });
}

// The LBVH broad phase and the narrow pass over its candidates.  The trees
// are rebuilt at the first detection of a step (b.bvh_rebuild) and refitted
// at the refresh detections within it (same leaf order, new boxes).
void lbvh_tree(cv_gpu& g, BvhBufs& t, const double* pbox, uint32_t N, double scale, bool rebuild) {
  hipStream_t st = g.stream;
  if (N == 0) {
    t.N = 0;
    return;
  }
  if (rebuild || t.N != N) {
    t.N = N;
    t.key.reserve(N);
    t.key_sorted.reserve(N);
    t.idx.reserve(N);
    t.prim.reserve(N);
    t.child.reserve(2 * (size_t)std::max<uint32_t>(N - 1, 1));
    t.range.reserve(2 * (size_t)std::max<uint32_t>(N - 1, 1));
    t.parent.reserve(2 * (size_t)N);
    t.box.reserve(6 * (2 * (size_t)N));
    t.flag.reserve(std::max<uint32_t>(N - 1, 1) + 1);  // arrival counts, then the top-list count
    t.top.reserve(2 * (size_t)N);
    const double half = scale > 0 ? scale : 1.0;  // box centres lie in [-scale, scale]^3
    k_morton<<<blocks(N), TPB, 0, st>>>(pbox, N, -half, 2097152.0 / (2.0 * half), t.key.p, t.idx.p);
    CV_HIP(hipGetLastError());
    with_temp(g, [&](void* tmp, size_t& bytes) {
      return rocprim::radix_sort_pairs(tmp, bytes, t.key.p, t.key_sorted.p, t.idx.p, t.prim.p, (size_t)N, 0, 63,
                                       st);
    });
    k_karras<<<blocks(std::max<uint32_t>(N - 1, 1)), TPB, 0, st>>>(t.key_sorted.p, (int)N, t.child.p, t.parent.p,
                                                                  t.range.p);
    CV_HIP(hipGetLastError());
  }
  uint32_t* n_top = t.flag.p + std::max<uint32_t>(N - 1, 1);
  CV_HIP(hipMemsetAsync(t.flag.p, 0, (std::max<uint32_t>(N - 1, 1) + 1) * sizeof(uint32_t), st));
  const unsigned nb = (N + kRefitChunk - 1) / kRefitChunk;
  k_refit_chunks<<<nb, kRefitChunk, 0, st>>>(pbox, t.prim.p, t.child.p, t.parent.p, t.range.p, N, t.flag.p, t.box.p,
                                             t.top.p, n_top);
  CV_HIP(hipGetLastError());
  if (nb > 1) {
    k_refit_top<<<1, kRefitChunk, 0, st>>>(t.child.p, t.parent.p, N, t.flag.p, t.box.p, t.top.p, n_top);
    CV_HIP(hipGetLastError());
  }
}

void lbvh_broad_narrow(cv_gpu& g, const DetectArgs& a, double scale, uint2* vf, uint2* ee) {
  DevSolver& d = g.dev;
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  const bool rebuild = b.bvh_rebuild;
  b.bvh_rebuild = false;
  lbvh_tree(g, b.bvh_faces, b.fbox.p, d.F, scale, rebuild);
  lbvh_tree(g, b.bvh_edges, b.ebox.p, d.E, scale, rebuild);
  BvhBufs& pts = b.bvh_points;
  if (rebuild || pts.N != d.n_total) {
    // Points in Morton order of their positions at the step's first detection.
    const uint32_t N = d.n_total;
    pts.N = N;
    pts.key.reserve(N);
    pts.key_sorted.reserve(N);
    pts.idx.reserve(N);
    pts.prim.reserve(N);
    const double half = scale > 0 ? scale : 1.0;
    k_morton_points<<<blocks(N), TPB, 0, st>>>(a.x, N, -half, 2097152.0 / (2.0 * half), pts.key.p, pts.idx.p);
    CV_HIP(hipGetLastError());
    with_temp(g, [&](void* tmp, size_t& bytes) {
      return rocprim::radix_sort_pairs(tmp, bytes, pts.key.p, pts.key_sorted.p, pts.idx.p, pts.prim.p, (size_t)N,
                                       0, 63, st);
    });
  }
  Bvh tf{b.bvh_faces.box.p, b.bvh_faces.child.p, b.bvh_faces.prim.p, d.F};
  Bvh te{b.bvh_edges.box.p, b.bvh_edges.child.p, b.bvh_edges.prim.p, d.E};
  if (d.F)
    k_bvh_walk<0><<<blocks(d.n_total), TPB, 0, st>>>(a, tf, pts.prim.p, d.n_total, b.cand_cap_vf, vf, b.list_ctr.p);
  if (d.E)
    k_bvh_walk<1><<<blocks(d.E), TPB, 0, st>>>(a, te, b.bvh_edges.prim.p, d.E, b.cand_cap_ee, ee, b.list_ctr.p);
  CV_HIP(hipGetLastError());
  k_narrow_lists<<<8 * 256, TPB, 0, st>>>(a, vf, b.cand_cap_vf, ee, b.cand_cap_ee, b.list_ctr.p, 1);
  CV_HIP(hipGetLastError());
}

void gpu_detect(cv_gpu& g, const cv_solver& s, double rq, const cv_ogc_settings& ogc, size_t cand_cap) {
  DevSolver& d = g.dev;
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  // Padded radius (src/ogc/world.rs:352-358): scale = max(rq, max |x|_inf).
  // The step prologue stored max |x| of the moving part in sc->scale.
  double scale = std::max(std::max(rq, g.h_scal->scale), d.fixed_abs_max);
  double pr = rq + 64.0 * 2.220446049250313e-16 * scale;
  double pr2 = pr * pr;
  // Grid registration inflates by slightly more than pr so rounding of
  // box +- pr can never drop a pair the exact predicate accepts.
  double pr_reg = pr * (1.0 + 1e-9) + 1e-12;
  double h = std::max(d.cell, 2.0 * pr_reg);
  // Reset per-detection counters.
  Scalars& hs = *g.h_scal;
  hs.vf_count = hs.ee_count = hs.vf_cand = hs.ee_cand = hs.grid_tests = 0;
  hs.raw_contacts = hs.contacts = hs.active = hs.incidences = 0;
  hs.cand_export_vf = hs.cand_export_ee = 0;
  hs.vf_list_n = hs.ee_list_n = hs.vf_list_max = hs.ee_list_max = 0;
  hs.err_degenerate = hs.err_overflow = 0;
  hs.rq = rq;
  hs.padded_radius2 = pr2;
  CV_HIP(hipMemcpyAsync(g.d_scal, g.h_scal, sizeof(Scalars), hipMemcpyHostToDevice, st));
  b.fbox.reserve(6 * (size_t)d.F);
  b.ebox.reserve(6 * (size_t)d.E);
  k_boxes<<<blocks(d.F + d.E), TPB, 0, st>>>(b.x64.p, d.faces.p, d.face_moving.p, d.F, d.edges.p, d.E,
                                             b.fbox.p, b.ebox.p, g.d_scal);
  CV_HIP(hipGetLastError());
  b.dist_bits.reserve(d.n_total);
  k_init_dist<<<blocks(d.n_total), TPB, 0, st>>>(b.dist_bits.p, d.n_total, rq);
  if (cand_cap) {
    b.cand_vf.reserve(2 * cand_cap);
    b.cand_ee.reserve(2 * cand_cap);
  }
  DetectArgs a{};
  a.x = b.x64.p;
  a.n = d.n;
  a.n_total = d.n_total;
  a.F = d.F;
  a.E = d.E;
  a.faces = d.faces.p;
  a.edges = d.edges.p;
  a.tri_edges = d.faces.p + 3 * (size_t)d.F;  // packed after faces (see ensure_device_solver)
  a.opp = d.opposites.p;
  a.nb_off = d.nb_off.p;
  a.nb = d.nb.p;
  a.ex_off = d.ex_off.p;
  a.ex = d.ex.p;
  a.face_moving = d.face_moving.p;
  a.edge_moving = d.edge_moving.p;
  a.rest_len2 = d.rest_len2.p;
  a.fbox = b.fbox.p;
  a.ebox = b.ebox.p;
  a.rq = rq;
  a.pr = pr_reg;
  a.pr2 = pr2;
  a.inv_h = 1.0 / h;
  a.dist_bits = b.dist_bits.p;
  a.raw_ids = b.raw_ids.p;
  a.raw_count = b.raw_count.p;
  a.raw_eps = b.raw_eps.p;
  a.cap = b.contact_cap;
  a.cand_vf = cand_cap ? b.cand_vf.p : nullptr;
  a.cand_ee = cand_cap ? b.cand_ee.p : nullptr;
  a.cand_cap = cand_cap;
  a.sc = g.d_scal;
  // Candidate regions of the broad pass: VF [0, cap_vf), EE [cap_vf,
  // cap_vf + cap_ee), sized from the last detection's counts (grown and
  // redone on overflow).
  if (b.cand_cap_vf == 0) b.cand_cap_vf = 8ull * d.n_total + (1u << 20);
  if (b.cand_cap_ee == 0) b.cand_cap_ee = 8ull * d.E + (1u << 20);
  b.pair_list.reserve(b.cand_cap_vf + b.cand_cap_ee);
  b.list_ctr.reserve(kCtrCount * kCtrStride);
  CV_HIP(hipMemsetAsync(b.list_ctr.p, 0, kCtrCount * kCtrStride * sizeof(unsigned long long), st));
  uint2* vf = b.pair_list.p;
  uint2* ee = b.pair_list.p + b.cand_cap_vf;
  if (g.broad_phase == 1 && !g.b.bvh_fallback) {
    lbvh_broad_narrow(g, a, scale, vf, ee);
  } else {
    // Faces and edges registered with the padded inflation; points look up
    // their own cell, edges the cells of their own box.
    const RegPrims rp{b.fbox.p, b.ebox.p, b.x64.p, d.edge_moving.p, d.F, d.E, d.n_total, pr_reg};
    a.log2t = build_grid(g, rp, h);
    a.ent = b.cell_entry.p;
    a.bstart = b.bstart.p;
    k_pairs<<<8 * 256, TPB, 0, st>>>(a, b.cand_cap_vf, vf, b.cand_cap_ee, ee, b.list_ctr.p);
    CV_HIP(hipGetLastError());
    k_narrow_lists<<<8 * 256, TPB, 0, st>>>(a, vf, b.cand_cap_vf, ee, b.cand_cap_ee, b.list_ctr.p, kSub);
    CV_HIP(hipGetLastError());
  }
  read_scalars(g);
  if (g.h_scal->err_overflow & 16u) {
    // An LBVH walk ran out of stack (a tree deeper than kBvhStack along one
    // path) or out of its pop bound: this detection is redone with the grid.
    g.b.bvh_fallback = true;
    g.b.bvh_fallbacks++;
    gpu_detect(g, s, rq, ogc, cand_cap);
    g.b.bvh_fallback = false;
    return;
  }
  if (g.broad_phase != 1 || g.b.bvh_fallback) b.grid_last = g.h_scal->grid_entries;
  if (g.h_scal->err_overflow & 40u) {
    // More grid registrations (8) than room, or a candidate region too small
    // (32): grow to the counts this detection found and detect again.
    const Scalars& hs = *g.h_scal;
    if (hs.err_overflow & 8u) b.grid_cap = b.grid_last + b.grid_last / 2 + 4096;
    const uint64_t ns = g.broad_phase == 1 && !g.b.bvh_fallback ? 1 : kSub;  // parts per kind
    if (hs.vf_list_max * ns > b.cand_cap_vf) b.cand_cap_vf = ns * (hs.vf_list_max + hs.vf_list_max / 4 + 4096);
    if (hs.ee_list_max * ns > b.cand_cap_ee) b.cand_cap_ee = ns * (hs.ee_list_max + hs.ee_list_max / 4 + 4096);
    gpu_detect(g, s, rq, ogc, cand_cap);
    return;
  }
  if (g.h_scal->err_overflow)
    fail(CV_ERR_CAPACITY, "OGC contact buffer or grid capacity exceeded");
  uint64_t m = g.h_scal->raw_contacts;
  // Canonical order and dedup (src/ogc/world.rs:377-381).
  b.key_hi.reserve(m);
  b.key_lo.reserve(m);
  b.key_tmp.reserve(m);
  b.key_hi_sorted.reserve(m);
  b.perm.reserve(m);
  b.perm_tmp.reserve(m);
  b.uniq_flag.reserve(m + 1);
  b.uniq_pos.reserve(m + 1);
  if (m) {
    k_contact_keys<<<blocks(m), TPB, 0, st>>>(b.raw_ids.p, m, b.key_hi.p, b.key_lo.p, b.perm.p);
    with_temp(g, [&](void* t, size_t& bytes) {  // by low key
      return rocprim::radix_sort_pairs(t, bytes, b.key_lo.p, b.key_tmp.p, b.perm.p, b.perm_tmp.p,
                                       (size_t)m, 0, 64, st);
    });
    k_gather_u64<<<blocks(m), TPB, 0, st>>>(b.key_hi.p, b.perm_tmp.p, m, b.key_tmp.p);
    with_temp(g, [&](void* t, size_t& bytes) {  // stable by high key
      return rocprim::radix_sort_pairs(t, bytes, b.key_tmp.p, b.key_hi_sorted.p, b.perm_tmp.p, b.perm.p,
                                       (size_t)m, 0, 64, st);
    });
    // key_hi_sorted holds the high keys in final order, perm the permutation.
    k_unique_flags<<<blocks(m), TPB, 0, st>>>(b.key_hi_sorted.p, b.key_lo.p, b.perm.p, m, b.uniq_flag.p);
    CV_HIP(hipMemsetAsync(b.uniq_flag.p + m, 0, sizeof(uint32_t), st));
    scan_u32(g, b.uniq_flag.p, b.uniq_pos.p, (size_t)m + 1);
    k_unique_write<<<blocks(m), TPB, 0, st>>>(b.raw_ids.p, b.raw_count.p, b.raw_eps.p, b.perm.p,
                                              b.uniq_flag.p, b.uniq_pos.p, m, b.c_ids.p, b.c_count.p,
                                              b.c_eps.p);
    // The unique count lives in the device scalars like every other counter
    // (no read-back here; the next scalar read-back carries it).
    k_store_count<<<1, 1, 0, st>>>(b.uniq_pos.p + m, &g.d_scal->contacts);
    // The contact list for the host colouring, copied with room for all m
    // raw contacts so the copy needs no count first.
    g.h_cont_valid = g.host_contacts;
    if (g.host_contacts && g.h_cont_cap < m) {
      if (g.h_cont) CV_HIP(hipHostFree(g.h_cont));
      g.h_cont = nullptr;
      size_t cap = m + m / 2 + 1024;
      CV_HIP(hipHostMalloc((void**)&g.h_cont, 6 * cap * sizeof(uint32_t)));
      g.h_cont_cap = cap;
    }
    if (g.host_contacts) {
      CV_HIP(hipMemcpyAsync(g.h_cont, b.c_ids.p, 5 * m * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      CV_HIP(hipMemcpyAsync(g.h_cont + 5 * g.h_cont_cap, b.c_count.p, m * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, st));
    }
  } else {
    g.h_cont_valid = true;
    k_store_u64<<<1, 1, 0, st>>>(&g.d_scal->contacts, 0ull);
  }
  g.detected = true;
  CV_HIP(hipGetLastError());
  // Incidence, active count, per-vertex contact records.
  b.vcount.reserve(d.n + 1);
  b.coff.reserve(d.n + 1);
  CV_HIP(hipMemsetAsync(b.vcount.p, 0, sizeof(uint32_t) * (d.n + 1), st));
  if (m) {
    b.inc_key.reserve(4 * m);
    b.inc_key_sorted.reserve(4 * m);
    k_contact_post<<<blocks(m), TPB, 0, st>>>(b.x64.p, b.c_ids.p, b.c_count.p, b.c_eps.p, d.n, ogc.radius,
                                              ogc.stiffness, b.vcount.p, b.inc_key.p, g.d_scal);
  }
  scan_u32(g, b.vcount.p, b.coff.p, (size_t)d.n + 1);
  b.radii64.reserve(d.n);
  k_store_u64<<<1, 1, 0, st>>>((unsigned long long*)&g.d_scal->min_distance,
                               (unsigned long long)__builtin_bit_cast(uint64_t, rq));
  k_radii<<<blocks(d.n), TPB, 0, st>>>(b.dist_bits.p, d.n, ogc.relaxation, b.radii64.p, g.d_scal);
  read_scalars(g);
  uint64_t ninc = g.h_scal->incidences;
  b.crec.reserve(std::max<uint64_t>(ninc, 1));
  b.crec_d.reserve(std::max<uint64_t>(ninc, 1));
  if (ninc) {
    with_temp(g, [&](void* t, size_t& bytes) {
      return rocprim::radix_sort_keys(t, bytes, b.inc_key.p, b.inc_key_sorted.p, (size_t)ninc, 0, 64, st);
    });
    if (g.f64)
      k_contact_records<double><<<blocks(ninc), TPB, 0, st>>>(b.inc_key_sorted.p, ninc, b.c_ids.p,
                                                              b.c_count.p, b.c_eps.p, b.crec_d.p);
    else
      k_contact_records<float><<<blocks(ninc), TPB, 0, st>>>(b.inc_key_sorted.p, ninc, b.c_ids.p,
                                                             b.c_count.p, b.c_eps.p, b.crec.p);
    CV_HIP(hipGetLastError());
  }
  if (g.h_scal->err_degenerate) fail(CV_ERR_DEGENERATE_FACE, "OGC degenerate face; step rejected");
}

void gpu_diagnostics(cv_gpu& g, const double* dx) {
  DevSolver& d = g.dev;
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  k_store_u64<<<1, 1, 0, st>>>(&g.d_scal->diag_gap_bits, 0ull);
  k_store_u64<<<1, 1, 0, st>>>((unsigned long long*)&g.d_scal->diag_p95, 0ull);
  k_store_u64<<<1, 1, 0, st>>>((unsigned long long*)&g.d_scal->diag_max, 0ull);
  if (d.P) k_seam_gap<<<blocks(d.P), TPB, 0, st>>>(dx, d.dg_pr_off.p, d.dg_pr_v.p, d.dg_pr_w.p, d.P, g.d_scal);
  uint32_t T = d.T;
  if (T) {
    b.ext_key.reserve(T);
    b.ext_key_sorted.reserve(T);
    b.ext_area.reserve(T);
    b.ext_area_sorted.reserve(T);
    b.ext_prefix.reserve(T);
    k_ext_values<<<blocks(T), TPB, 0, st>>>(dx, d.dg_tri.p, d.dg_inv.p, d.dg_cs.p, d.dg_area.p, T, b.ext_key.p,
                                            b.ext_area.p);
    with_temp(g, [&](void* t, size_t& bytes) {
      return rocprim::radix_sort_pairs(t, bytes, b.ext_key.p, b.ext_key_sorted.p, b.ext_area.p,
                                       b.ext_area_sorted.p, (size_t)T, 0, 64, st);
    });
    with_temp(g, [&](void* t, size_t& bytes) {
      return rocprim::inclusive_scan(t, bytes, b.ext_area_sorted.p, b.ext_prefix.p, (size_t)T,
                                     rocprim::plus<double>(), st);
    });
    k_ext_pick<<<blocks(T), TPB, 0, st>>>(b.ext_key_sorted.p, b.ext_prefix.p, T, g.d_scal);
  }
  CV_HIP(hipGetLastError());
}
