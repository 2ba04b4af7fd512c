// vbd_device.hip — the device VBD step (GpuVbd::step, src/gpu_vbd.rs:125-472,
// with the semantics of VbdSolver::step, src/vbd.rs:250-422).
//
// Per-colour Gauss-Seidel sweep: a group of L lanes serves one vertex.  Each
// lane evaluates the exact energy/gradient/Hessian of its incident terms
// (records stored contiguously per vertex, so a group reads one coalesced
// run), the group reduces with butterfly shuffles (every lane ends with the
// identical sum), every lane forms the same Gershgorin-shifted Newton step,
// and the ball-truncated backtracking search re-evaluates only energies.
//
// Positions, targets and ball centres are stored in f64, relative to the
// garment centroid.  The sweep is templated on its arithmetic: f64 (default)
// reproduces the f64 reference step to round-off; f32 solves each vertex in
// its own frame (coordinates relative to the vertex's position, converted
// from f64 once per load), so f32 rounds lengths of the one-ring's size
// rather than of the scene's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include <hip/hip_ext.h>

#include "gpu.hpp"

using cvm::V3;

namespace {

constexpr int TPB = 256;
inline unsigned blocks(uint64_t n, int per = TPB) { return (unsigned)((n + per - 1) / per); }

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <class R>
struct GrabArgs {
  int active;
  uint32_t v[3];
  R w[3];
  double target[3];  // relative to the origin
  R k;
};

template <class R>
struct SweepArgs {
  const uint4* vinfo;  // per colour-ordered slot: vertex, s0, ns, c0 (k_vertex_info)
  const uint32_t* vnc;  // and its contact-record count
  double4* pos;  // xyz relative to the origin, w: pinned flag
  const double4* tgt;  // target, inertia
  const double4* ball;  // centre, radius
  const uint32_t* term_off;
  const cvm::TermRecT<R>* terms;
  const uint32_t* coff;
  const cvm::TermRecT<R>* crec;
  double* gsq;  // per-vertex |g|^2 of the sweep (f64 in both precisions)
  unsigned long long* clipped;
  const unsigned int* stop;  // set when a refresh is pending: the launch is a no-op
  // 32-lane sweep only (else null): exclusive scan over the slots of the
  // "more than 32 records" flag, and those slots in order (k_heavy_*).
  const uint32_t* hpos;
  const uint32_t* hlist;
  R cr, ck;
  GrabArgs<R> grab;
};

// Origin of a vertex solve's coordinates: the stored origin itself in f64
// (no conversion: the f64 sweep works on the stored values), the solved
// vertex's own position in f32.
struct Frame {
  double x, y, z;
};
template <class R>
__device__ __forceinline__ V3<R> rel(double4 q, const Frame& f) {
  if constexpr (sizeof(R) == 8)
    return V3<R>{q.x, q.y, q.z};
  else
    return V3<R>{(R)(q.x - f.x), (R)(q.y - f.y), (R)(q.z - f.z)};
}
template <class R>
__device__ __forceinline__ V3<R> ldp(const double4* pos, uint32_t i, const Frame& f) {
  return rel<R>(pos[i], f);
}
template <class R>
__device__ __forceinline__ Frame frame_of(double4 p) {
  if constexpr (sizeof(R) == 8)
    return Frame{0.0, 0.0, 0.0};
  else
    return Frame{p.x, p.y, p.z};
}
// A solved position in the frame back to stored f64 coordinates.
template <class R>
__device__ __forceinline__ double4 stored(V3<R> p, const Frame& f) {
  if constexpr (sizeof(R) == 8)
    return make_double4(p.x, p.y, p.z, 0.0);
  else
    return make_double4(f.x + (double)p.x, f.y + (double)p.y, f.z + (double)p.z, 0.0);
}

// Value of v in another lane of the same 16-lane row, by a DPP move (no LDS
// round trip); CTRL is the DPP control (quad permute, row mirrors).  All
// lanes of the row are active wherever a group of >= 8 lanes reduces.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, CTRL, 0xF, 0xF, false));
}
// Sum over a group of L >= 8 aligned lanes, the same bits in every lane.
// Levels pair lanes whose partial sums cover the same terms: distance 1 and 2
// in a quad (quad permutes [1,0,3,2] and [2,3,0,1]), the mirrored quads of a
// half-row, the mirrored half-rows of a row, then lane shuffles across rows.
template <int L, class R>
__device__ __forceinline__ R group_sum(R v) {
  static_assert(L >= 8 && L <= 64 && (L & (L - 1)) == 0, "group of 8, 16, 32 or 64 lanes");
  v += dpp_move<0xB1>(v);  // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);  // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  if (L >= 16) v += dpp_move<0x140>(v);  // row_mirror
#pragma unroll
  for (int m = 16; m < L; m <<= 1) v += __shfl_xor(v, m, L);
  return v;
}

// Grab term of vertex v (src/vbd.rs:237-246): d = sum w_j x_j - target.
template <class R>
__device__ __forceinline__ bool grab_d(const GrabArgs<R>& g, uint32_t v, V3<R> xv, const double4* pos,
                                       const Frame& f, V3<R>& d, R& wv) {
  if (!g.active) return false;
  int s = -1;
  for (int k = 0; k < 3; ++k)
    if (g.v[k] == v) s = k;
  if (s < 0) return false;
  d = xv - rel<R>(make_double4(g.target[0], g.target[1], g.target[2], 0.0), f);
  for (int k = 0; k < 3; ++k)
    if (k != s) d = d + (ldp<R>(pos, g.v[k], f) - xv) * g.w[k];
  wv = g.w[s];
  return true;
}

// The first term record of each lane and the positions of its four vertices
// are loaded once per vertex solve and kept in registers for the derivative
// and every line-search energy (the neighbours do not move during the solve;
// the vertex's own slot is replaced by the candidate).  Unused ids of a
// record are 0, a valid vertex.  Records past the first L are re-read.
template <class R>
struct LaneTerm {
  cvm::TermRecT<R> rec;
  V3<R> q[4];
  bool live;
};
template <class R>
__device__ __forceinline__ LaneTerm<R> load_lane_term(const SweepArgs<R>& a, int lane, uint32_t s0, uint32_t ns,
                                                      uint32_t c0, uint32_t nt, const Frame& f) {
  LaneTerm<R> t;
  t.live = (uint32_t)lane < nt;
  if (t.live) {
    t.rec = (uint32_t)lane < ns ? a.terms[s0 + lane] : a.crec[c0 + (lane - ns)];
#pragma unroll
    for (int k = 0; k < 4; ++k) t.q[k] = ldp<R>(a.pos, t.rec.ids[k], f);
  }
  return t;
}

template <int L, class R>
__device__ __forceinline__ R local_energy(const SweepArgs<R>& a, const LaneTerm<R>& lt, uint32_t v, V3<R> xv, int lane,
                          uint32_t s0, uint32_t ns, uint32_t c0, uint32_t nt, const Frame& f) {
  auto P = [&](int, uint32_t i) { return ldp<R>(a.pos, i, f); };
  auto Q = [&](int k, uint32_t) { return lt.q[k]; };
  R e = lt.live ? cvm::term_energy<R>(lt.rec, xv, a.cr, a.ck, Q) : R(0);
  for (uint32_t i = lane + L; i < nt; i += L) {
    const cvm::TermRecT<R>& t = i < ns ? a.terms[s0 + i] : a.crec[c0 + (i - ns)];
    e += cvm::term_energy<R>(t, xv, a.cr, a.ck, P);
  }
  if (lane == 0) {
    V3<R> d;
    R wv;
    if (grab_d(a.grab, v, xv, a.pos, f, d, wv)) e += R(0.5) * a.grab.k * cvm::dot(d, d);
  }
  return group_sum<L>(e);
}

// f32 line search: the change of the local energy (without the inertia term)
// from x to the candidate xc, summed over the group: per lane, its records'
// energies at xc less ex (their energies at x, from the derivative pass),
// with stitches and the grab in the exact difference form (quad_change).
template <int L, class R>
__device__ __forceinline__ R local_change(const SweepArgs<R>& a, const LaneTerm<R>& lt, uint32_t v, V3<R> x, V3<R> xc,
                                          int lane, uint32_t s0, uint32_t ns, uint32_t c0, uint32_t nt,
                                          const Frame& f, R ex) {
  auto P = [&](int, uint32_t i) { return ldp<R>(a.pos, i, f); };
  auto Q = [&](int k, uint32_t) { return lt.q[k]; };
  R e = -ex;
  if (lt.live) e += cvm::term_change_part<R>(lt.rec, x, xc, a.cr, a.ck, Q);
  for (uint32_t i = lane + L; i < nt; i += L) {
    const cvm::TermRecT<R>& t = i < ns ? a.terms[s0 + i] : a.crec[c0 + (i - ns)];
    e += cvm::term_change_part<R>(t, x, xc, a.cr, a.ck, P);
  }
  if (lane == 0) {
    V3<R> d;
    R wv;
    if (grab_d(a.grab, v, x, a.pos, f, d, wv)) e += cvm::quad_change<R>(d, wv, a.grab.k, xc - x);
  }
  return group_sum<L>(e);
}

// Derivatives and the local energy at x, the energy summed from the terms'
// own values in the per-lane order and reduction local_energy uses, so it
// serves as the line search's reference energy without a second pass over
// the terms.  It can differ from local_energy at x in the last bits (the two
// routines may contract differently); the accept slack (accept_slack) covers that.
// E = false leaves acc.e undefined.  With X, *ex receives the lane's own
// (unreduced) energy at x of its records other than stitches: what the f32
// line search subtracts from their energies at a candidate.
template <int L, class R, bool E = true, bool X = false>
__device__ __forceinline__ cvm::Deriv<R> local_deriv(const SweepArgs<R>& a, const LaneTerm<R>& lt, uint32_t v, V3<R> x,
                                     int lane, uint32_t s0, uint32_t ns, uint32_t c0, uint32_t nt,
                                     const Frame& f, R* ex = nullptr) {
  auto P = [&](int, uint32_t i) { return ldp<R>(a.pos, i, f); };
  auto Q = [&](int k, uint32_t) { return lt.q[k]; };
  cvm::Deriv<R> acc = cvm::deriv_zero<R>();
  R e_other = R(0);
  if (lt.live) {
    acc = cvm::term_deriv<R>(lt.rec, x, a.cr, a.ck, Q);
    if (X && cvm::rec_kind(lt.rec.head) != cvm::TERM_STITCH) e_other = acc.e;
  }
  for (uint32_t i = lane + L; i < nt; i += L) {
    const cvm::TermRecT<R>& t = i < ns ? a.terms[s0 + i] : a.crec[c0 + (i - ns)];
    const cvm::Deriv<R> dt = cvm::term_deriv<R>(t, x, a.cr, a.ck, P);
    if (X && cvm::rec_kind(t.head) != cvm::TERM_STITCH) e_other += dt.e;
    cvm::deriv_add(acc, dt);
  }
  if (X) *ex = e_other;
  if (lane == 0) {
    V3<R> d;
    R wv;
    if (grab_d(a.grab, v, x, a.pos, f, d, wv)) cvm::deriv_add(acc, cvm::stitch_d<R>(d, wv, a.grab.k));
  }
  if (E && !X) acc.e = group_sum<L>(acc.e);
  acc.g.x = group_sum<L>(acc.g.x);
  acc.g.y = group_sum<L>(acc.g.y);
  acc.g.z = group_sum<L>(acc.g.z);
#pragma unroll
  for (int k = 0; k < 6; ++k) acc.h.a[k] = group_sum<L>(acc.h.a[k]);
  return acc;
}

// The local solve of vertex vi.x by its group of L lanes (src/vbd.rs:337-395);
// vi = (vertex, first term record, term count, first contact record), nc the
// contact-record count.
template <int L, class R>
__device__ __forceinline__ void solve_vertex(const SweepArgs<R>& a, uint4 vi, uint32_t nc, int lane) {
  const uint32_t v = vi.x, s0 = vi.y, ns = vi.z, c0 = vi.w, nt = ns + nc;
  // The vertex's data and the lane's first term record are loaded at once.
  const double4 p4 = a.pos[v];
  const double4 tg = a.tgt[v];
  const double4 b4 = a.ball[v];
  const Frame f = frame_of<R>(p4);
  const LaneTerm<R> lt = load_lane_term(a, lane, s0, ns, c0, nt, f);
  if (p4.w > 0.5) {  // pinned: no update, zero residual contribution
    if (lane == 0) a.gsq[v] = 0.0;
    return;
  }
  V3<R> x = rel<R>(p4, f);
  constexpr bool F32 = sizeof(R) == 4;
  R ex = R(0);
  cvm::Deriv<R> acc = local_deriv<L, R, true, F32>(a, lt, v, x, lane, s0, ns, c0, nt, f, &ex);
  V3<R> target = rel<R>(tg, f);
  R inertia = (R)tg.w;
  // f64: the reference's test E <= E_old + 1e-14 on energies from one
  // evaluation routine (a candidate equal to x reproduces `old` exactly).
  // f32: the same test on the change E - E_old, accumulated as changes (see
  // local_change; the inertia term as m/2 dx.(c + x - 2 t)), with a slack of
  // four f32 ulps of the non-quadratic energies it differences.
  V3<R> dx = x - target;
  R old = R(0), slack = R(0);
  if constexpr (F32) {
    slack = R(1e-14) + R(4.0 * 1.1920928955078125e-07) * group_sum<L>(fabs(ex));
  } else {
    old = acc.e + R(0.5) * inertia * cvm::dot(dx, dx);
    slack = cvm::accept_slack<R>(old);
  }
  V3<R> delta = cvm::newton_direction<R>(acc, x, target, inertia);
  R gg = cvm::dot(acc.g, acc.g);
  V3<R> center = rel<R>(b4, f);
  const R radius = (R)b4.w;
  bool truncated = false;
  V3<R> accepted = x;
  R scale = R(1);
  for (int attempt = 0; attempt < 10; ++attempt) {
    V3<R> cand = cvm::truncate_ball<R>(x + delta * scale, center, radius, truncated);
    V3<R> dc = cand - target;
    R e;
    if constexpr (F32)
      e = local_change<L>(a, lt, v, x, cand, lane, s0, ns, c0, nt, f, ex) +
          R(0.5) * inertia * cvm::dot(cand - x, dc + dx);
    else
      e = local_energy<L>(a, lt, v, cand, lane, s0, ns, c0, nt, f) + R(0.5) * inertia * cvm::dot(dc, dc);
    if (cvm::finite3(cand) && isfinite(e) && e <= old + slack) {
      accepted = cand;
      break;
    }
    scale *= R(0.5);
  }
  if (lane == 0) {
    a.pos[v] = stored<R>(accepted, f);
    a.gsq[v] = (double)gg;
    if (truncated) atomicAdd(a.clipped, 1ull);
  }
}

// End of iteration `it` (it = -1: after the truncated guess), on the device
// so the host need not read the clipped count back every iteration: counts
// the iteration, accumulates truncations (c, the iteration's clipped count)
// and, when the clipped count reaches the refresh threshold
// (src/vbd.rs:326-336), raises `stop` so the launches already queued behind
// it are no-ops until the host has re-detected; with the residual tolerance,
// a converged sweep ends the step.
__device__ __forceinline__ void iter_end(Scalars* sc, unsigned long long c, int it, double threshold, double tol,
                                         double nfree3) {
  if (sc->stop) return;
  if (it >= 0) {
    sc->iters_run += 1;
    sc->truncated_total += c;
  }
  sc->clipped = 0;
  // Residual-tolerance extension: converged after this sweep ends the step.
  if (it >= 0 && tol > 0.0 && sqrt(sc->true_residual_sum / nfree3) <= tol) {
    sc->stop = 2;
    sc->stop_it = (unsigned int)(it + 1);
    return;
  }
  if (c > 0 && (double)c >= threshold) {
    sc->stop = 1;
    sc->stop_it = (unsigned int)(it + 1);  // the next iteration to run after the refresh
  }
}
__global__ void k_iter_end(Scalars* sc, int it, double threshold, double tol, double nfree3) {
  if (it < 0) sc->sweep_blocks = 0;
  iter_end(sc, sc->clipped, it, threshold, tol, nfree3);
}
// The iteration's end folded into its last sweep launch: the last block to
// finish (device-scope counter) runs iter_end.  Not on iterations after which
// the residual tolerance is checked, which need a residual pass between the
// sweep and the decision.
struct IterTail {
  Scalars* sc;  // null: no tail
  int it;
  double threshold;
};
__device__ __forceinline__ void iter_tail(const IterTail& t) {
  __syncthreads();  // the block's clipped-count atomics are done
  if (threadIdx.x != 0) return;
  __threadfence();
  if (atomicAdd(&t.sc->sweep_blocks, 1u) != gridDim.x - 1) return;
  __threadfence();
  t.sc->sweep_blocks = 0;
  // Read through to the device-coherent value (this CU's L1 may hold a stale
  // line of the scalars).
  unsigned long long c = __hip_atomic_load(&t.sc->clipped, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  iter_end(t.sc, c, t.it, t.threshold, 0.0, 1.0);
}

// A vertex with more records than a 32-lane group holds would have its
// lanes re-read the extra records on every pass (the records past the first
// L are not kept in registers), which makes it several times slower than the
// rest and, although such vertices are rare (~0.1 % on a garment in
// contact), the slowest wavefront of most launches.  The 32-lane sweep
// therefore leaves them to kHeavyBlocks extra blocks of the same launch,
// where each takes a whole 64-lane wavefront.
constexpr uint32_t kHeavyBlocks = 16;
constexpr uint32_t kHeavyWaves = TPB / 64;
__global__ void k_heavy_flags(const uint4* vinfo, const uint32_t* vnc, uint32_t n, uint32_t* flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) flag[i] = i < n && vinfo[i].z + vnc[i] > 32u ? 1u : 0u;
}
__global__ void k_heavy_list(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* list) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) list[pos[i]] = i;
}
// Heavy wavefront w of W sweeps entries w, w + W, ... of [h0, h1) of the
// heavy list (the heavy slots of slot range [off, off + count)).
template <class R>
__device__ __forceinline__ void sweep_heavy(const SweepArgs<R>& a, uint32_t h0, uint32_t h1, uint32_t w,
                                            uint32_t W) {
  const int lane = threadIdx.x & 63;
  for (uint32_t k = h0 + w; k < h1; k += W) {
    const uint32_t slot = a.hlist[k];
    solve_vertex<64, R>(a, a.vinfo[slot], a.vnc[slot], lane);
  }
}

// One colour of one iteration.  hb > 0 (32 lanes per vertex only): the last
// hb logical blocks sweep the colour's vertices with more than 32 records,
// 64 lanes each, and the 32-lane groups skip them.
template <int L, class R>
__global__ __launch_bounds__(TPB, sizeof(R) == 4 ? 4 : 2) void k_sweep(SweepArgs<R> a, uint32_t off,
                                                                       uint32_t count, IterTail tail, uint32_t hb) {
  const int lane = threadIdx.x & (L - 1);
  // Vertices of a colour are in index order, so neighbouring groups share
  // neighbours: keep consecutive groups on one XCD's L2.
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x), nb = gridDim.x - hb;
  if (lb < nb) {
    const uint32_t grp = (lb * blockDim.x + threadIdx.x) / L;
    if (grp < count) {
      // The vertex's slot data and the stop flag are loaded together (one
      // memory round trip instead of two before the solve starts).
      const uint4 vi = a.vinfo[off + grp];
      const uint32_t nc = a.vnc[off + grp];
      const bool heavy = L == 32 && hb && vi.z + nc > 32u;
      if (!heavy && !*a.stop) solve_vertex<L, R>(a, vi, nc, lane);
    }
  } else if (L == 32 && !*a.stop) {
    sweep_heavy(a, a.hpos[off], a.hpos[off + count], (lb - nb) * kHeavyWaves + threadIdx.x / 64,
                hb * kHeavyWaves);
  }
  if (tail.sc) iter_tail(tail);  // block-uniform
}

// One colour index of one iteration of many solvers in a single launch: slot
// k covers groups [start_k, start_k + count) of solver inst_k's colour
// starting at `off` in its order.  Each solver's vertices see exactly the
// work of its own k_sweep launch.
struct BatchSlot {
  uint32_t start, off, count, inst;
};
template <int L, class R>
__global__ __launch_bounds__(TPB) void k_sweep_batch(const SweepArgs<R>* args, const BatchSlot* slots,
                                                     uint32_t nslots, uint32_t total, uint32_t hb) {
  const int lane = threadIdx.x & (L - 1);
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x), nb = gridDim.x - hb;
  if (L == 32 && lb >= nb) {
    // The heavy vertices of every slot, as in k_sweep: heavy wavefront w of W
    // takes entries w, w + W, ... of the slots' heavy ranges laid end to end
    // (the block size is the batch's tuned one, a multiple of 64).
    const uint32_t wpb = blockDim.x / 64;
    const uint32_t w = (lb - nb) * wpb + threadIdx.x / 64, W = hb * wpb;
    uint32_t base = 0;
    for (uint32_t s = 0; s < nslots; ++s) {
      const BatchSlot sl = slots[s];
      const SweepArgs<R>& a = args[sl.inst];
      const uint32_t h0 = a.hpos[sl.off], h1 = a.hpos[sl.off + sl.count];
      const uint32_t first = base + (w >= base ? w - base : (W - (base - w) % W) % W);
      if (!*a.stop && first < base + (h1 - h0)) sweep_heavy(a, h0 + (first - base), h1, 0, W);
      base += h1 - h0;
    }
    return;
  }
  const uint32_t grp = (lb * blockDim.x + threadIdx.x) / L;
  if (grp >= total) return;
  uint32_t lo = 0, hi = nslots;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (slots[mid].start <= grp) lo = mid;
    else hi = mid;
  }
  const BatchSlot sl = slots[lo];
  const SweepArgs<R>& a = args[sl.inst];
  if (*a.stop) return;
  const uint32_t slot = sl.off + grp - sl.start;
  const uint4 vi = a.vinfo[slot];
  const uint32_t nc = a.vnc[slot];
  if (L == 32 && hb && vi.z + nc > 32u) return;  // swept by the heavy blocks
  solve_vertex<L, R>(a, vi, nc, lane);
}

struct IterEnd {
  Scalars* sc;
  int it;
  double threshold, tol, nfree3;
};
__global__ void k_iter_end_batch(const IterEnd* e, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const IterEnd q = e[i];
  iter_end(q.sc, q.sc->clipped, q.it, q.threshold, q.tol, q.nfree3);
}

// Gradient probe (src/vbd.wgsl:142-147): (g, E) of the local objective.
template <int L, class R>
__global__ void k_probe(SweepArgs<R> a, uint32_t n, double* out) {
  const int lane = threadIdx.x & (L - 1);
  const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  if (v >= n) return;
  const double4 p4 = a.pos[v];
  const Frame f = frame_of<R>(p4);
  V3<R> x = rel<R>(p4, f);
  uint32_t s0 = a.term_off[v], ns = a.term_off[v + 1] - s0;
  uint32_t c0 = a.coff[v], nt = ns + (a.coff[v + 1] - c0);
  const LaneTerm<R> lt = load_lane_term(a, lane, s0, ns, c0, nt, f);
  cvm::Deriv<R> acc = local_deriv<L>(a, lt, v, x, lane, s0, ns, c0, nt, f);
  if (lane == 0) {
    const double4 tg = a.tgt[v];
    const R m = (R)tg.w;
    V3<R> dx = x - rel<R>(tg, f);
    out[4 * v] = acc.g.x + dx.x * m;
    out[4 * v + 1] = acc.g.y + dx.y * m;
    out[4 * v + 2] = acc.g.z + dx.z * m;
    out[4 * v + 3] = acc.e + R(0.5) * m * cvm::dot(dx, dx);
  }
}

// True residual contribution |g_v|^2 of the local objective at the current
// positions (g = elastic + contact + grab gradient + inertia (x - target)),
// 0 for pinned vertices: the Newton-style rms of src/solver.rs:229-231.
template <int L, class R>
__global__ __launch_bounds__(TPB) void k_true_grad(SweepArgs<R> a, uint32_t n, double* out) {
  const int lane = threadIdx.x & (L - 1);
  const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  if (v >= n) return;
  const double4 p4 = a.pos[v];
  if (p4.w > 0.5) {
    if (lane == 0) out[v] = 0.0;
    return;
  }
  const Frame f = frame_of<R>(p4);
  V3<R> x = rel<R>(p4, f);
  uint32_t s0 = a.term_off[v], ns = a.term_off[v + 1] - s0;
  uint32_t c0 = a.coff[v], nt = ns + (a.coff[v + 1] - c0);
  const LaneTerm<R> lt = load_lane_term(a, lane, s0, ns, c0, nt, f);
  cvm::Deriv<R> acc = local_deriv<L, R, false>(a, lt, v, x, lane, s0, ns, c0, nt, f);
  if (lane == 0) {
    const double4 tg = a.tgt[v];
    V3<R> g = acc.g + (x - rel<R>(tg, f)) * (R)tg.w;
    out[v] = (double)cvm::dot(g, g);
  }
}
// The true residual of an iteration in two launches: k_true_resid sums the
// |g_v|^2 of each block's vertices in a fixed order into part[block];
// k_resid_end sums the partials in block order into sc->true_residual_sum
// and, with `end`, runs the iteration's end (iter_end).  With `end`, a
// pending refresh or a converged step (the stop flag, set before the
// launches) makes both no-ops, as it makes the sweeps: the residual would not
// be read.  (A last-block ticket would save the second launch but needs a
// device-scope release per block, an L2 write-back each on this GPU.)
struct ResidEnd {
  int end;
  int it;
  double threshold, tol, nfree3;
};
template <int L, class R>
__global__ __launch_bounds__(TPB) void k_true_resid(SweepArgs<R> a, uint32_t n, double* part, int honor_stop) {
  if (honor_stop && *a.stop) return;  // grid-uniform
  constexpr int G = TPB / L;
  __shared__ double red[G];
  const int lane = threadIdx.x & (L - 1);
  const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  double gg = 0.0;
  if (v < n) {
    const double4 p4 = a.pos[v];
    if (p4.w <= 0.5) {
      const Frame f = frame_of<R>(p4);
      V3<R> x = rel<R>(p4, f);
      uint32_t s0 = a.term_off[v], ns = a.term_off[v + 1] - s0;
      uint32_t c0 = a.coff[v], nt = ns + (a.coff[v + 1] - c0);
      const LaneTerm<R> lt = load_lane_term(a, lane, s0, ns, c0, nt, f);
      cvm::Deriv<R> acc = local_deriv<L, R, false>(a, lt, v, x, lane, s0, ns, c0, nt, f);
      const double4 tg = a.tgt[v];
      V3<R> g = acc.g + (x - rel<R>(tg, f)) * (R)tg.w;
      gg = (double)cvm::dot(g, g);
    }
  }
  if (lane == 0) red[threadIdx.x / L] = gg;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int k = 0; k < G; ++k) sum += red[k];
    part[blockIdx.x] = sum;
  }
}
__global__ __launch_bounds__(TPB) void k_resid_end(const double* part, uint32_t np, Scalars* sc, ResidEnd e) {
  if (e.end && sc->stop) return;
  __shared__ double red[TPB];
  double sum = 0.0;
  for (uint32_t k = threadIdx.x; k < np; k += TPB) sum += part[k];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int k = TPB / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc->true_residual_sum = red[0];
    if (e.end) iter_end(sc, sc->clipped, e.it, e.threshold, e.tol, e.nfree3);
  }
}

// Deterministic two-level sum: block b of k_sum_part sums the contiguous
// chunk [b c, (b + 1) c) in a fixed order; k_sum_final sums the partials in
// block order.
constexpr uint32_t kSumChunk = 4096;
__global__ __launch_bounds__(256) void k_sum_part(const double* a, uint32_t n, double* part) {
  __shared__ double red[256];
  const uint32_t lo = blockIdx.x * kSumChunk, hi = min(n, lo + kSumChunk);
  double s = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) s += a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_sum_f64(const double* a, uint32_t n, double* out) {
  __shared__ double red[1024];
  double s = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Warm-start features of each moving vertex (12 floats, dimensionless), with
// L = |g| dt^2 / (1 + c dt), the free-fall displacement of one step:
//   0-2  offset of the step's own guess from x / L: the adaptive inertial
//        prediction of Prediction::guess (src/vbd/prediction.rs:15-41),
//        (v dt + f g dt^2) / (1 + c dt), with f = 1 without history
//   3-5  previous velocity v dt / L
//   6-7  mean Green strain of the incident triangles along warp and weft, x 100
//   8-10 mean velocity of the one-ring neighbours, v dt / L
//   11   distance to the nearest contact primitive / query radius (<= 1)
__global__ void k_warm_features(const double* x, const double* v, const double* guess, const uint32_t* term_off,
                                const cvm::TermRec64* terms, const uint32_t* nb_off, const uint32_t* nb,
                                const unsigned long long* dist_bits, uint32_t n, double dt, double L, double rq,
                                float* feat) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]}, w[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  float* f = feat + 12 * (size_t)i;
  for (int k = 0; k < 3; ++k) {
    f[k] = (float)((guess[3 * i + k] - p[k]) / L);
    f[3 + k] = (float)(w[k] * dt / L);
  }
  double ew = 0, ev = 0;
  int tri = 0;
  for (uint32_t k = term_off[i]; k < term_off[i + 1]; ++k) {
    const cvm::TermRec64& r = terms[k];
    if (cvm::rec_kind(r.head) != cvm::TERM_MEMBRANE) continue;
    double e1[3], e2[3], a[3], b[3];
    for (int c = 0; c < 3; ++c) {
      e1[c] = x[3 * r.ids[0] + c] - p[c];
      e2[c] = x[3 * r.ids[1] + c] - p[c];
      a[c] = e1[c] * r.f[0] + e2[c] * r.f[1];
      b[c] = e1[c] * r.f[2] + e2[c] * r.f[3];
    }
    ew += 0.5 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2] - 1.0);
    ev += 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2] - 1.0);
    ++tri;
  }
  f[6] = (float)(tri ? 100.0 * ew / tri : 0.0);
  f[7] = (float)(tri ? 100.0 * ev / tri : 0.0);
  double mv[3] = {0, 0, 0};
  int cnt = 0;
  for (uint32_t k = nb_off[i]; k < nb_off[i + 1]; ++k) {
    uint32_t j = nb[k];
    if (j >= n) continue;
    for (int c = 0; c < 3; ++c) mv[c] += v[3 * j + c];
    ++cnt;
  }
  for (int c = 0; c < 3; ++c) f[8 + c] = (float)(cnt ? mv[c] / cnt * dt / L : 0.0);
  double d = __longlong_as_double((long long)dist_bits[i]);
  f[11] = (float)(fmin(d, rq) / rq);
}
// One pass of one-ring averaging of the network's correction (moving
// neighbours only): out_i = (in_i + sum_j in_j) / (1 + #j).
__global__ void k_warm_smooth(const float* in, const uint32_t* nb_off, const uint32_t* nb, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s[3] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
  int cnt = 1;
  for (uint32_t k = nb_off[i]; k < nb_off[i + 1]; ++k) {
    uint32_t j = nb[k];
    if (j >= n) continue;
    for (int c = 0; c < 3; ++c) s[c] += in[3 * j + c];
    ++cnt;
  }
  for (int c = 0; c < 3; ++c) out[3 * i + c] = s[c] / (float)cnt;
}
// Learned guess for moving vertices: the step's own guess (the adaptive
// inertial prediction) plus scale * L times the network's correction, or with
// normal_only its component along the area-weighted vertex normal at x (the
// sum of the incident triangles' (x1 - x) x (x2 - x)).
__global__ void k_warm_guess(const float* pred, const uint8_t* pinned, const double* x, const uint32_t* term_off,
                             const cvm::TermRec64* terms, uint32_t n, double L, int normal_only, double* guess) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || pinned[i]) return;
  double c[3] = {L * (double)pred[3 * i], L * (double)pred[3 * i + 1], L * (double)pred[3 * i + 2]};
  if (normal_only) {
    double nx = 0, ny = 0, nz = 0;
    for (uint32_t k = term_off[i]; k < term_off[i + 1]; ++k) {
      const cvm::TermRec64& r = terms[k];
      if (cvm::rec_kind(r.head) != cvm::TERM_MEMBRANE) continue;
      double a[3], b[3];
      for (int q = 0; q < 3; ++q) {
        a[q] = x[3 * r.ids[0] + q] - x[3 * i + q];
        b[q] = x[3 * r.ids[1] + q] - x[3 * i + q];
      }
      nx += a[1] * b[2] - a[2] * b[1];
      ny += a[2] * b[0] - a[0] * b[2];
      nz += a[0] * b[1] - a[1] * b[0];
    }
    double nn = nx * nx + ny * ny + nz * nz;
    double d = nn > 0 ? (c[0] * nx + c[1] * ny + c[2] * nz) / nn : 0.0;
    c[0] = d * nx;
    c[1] = d * ny;
    c[2] = d * nz;
  }
  for (int k = 0; k < 3; ++k) guess[3 * i + k] += c[k];
}
// Step prologue reductions (single block, fixed order => deterministic):
// origin = centroid (src/gpu_vbd.rs:156), predicted motion max|target - x|
// (src/vbd.rs:298-302) and max |x|_inf of the moving part (padded radius).
// Centroid (the origin of the sweep's relative positions), largest inertial
// motion |target - x| (for rq, src/vbd.rs:298-304) and largest |x|_inf (the
// detection's padding scale), in two deterministic passes: fixed contiguous
// chunks per block, then the chunks' partials in block order.
constexpr uint32_t kPrologueChunk = 2048;
__global__ __launch_bounds__(256) void k_prologue_partial(const double* x, const double* v, uint32_t n, double dt,
                                                          double damp, double gx, double gy, double gz,
                                                          double* part) {
  __shared__ double red[5][256];
  double ax = 0, ay = 0, az = 0, mo = 0, mabs = 0;
  uint32_t i0 = blockIdx.x * kPrologueChunk, i1 = min(n, i0 + kPrologueChunk);
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
    ax += px;
    ay += py;
    az += pz;
    double tx = px + (v[3 * i] * dt + gx * dt * dt) / damp;
    double ty = py + (v[3 * i + 1] * dt + gy * dt * dt) / damp;
    double tz = pz + (v[3 * i + 2] * dt + gz * dt * dt) / damp;
    double dx = tx - px, dy = ty - py, dz = tz - pz;
    mo = fmax(mo, sqrt(dx * dx + dy * dy + dz * dz));
    mabs = fmax(mabs, fmax(fabs(px), fmax(fabs(py), fabs(pz))));
  }
  red[0][threadIdx.x] = ax;
  red[1][threadIdx.x] = ay;
  red[2][threadIdx.x] = az;
  red[3][threadIdx.x] = mo;
  red[4][threadIdx.x] = mabs;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
      for (int k = 3; k < 5; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 5) part[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void k_prologue_final(const double* part, uint32_t nparts, uint32_t n, Scalars* sc) {
  if (threadIdx.x != 0) return;
  double ax = 0, ay = 0, az = 0, mo = 0, mabs = 0;
  for (uint32_t b = 0; b < nparts; ++b) {
    ax += part[5 * b];
    ay += part[5 * b + 1];
    az += part[5 * b + 2];
    mo = fmax(mo, part[5 * b + 3]);
    mabs = fmax(mabs, part[5 * b + 4]);
  }
  sc->origin[0] = ax / n;
  sc->origin[1] = ay / n;
  sc->origin[2] = az / n;
  sc->motion = mo;
  sc->scale = mabs;
}

// Targets, inertia, adaptive guess (src/vbd.rs:283-297, src/vbd/prediction.rs:15-41).
__global__ void k_prologue(const double* x, const double* v, const double* acc, const double* masses,
                           const uint8_t* pinned, uint32_t n, const double* fixed, uint32_t n_fixed, double dt,
                           double damping, double gx, double gy, double gz, int pred_valid, const Scalars* sc,
                           double4* pos, double4* tgt, double* guess, double* x_old, double* v_old) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double ox = sc->origin[0], oy = sc->origin[1], oz = sc->origin[2];
  if (i >= n) {
    uint32_t j = i - n;
    if (j < n_fixed)
      pos[i] = make_double4(fixed[3 * j] - ox, fixed[3 * j + 1] - oy, fixed[3 * j + 2] - oz, 1.0);
    return;
  }
  double damp = 1.0 + damping * dt;
  double p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  double w[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  double g[3] = {gx, gy, gz};
  double t[3];
  for (int k = 0; k < 3; ++k) t[k] = p[k] + (w[k] * dt + g[k] * dt * dt) / damp;
  double inertia = masses[i] * damp / (dt * dt);
  tgt[i] = make_double4(t[0] - ox, t[1] - oy, t[2] - oz, inertia);
  double g2 = gx * gx + gy * gy + gz * gz;
  double factor = 1.0;
  if (pred_valid && g2 > 1e-20) {
    double ag = acc[3 * i] * gx + acc[3 * i + 1] * gy + acc[3 * i + 2] * gz;
    factor = fmin(fmax(ag / g2, 0.0), 1.0);
  }
  for (int k = 0; k < 3; ++k) {
    guess[3 * i + k] = p[k] + (w[k] * dt + (factor * g[k]) * (dt * dt)) / (1.0 + damping * dt);
    x_old[3 * i + k] = p[k];
    v_old[3 * i + k] = w[k];
  }
  if (pinned[i]) pos[i] = make_double4(p[0] - ox, p[1] - oy, p[2] - oz, 1.0);
}

// Balls from the last detection.  For the f32 sweep the radius is shrunk by
// four f32 ulps of itself: a point truncated in a vertex's frame (where the
// centre is within the radius of the origin) and converted back then never
// leaves the f64 ball (the f32 safety margin of src/gpu_vbd.rs:258-274); the
// f64 sweep uses the radius unchanged.
template <class R>
__global__ void k_balls(const double* xdet, const double* radii, uint32_t n, const Scalars* sc, double4* ball,
                        double* centers, double* radii_used) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cx = xdet[3 * i] - sc->origin[0], cy = xdet[3 * i + 1] - sc->origin[1],
         cz = xdet[3 * i + 2] - sc->origin[2];
  double r = radii[i];
  if (sizeof(R) == 4) r -= 4.0 * 1.1920928955078125e-07 * r;
  ball[i] = make_double4(cx, cy, cz, r);
  centers[3 * i] = xdet[3 * i];
  centers[3 * i + 1] = xdet[3 * i + 1];
  centers[3 * i + 2] = xdet[3 * i + 2];
  radii_used[i] = r;
}

// Conservative warm start: the guess clipped into its ball (src/vbd.rs:316-324).
__global__ void k_truncate_guess(const double* guess, const double* centers, const double* radii_used,
                                 const uint8_t* pinned, uint32_t n, const Scalars* sc, double4* pos,
                                 unsigned long long* clipped) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || pinned[i]) return;
  double d[3], len2 = 0;
  for (int k = 0; k < 3; ++k) {
    d[k] = guess[3 * i + k] - centers[3 * i + k];
    len2 += d[k] * d[k];
  }
  double len = sqrt(len2), r = radii_used[i];
  double p[3];
  bool clip = len > r;
  for (int k = 0; k < 3; ++k) p[k] = clip ? centers[3 * i + k] + d[k] * (r / len) : guess[3 * i + k];
  pos[i] = make_double4(p[0] - sc->origin[0], p[1] - sc->origin[1], p[2] - sc->origin[2], 0.0);
  if (clip) atomicAdd(clipped, 1ull);
}

// Current f64 positions for a mid-step detection.
__global__ void k_current(const double4* pos, const double* x_old, const uint8_t* pinned, uint32_t n,
                          const Scalars* sc, double* x64) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pinned[i]) {
    for (int k = 0; k < 3; ++k) x64[3 * i + k] = x_old[3 * i + k];
    return;
  }
  auto q = pos[i];
  x64[3 * i] = sc->origin[0] + (double)q.x;
  x64[3 * i + 1] = sc->origin[1] + (double)q.y;
  x64[3 * i + 2] = sc->origin[2] + (double)q.z;
}

// Commit (src/vbd.rs:399-411, src/vbd/prediction.rs:43-54): finiteness,
// trust-region invariant, velocity, and the acceleration history.
__global__ void k_finalize(const double4* pos, const double* x_old, const double* v_old,
                           const double* centers, const double* radii, const uint8_t* pinned, uint32_t n,
                           double dt, double damping, double tol, double* x_out, double* v_out, double* acc,
                           Scalars* sc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xn[3];
  if (pinned[i]) {
    for (int k = 0; k < 3; ++k) xn[k] = x_old[3 * i + k];
  } else {
    auto q = pos[i];
    xn[0] = sc->origin[0] + (double)q.x;
    xn[1] = sc->origin[1] + (double)q.y;
    xn[2] = sc->origin[2] + (double)q.z;
  }
  bool fin = isfinite(xn[0]) && isfinite(xn[1]) && isfinite(xn[2]);
  if (!fin) atomicOr(&sc->err_nonfinite, 1u);
  double d2 = 0;
  for (int k = 0; k < 3; ++k) {
    double d = xn[k] - centers[3 * i + k];
    d2 += d * d;
  }
  if (fin && !pinned[i] && sqrt(d2) > radii[i] + tol) atomicOr(&sc->err_trust, 1u);
  for (int k = 0; k < 3; ++k) {
    double vn = (xn[k] - x_old[3 * i + k]) / dt;
    x_out[3 * i + k] = xn[k];
    v_out[3 * i + k] = vn;
    acc[3 * i + k] = (vn - v_old[3 * i + k]) / dt + damping * vn;
  }
}


void read_scalars(cv_gpu& g) {
  CV_HIP(hipMemcpyAsync(g.h_scal, g.d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, g.stream));
  CV_HIP(hipStreamSynchronize(g.stream));
}
void write_scalars(cv_gpu& g) {
  CV_HIP(hipMemcpyAsync(g.d_scal, g.h_scal, sizeof(Scalars), hipMemcpyHostToDevice, g.stream));
}

// Device arrays of the sweep in precision R.
template <class R>
struct PBufs;
template <>
struct PBufs<float> {
  static double4* pos(cv_gpu& g) { return g.b.pos_d.p; }
  static double4* tgt(cv_gpu& g) { return g.b.tgt_d.p; }
  static double4* ball(cv_gpu& g) { return g.b.ball_d.p; }
  static const cvm::TermRec* terms(cv_gpu& g) { return g.dev.terms.p; }
  static const cvm::TermRec* crec(cv_gpu& g) { return g.b.crec.p; }
};
template <>
struct PBufs<double> {
  static double4* pos(cv_gpu& g) { return g.b.pos_d.p; }
  static double4* tgt(cv_gpu& g) { return g.b.tgt_d.p; }
  static double4* ball(cv_gpu& g) { return g.b.ball_d.p; }
  static const cvm::TermRec64* terms(cv_gpu& g) { return g.dev.terms_d.p; }
  static const cvm::TermRec64* crec(cv_gpu& g) { return g.b.crec_d.p; }
};

template <class R>
SweepArgs<R> sweep_args(cv_gpu& g, const cv_settings* set, const cv_grab* grab) {
  SweepArgs<R> a{};
  a.vinfo = g.b.vinfo.p;
  a.vnc = g.b.vnc.p;
  a.pos = PBufs<R>::pos(g);
  a.tgt = PBufs<R>::tgt(g);
  a.ball = PBufs<R>::ball(g);
  a.term_off = g.dev.term_off.p;
  a.terms = PBufs<R>::terms(g);
  a.coff = g.b.coff.p;
  a.crec = PBufs<R>::crec(g);
  a.gsq = g.b.gsq.p;
  a.clipped = &g.d_scal->clipped;
  a.stop = &g.d_scal->stop;
  a.hpos = g.lanes == 32 ? g.b.hpos.p : nullptr;
  a.hlist = g.lanes == 32 ? g.b.hlist.p : nullptr;
  a.cr = (R)set->contact.radius;
  a.ck = (R)set->contact.stiffness;
  a.grab.active = grab ? 1 : 0;
  if (grab) {
    for (int k = 0; k < 3; ++k) {
      a.grab.v[k] = grab->vertices[k];
      a.grab.w[k] = (R)grab->weights[k];
      a.grab.target[k] = grab->target[k] - g.h_scal->origin[k];
    }
    a.grab.k = (R)grab->stiffness;
  }
  return a;
}

// With profiling on, the launch carries start/stop events that the runtime
// stamps at the kernel's own begin and end (hipExtLaunchKernelGGL), so the
// summed time is kernel execution only, as a kernel trace reports it.
template <int L, class R>
void launch_sweep(const SweepArgs<R>& a, uint32_t off, uint32_t count, const IterTail& tail, hipStream_t st,
                  hipEvent_t e0, hipEvent_t e1) {
  const uint32_t hb = L == 32 && a.hpos ? kHeavyBlocks : 0;
  dim3 grid(blocks((uint64_t)count * L) + hb), block(TPB);
  if (e0)
    hipExtLaunchKernelGGL(k_sweep<L, R>, grid, block, 0, st, e0, e1, 0, a, off, count, tail, hb);
  else
    k_sweep<L, R><<<grid, block, 0, st>>>(a, off, count, tail, hb);
}
template <class R>
void sweep_color(cv_gpu& g, const SweepArgs<R>& a, uint32_t off, uint32_t count, const IterTail& tail = IterTail{},
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  switch (g.lanes) {
    case 8: launch_sweep<8>(a, off, count, tail, g.stream, e0, e1); break;
    case 16: launch_sweep<16>(a, off, count, tail, g.stream, e0, e1); break;
    case 64: launch_sweep<64>(a, off, count, tail, g.stream, e0, e1); break;
    default: launch_sweep<32>(a, off, count, tail, g.stream, e0, e1); break;
  }
}

// Host greedy colouring of the current contact set (the detection left the
// contact list in g.h_cont and its count in the scalars it read back).
// Colours of the vertices an incremental colouring changed: (vertex, colour) pairs.
__global__ void k_apply_colors(const uint32_t* chg, uint32_t k, uint32_t* color) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) color[chg[2 * i]] = chg[2 * i + 1];
}
uint32_t color_on_host(cv_gpu& g, const cv_solver& s, const cv_grab* grab) {
  uint64_t m = g.h_scal->contacts;
  const uint32_t* cids = g.h_cont;
  const uint32_t* ccount = g.h_cont + 5 * g.h_cont_cap;
  uint32_t n = s.n;
  // The colouring is a function of the solver, the contact list and the grab
  // vertices: when all three equal the last colouring's, it is reused (the
  // order already sits on the device).  Contact-free steps always reuse it.
  uint32_t key[4] = {grab ? 1u : 0u, grab ? grab->vertices[0] : 0u,
                     grab ? grab->vertices[1] : 0u, grab ? grab->vertices[2] : 0u};
  if (g.color_valid && g.color_generation == s.generation && g.color_m == m &&
      std::memcmp(g.color_key, key, sizeof key) == 0 &&
      std::equal(cids, cids + 5 * m, g.color_cids.begin()) &&
      std::equal(ccount, ccount + m, g.color_cids.begin() + 5 * m))
    return (uint32_t)(g.color_off.size() - 1);
  g.color_valid = false;
  g.color_cids.assign(cids, cids + 5 * m);
  g.color_cids.insert(g.color_cids.end(), ccount, ccount + m);
  ColoringState& cs = g.coloring;
  uint32_t k = update_colors(s, cs, cids, ccount, m, grab, false, g.color_off);
  // The device keeps every vertex's colour; the colour-ordered slot table is
  // its stable sort by colour (vertex order within a colour, as the host
  // greedy lists them).  A full colouring uploads all colours, an incremental
  // one only the vertices it recoloured.
  // b.order is rebuilt from all colours when another colouring (the fast
  // one) wrote it last or the previous update did not finish.
  const bool resync = cs.full || !g.order_host;
  g.order_host = false;
  g.b.color_dev.reserve(n);
  if (resync) {
    g.b.color_dev.upload(cs.color_of, g.stream);
  } else if (!cs.changed.empty()) {
    const size_t nc = cs.changed.size();
    if (g.h_cchg_cap < nc) {
      if (g.h_cchg) CV_HIP(hipHostFree(g.h_cchg));
      g.h_cchg = nullptr;
      const size_t cap = nc + nc / 2 + 1024;
      CV_HIP(hipHostMalloc((void**)&g.h_cchg, 2 * cap * sizeof(uint32_t)));
      g.h_cchg_cap = cap;
    }
    // The previous call's copy out of this buffer finished at the detection's
    // read-back, so it can be rewritten.
    for (size_t i = 0; i < nc; ++i) {
      g.h_cchg[2 * i] = cs.changed[i];
      g.h_cchg[2 * i + 1] = cs.color_of[cs.changed[i]];
    }
    g.b.color_chg.reserve(2 * nc);
    CV_HIP(hipMemcpyAsync(g.b.color_chg.p, g.h_cchg, 2 * nc * sizeof(uint32_t), hipMemcpyHostToDevice, g.stream));
    k_apply_colors<<<blocks((uint32_t)nc), TPB, 0, g.stream>>>(g.b.color_chg.p, (uint32_t)nc, g.b.color_dev.p);
    CV_HIP(hipGetLastError());
  }
  if (resync || !cs.changed.empty()) device_color_order(g, n, k);
  g.order_host = true;
  g.color_valid = true;
  g.color_generation = s.generation;
  g.color_m = m;
  std::memcpy(g.color_key, key, sizeof key);
  return k;
}

// Per colour-ordered slot i: the vertex order[i], its first term record and
// term count, its first contact record and contact-record count, so a sweep
// group loads them in one step and its term records in the next.
__global__ void k_vertex_info(const uint32_t* order, const uint32_t* term_off, const uint32_t* coff, uint32_t n,
                              uint4* vinfo, uint32_t* vnc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v = order[i];
  vinfo[i] = make_uint4(v, term_off[v], term_off[v + 1] - term_off[v], coff[v]);
  vnc[i] = coff[v + 1] - coff[v];
}
void vertex_info(cv_gpu& g, uint32_t n) {
  g.b.vinfo.reserve(n);
  g.b.vnc.reserve(n);
  k_vertex_info<<<blocks(n), TPB, 0, g.stream>>>(g.b.order.p, g.dev.term_off.p, g.b.coff.p, n, g.b.vinfo.p,
                                                 g.b.vnc.p);
  CV_HIP(hipGetLastError());
  if (g.lanes == 32) {
    g.b.hflag.reserve(n + 1);
    g.b.hpos.reserve(n + 1);
    g.b.hlist.reserve(std::max<uint32_t>(n, 1));
    k_heavy_flags<<<blocks(n + 1), TPB, 0, g.stream>>>(g.b.vinfo.p, g.b.vnc.p, n, g.b.hflag.p);
    CV_HIP(hipGetLastError());
    device_scan_u32(g, g.b.hflag.p, g.b.hpos.p, (size_t)n + 1);
    k_heavy_list<<<blocks(n), TPB, 0, g.stream>>>(g.b.hflag.p, g.b.hpos.p, n, g.b.hlist.p);
    CV_HIP(hipGetLastError());
  }
}

// The fast colouring (on the device, gpu_fast_colors) of the last detection;
// reused while the detections find no contacts and the grab is unchanged.
uint32_t color_on_device(cv_gpu& g, const cv_solver& s, const cv_grab* grab) {
  const uint64_t m = g.h_scal->contacts;
  uint32_t key[4] = {(grab ? 1u : 0u) | 4u, grab ? grab->vertices[0] : 0u, grab ? grab->vertices[1] : 0u,
                     grab ? grab->vertices[2] : 0u};
  if (g.color_valid && g.color_generation == s.generation && g.color_m == 0 && m == 0 &&
      std::memcmp(g.color_key, key, sizeof key) == 0)
    return (uint32_t)(g.color_off.size() - 1);
  g.color_valid = false;
  g.color_cids.clear();
  g.order_host = false;  // b.order is about to hold the fast colouring's order
  uint32_t k = gpu_fast_colors(g, s, grab, g.color_off);
  g.color_valid = true;
  g.color_generation = s.generation;
  g.color_m = m;
  std::memcpy(g.color_key, key, sizeof key);
  return k;
}

// Origin, motion and scale of the positions in b.x64 into the scalars.
void prologue_reduce(cv_gpu& g, uint32_t n, double dt, double damp, const double* gravity) {
  uint32_t parts = std::max<uint32_t>(1, (n + kPrologueChunk - 1) / kPrologueChunk);
  g.b.red_parts.reserve(5 * (size_t)parts);
  k_prologue_partial<<<parts, 256, 0, g.stream>>>(g.b.x64.p, g.b.v64.p, n, dt, damp, gravity[0], gravity[1],
                                                  gravity[2], g.b.red_parts.p);
  k_prologue_final<<<1, 64, 0, g.stream>>>(g.b.red_parts.p, parts, n, g.d_scal);
  CV_HIP(hipGetLastError());
}

// Deterministic sum of a[0, n) into *out on stream st.
void sum_f64(cv_gpu& g, const double* a, uint32_t n, double* out, hipStream_t st) {
  const uint32_t parts = std::max<uint32_t>(1, (n + kSumChunk - 1) / kSumChunk);
  g.b.sum_parts.reserve(parts);
  k_sum_part<<<parts, 256, 0, st>>>(a, n, g.b.sum_parts.p);
  k_sum_f64<<<1, 256, 0, st>>>(g.b.sum_parts.p, parts, out);
  CV_HIP(hipGetLastError());
}
template <class R>
void true_residual(cv_gpu& g, const SweepArgs<R>& a, uint32_t n, hipStream_t st) {
  // A derivative pass without the line search: 16 lanes per vertex (two
  // records per lane for a typical vertex) unless the sweep uses fewer.
  if (g.lanes == 8)
    k_true_grad<8, R><<<blocks((uint64_t)n * 8), TPB, 0, st>>>(a, n, g.b.gtrue.p);
  else
    k_true_grad<16, R><<<blocks((uint64_t)n * 16), TPB, 0, st>>>(a, n, g.b.gtrue.p);
  sum_f64(g, g.b.gtrue.p, n, &g.d_scal->true_residual_sum, st);
}
template <class R>
void true_residual(cv_gpu& g, const SweepArgs<R>& a, uint32_t n) {
  true_residual<R>(g, a, n, g.stream);
}
// The two-launch form on the solver's stream: residual, sum and (with
// e.end) the iteration's end.
template <class R>
void true_residual_fused(cv_gpu& g, const SweepArgs<R>& a, uint32_t n, ResidEnd e) {
  const int L = g.lanes == 8 ? 8 : 16;
  const unsigned nb = blocks((uint64_t)std::max<uint32_t>(n, 1) * L);
  g.b.sum_parts.reserve(nb);
  if (L == 8)
    k_true_resid<8, R><<<nb, TPB, 0, g.stream>>>(a, n, g.b.sum_parts.p, e.end);
  else
    k_true_resid<16, R><<<nb, TPB, 0, g.stream>>>(a, n, g.b.sum_parts.p, e.end);
  k_resid_end<<<1, TPB, 0, g.stream>>>(g.b.sum_parts.p, nb, g.d_scal, e);
  CV_HIP(hipGetLastError());
}

struct StepOut {
  uint64_t detections = 0, truncated = 0;
  uint32_t colors = 0, iterations_run = 0;
  uint64_t sweep_launches = 0, sweep_timed_launches = 0;
  double sweep_ms = 0, sweep_bytes = 0, sweep_bytes_min = 0, sweep_timed_bytes_min = 0;

  double detection_ms = 0, refit_ms = 0, candidates_ms = 0, narrow_ms = 0;
  uint64_t candidate_pairs = 0, grid_tests = 0;
  uint64_t bvh_fallbacks = 0;  // LBVH detections redone with the grid
  bool true_residual = false;
  bool diagnostics = false;  // garment diagnostics computed after the commit
};

// Detection + balls (+ colouring) on the positions held in b.x64.
template <class R>
void detect_and_color(cv_gpu& g, const cv_solver& s, const cv_settings* set, const cv_grab* grab, double rq,
                      StepOut& out, bool first) {
  double t0 = now_ms();
  const bool fast = (set->flags & CV_FLAG_FAST_COLORING) != 0;
  struct Restore {
    cv_gpu& g;
    ~Restore() { g.host_contacts = true; }
  } restore{g};
  g.host_contacts = !fast;  // the fast colouring reads the contacts on the device
  if (first) g.b.bvh_rebuild = true;  // LBVH: new trees per step, refits at refreshes
  const uint64_t fb = g.b.bvh_fallbacks;
  gpu_detect(g, s, rq, set->contact, 0);
  out.bvh_fallbacks += g.b.bvh_fallbacks - fb;
  out.detections++;
  out.candidate_pairs += g.h_scal->vf_cand + g.h_scal->ee_cand;
  out.grid_tests += g.h_scal->grid_tests;
  double md = g.h_scal->min_distance;
  if (first && md <= 1e-12)
    fail(CV_ERR_ZERO_SEPARATION, "OGC zero separation; step rejected, reset the arrangement");
  if (first && sizeof(R) == 4 && md < 2e-6)
    fail(CV_ERR_SEPARATION_F32,
         "GPU OGC separation below f32 safety margin (2 micrometres); use f64 reference or reset");
  uint32_t n = s.n;
  k_balls<R><<<blocks(n), TPB, 0, g.stream>>>(g.b.x64.p, g.b.radii64.p, n, g.d_scal, PBufs<R>::ball(g),
                                              g.b.centers64.p, g.b.guess64.p + 3 * (size_t)n);
  CV_HIP(hipGetLastError());
  double t1 = now_ms();
  out.colors = fast ? color_on_device(g, s, grab) : color_on_host(g, s, grab);
  vertex_info(g, n);  // the contact offsets changed, and perhaps the order
  out.detection_ms += now_ms() - t0;
  out.narrow_ms += t1 - t0;
}

// Free-fall displacement of one step, the length unit of the warm-start
// features and predictions.
inline double warm_length(const cv_settings* set) {
  double g = std::sqrt(set->gravity[0] * set->gravity[0] + set->gravity[1] * set->gravity[1] +
                       set->gravity[2] * set->gravity[2]);
  return std::max(g * set->dt * set->dt / (1.0 + set->damping * set->dt), 1e-9);
}
// Features at the positions in b.x64 / velocities in b.v64 and the guess in
// b.guess64 (k_prologue's), using the last detection's distances.
void warm_features(cv_gpu& g, const cv_solver& s, const cv_settings* set, double rq) {
  uint32_t n = s.n;
  g.feat.reserve(12 * (size_t)n);
  k_warm_features<<<blocks(n), TPB, 0, g.stream>>>(g.b.x64.p, g.b.v64.p, g.b.guess64.p, g.dev.term_off.p,
                                                   g.dev.terms_d.p, g.dev.nb_off.p, g.dev.nb.p, g.b.dist_bits.p, n,
                                                   set->dt, warm_length(set), rq, g.feat.p);
  CV_HIP(hipGetLastError());
}
// The network's correction for the features in g.feat, smoothed over the
// one-ring g.warm_smooth times, added to the guess in b.guess64.
void warm_guess(cv_gpu& g, const cv_solver& s, const cv_settings* set) {
  uint32_t n = s.n;
  g.wpred.reserve(6 * (size_t)n);
  mlp_infer_device(g, g.feat.p, n, g.wpred.p);
  float* a = g.wpred.p;
  float* b = g.wpred.p + 3 * (size_t)n;
  for (int k = 0; k < g.warm_smooth; ++k) {
    k_warm_smooth<<<blocks(n), TPB, 0, g.stream>>>(a, g.dev.nb_off.p, g.dev.nb.p, n, b);
    std::swap(a, b);
  }
  k_warm_guess<<<blocks(n), TPB, 0, g.stream>>>(a, g.dev.pinned.p, g.b.x64.p, g.dev.term_off.p, g.dev.terms_d.p, n,
                                                g.warm_scale * warm_length(set), g.warm_normal_only ? 1 : 0,
                                                g.b.guess64.p);
  CV_HIP(hipGetLastError());
}

// One VBD step of one solver as a sequence of phases, so the single-solver
// step and the batched step (many solvers sharing sweep launches) run the
// same device work in the same order per solver:
//   begin()      prologue, detection, colouring, truncated guess;
//   enqueue()    the sweeps of a range of iterations (or the batch driver's
//                shared launches) and their end-of-iteration kernels;
//   after_sync() once the stream has drained: accounting, and the refresh
//                (re-detection and recolouring) a raised `stop` asks for;
//   finish()     residuals, trust-region check, commit of the positions.
// Iterations are queued in batches with no read-back inside a batch: the
// first batch is iteration 0 alone (a refresh after it is common), the next
// runs to the end.  A pending refresh turns the rest of a batch into no-ops;
// the host then re-detects and queues again from the next iteration.
template <class R>
struct StepCtx {
  cv_gpu& g;
  cv_solver& s;
  const cv_settings* set;
  const cv_grab* grab;
  StepOut& out;
  SweepArgs<R> a{};
  double rq = 0, threshold = 0, tol = 0, nfree3 = 1;
  bool want_true = false;
  bool diag = true;  // garment diagnostics after the commit (the last step of a run)
  uint32_t iters = 0, next = 0, done = 0;
  bool finished = false;
  struct Launch {
    size_t ev;
    uint32_t it;
  };
  std::vector<Launch> launches;
  std::vector<uint32_t> timed_its;  // iterations whose launches carry events
  std::vector<size_t> pending_ev;   // completed timed launches not yet read
  size_t nev = 0;                   // event pairs used in this step
  static constexpr uint64_t kTimedStride = 4;

  StepCtx(cv_gpu& g_, cv_solver& s_, const cv_settings* set_, const cv_grab* grab_, StepOut& out_)
      : g(g_), s(s_), set(set_), grab(grab_), out(out_) {}

  void prologue_reduce() {
    double dt = set->dt, damp = 1.0 + set->damping * dt;
    ::prologue_reduce(g, s.n, dt, damp, set->gravity);
    CV_HIP(hipGetLastError());
    CV_HIP(hipMemcpyAsync(g.h_scal, g.d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, g.stream));
  }
  // After prologue_reduce() and a drain of the stream.
  void begin() {
    DevSolver& d = g.dev;
    StepBufs& b = g.b;
    hipStream_t st = g.stream;
    uint32_t n = s.n;
    double dt = set->dt, damp = 1.0 + set->damping * dt;
    double r = set->contact.radius;
    rq = std::min(std::max(r + g.h_scal->motion, 2.0 * r), 4.0 * r);
    g.h_scal->clipped = 0;
    g.h_scal->truncated_total = 0;
    g.h_scal->stop = g.h_scal->stop_it = g.h_scal->iters_run = 0;
    g.h_scal->err_nonfinite = g.h_scal->err_trust = 0;
    write_scalars(g);
    k_prologue<<<blocks(n + d.n_fixed), TPB, 0, st>>>(
        b.x64.p, b.v64.p, d.acc.p, d.masses.p, d.pinned.p, n, d.fixed_pos.p, d.n_fixed, dt, set->damping,
        set->gravity[0], set->gravity[1], set->gravity[2], g.pred_valid ? 1 : 0, g.d_scal, PBufs<R>::pos(g),
        PBufs<R>::tgt(g), b.guess64.p, b.x_old64.p, b.v_old64.p);
    CV_HIP(hipGetLastError());
    detect_and_color<R>(g, s, set, grab, rq, out, true);
    if (set->flags & CV_FLAG_WARM_START) {
      // Learned correction of the adaptive guess; clipped to the balls below
      // exactly as that guess is.
      if (!g.mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
      warm_features(g, s, set, rq);
      warm_guess(g, s, set);
    }
    k_truncate_guess<<<blocks(n), TPB, 0, st>>>(b.guess64.p, b.centers64.p, b.guess64.p + 3 * (size_t)n,
                                                   d.pinned.p, n, g.d_scal, PBufs<R>::pos(g), &g.d_scal->clipped);
    CV_HIP(hipGetLastError());
    threshold = set->contact.refresh_fraction * n;
    tol = set->residual_tolerance;
    nfree3 = 3.0 * std::max<uint32_t>(s.n_free, 1);
    want_true = tol > 0.0 || (set->flags & CV_FLAG_TRUE_RESIDUAL);
    if (want_true) b.gtrue.reserve(n);
    k_iter_end<<<1, 1, 0, st>>>(g.d_scal, -1, threshold, tol, nfree3);
    CV_HIP(hipGetLastError());
    a = sweep_args<R>(g, set, grab);
    iters = set->iterations;
  }
  // Queues the scalar read-back that after_sync() consumes.
  void read_back() {
    CV_HIP(hipMemcpyAsync(g.h_scal, g.d_scal, sizeof(Scalars), hipMemcpyDeviceToHost, g.stream));
  }
  // After read_back() and a drain: accounting; refresh if asked.  Returns the
  // iteration range [first, end) to queue next, or false when the step's
  // iterations are over.
  bool after_sync(uint32_t& first, uint32_t& end) {
    uint32_t ran = g.h_scal->iters_run;
    double inc = (double)g.h_scal->incidences;
    double it_bytes32 = g.static_bytes32 + 56.0 * inc;
    for (const Launch& l : launches) {
      if (l.it >= ran) continue;
      out.sweep_launches++;
      if (g.profile && l.ev != (size_t)-1) {
        pending_ev.push_back(l.ev);  // read by drain_timing() once more work is queued
        out.sweep_timed_launches++;
      }
    }
    for (uint32_t it : timed_its)
      if (it < ran) out.sweep_timed_bytes_min += it_bytes32;
    launches.clear();
    timed_its.clear();
    out.sweep_bytes += (ran - done) * (sizeof(R) == 4 ? g.static_bytes32own + 92.0 * inc
                                                      : g.static_bytes64 + 100.0 * inc);
    out.sweep_bytes_min += (ran - done) * it_bytes32;
    done = ran;
    if (g.h_scal->stop == 2) return finished = true, false;  // residual tolerance reached
    if (g.h_scal->stop) {
      next = g.h_scal->stop_it;
      if (next >= iters) return finished = true, false;
      // Refresh before iteration `next` at the current positions.
      g.h_scal->stop = 0;
      CV_HIP(hipMemcpyAsync(&g.d_scal->stop, &g.h_scal->stop, sizeof(unsigned int), hipMemcpyHostToDevice,
                            g.stream));
      k_current<<<blocks(s.n), TPB, 0, g.stream>>>(PBufs<R>::pos(g), g.b.x_old64.p, g.dev.pinned.p, s.n,
                                                      g.d_scal, g.b.x64.p);
      CV_HIP(hipGetLastError());
      detect_and_color<R>(g, s, set, grab, rq, out, false);
      a = sweep_args<R>(g, set, grab);
    } else if (next >= iters) {
      return finished = true, false;
    }
    first = next;
    // With a residual tolerance the batch after iteration 0 runs to the
    // iteration count the previous such step needed (a step that has not
    // converged there queues the rest after one more read-back): launches
    // queued past the convergence point would run as no-ops.
    end = next == 0 ? 1 : (tol > 0.0 && g.tol_iters > next && g.tol_iters < iters ? g.tol_iters : iters);
    next = end;
    return true;
  }
  // Sweeps of iterations [first, end) on this solver's own launches.
  void enqueue(uint32_t first, uint32_t end) {
    for (uint32_t it = first; it < end; ++it) {
      // With profiling on, every launch of every kTimedStride-th iteration is
      // stamped with begin/end events (stamping all of them costs more host
      // time per launch than the kernel runs).
      bool timed = g.profile && (g.timed_counter++ % kTimedStride == 0);
      if (timed) timed_its.push_back(it);
      // The iteration's end rides on its last (non-empty) colour's launch,
      // except where the residual tolerance is checked after it (a residual
      // pass comes between the sweep and the decision).
      size_t last = 0;
      for (size_t c = 0; c + 1 < g.color_off.size(); ++c)
        if (g.color_off[c + 1] > g.color_off[c]) last = c;
      const bool fold = !checks(it) && g.color_off.size() > 1 && g.color_off.back() > 0;
      for (size_t c = 0; c + 1 < g.color_off.size(); ++c) {
        uint32_t off = g.color_off[c], cnt = g.color_off[c + 1] - off;
        if (!cnt) continue;
        IterTail tail{};
        if (fold && c == last) tail = IterTail{g.d_scal, (int)it, threshold};
        if (timed) {
          if (g.sweep_events.size() < 2 * (nev + 1)) {
            hipEvent_t e0, e1;
            CV_HIP(hipEventCreate(&e0));
            CV_HIP(hipEventCreate(&e1));
            g.sweep_events.push_back(e0);
            g.sweep_events.push_back(e1);
          }
          sweep_color<R>(g, a, off, cnt, tail, g.sweep_events[2 * nev], g.sweep_events[2 * nev + 1]);
          launches.push_back({nev++, it});
        } else {
          sweep_color<R>(g, a, off, cnt, tail);
          launches.push_back({(size_t)-1, it});
        }
      }
      if (!fold) end_iteration(it);
    }
  }
  // Event times of completed timed launches; called after the next work is
  // queued so the host's event queries overlap the device's work.
  void drain_timing() {
    for (size_t ev : pending_ev) {
      float ms = 0;
      CV_HIP(hipEventElapsedTime(&ms, g.sweep_events[2 * ev], g.sweep_events[2 * ev + 1]));
      out.sweep_ms += ms;
    }
    pending_ev.clear();
  }
  // With a tolerance the residual is evaluated after sweeps k, 2k, ... and
  // the last (settings.residual_interval = k).
  bool checks(uint32_t it) const {
    const uint32_t every = std::max<uint32_t>(set->residual_interval, 1);
    return tol > 0.0 && ((it + 1) % every == 0 || it + 1 == iters);
  }
  void end_iteration(uint32_t it) {
    if (checks(it)) {
      true_residual_fused<R>(g, a, s.n, ResidEnd{1, (int)it, threshold, tol, nfree3});
      return;
    }
    k_iter_end<<<1, 1, 0, g.stream>>>(g.d_scal, (int)it, threshold, 0.0, nfree3);
    CV_HIP(hipGetLastError());
  }
  // Residuals and commit kernels; then read_back() and a drain, then check().
  void finish() {
    DevSolver& d = g.dev;
    StepBufs& b = g.b;
    uint32_t n = s.n;
    out.truncated += g.h_scal->truncated_total;
    out.iterations_run += g.h_scal->iters_run;
    if (want_true) {
      // Post-sweep residual of the final positions.
      true_residual_fused<R>(g, a, n, ResidEnd{0, 0, 0.0, 0.0, 1.0});
      out.true_residual = true;
    }
    sum_f64(g, b.gsq.p, n, &g.d_scal->residual_sum, g.stream);
    // The trust-region check tolerance of the device path (src/gpu_vbd.rs:436).
    k_finalize<<<blocks(n), TPB, 0, g.stream>>>(PBufs<R>::pos(g), b.x_old64.p, b.v_old64.p, b.centers64.p,
                                                   b.radii64.p, d.pinned.p, n, set->dt, set->damping, 1e-8,
                                                   b.xcur64.p, b.v64.p, d.acc.p, g.d_scal);
    CV_HIP(hipGetLastError());
    if (diag && !(set->flags & CV_FLAG_DIAGNOSTICS_OFF)) {
      gpu_diagnostics(g, b.xcur64.p);
      out.diagnostics = true;
    }
  }
  void check() const {
    if (g.h_scal->err_nonfinite) fail(CV_ERR_NONFINITE_RESULT, "Non-finite VBD result; state unchanged");
    if (g.h_scal->err_trust) fail(CV_ERR_TRUST_REGION, "GPU OGC trust-region check failed; state unchanged");
    if (tol > 0.0) g.tol_iters = out.iterations_run;
  }
};

// GpuVbd.check_gradients (src/gpu_vbd.rs:305-379), device half: the probe
// of every vertex's local objective (k_probe, the sweep's own derivative
// code) at the positions and targets the sweep starts from, read back with
// those positions, targets and the detection's contact records for the host
// comparison (check_probe), which throws on a mismatch before anything is
// committed.
template <class R>
void check_gradients_t(cv_gpu& g, const cv_solver& s, const SweepArgs<R>& a, const cv_settings* set,
                       const cv_grab* grab) {
  using T4 = double4;
  const size_t n = s.n, nall = n + g.dev.n_fixed;
  hipStream_t st = g.stream;
  g.probe_out.reserve(4 * n);
  k_probe<32, R><<<blocks(n * 32), TPB, 0, st>>>(a, (uint32_t)n, g.probe_out.p);
  CV_HIP(hipGetLastError());
  std::vector<double> probe(4 * n);
  std::vector<T4> pos(nall), tgt(n);
  std::vector<uint32_t> coff(n + 1);
  CV_HIP(hipMemcpyAsync(probe.data(), g.probe_out.p, probe.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CV_HIP(hipMemcpyAsync(pos.data(), a.pos, nall * sizeof(T4), hipMemcpyDeviceToHost, st));
  CV_HIP(hipMemcpyAsync(tgt.data(), a.tgt, n * sizeof(T4), hipMemcpyDeviceToHost, st));
  CV_HIP(hipMemcpyAsync(coff.data(), a.coff, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CV_HIP(hipStreamSynchronize(st));
  std::vector<cvm::TermRecT<R>> crec(coff[n]);
  if (!crec.empty()) {
    CV_HIP(hipMemcpyAsync(crec.data(), a.crec, crec.size() * sizeof(crec[0]), hipMemcpyDeviceToHost, st));
    CV_HIP(hipStreamSynchronize(st));
  }
  std::vector<double> x(3 * nall), t(4 * n);
  for (size_t i = 0; i < nall; ++i) {
    x[3 * i] = (double)pos[i].x;
    x[3 * i + 1] = (double)pos[i].y;
    x[3 * i + 2] = (double)pos[i].z;
  }
  for (size_t i = 0; i < n; ++i) {
    t[4 * i] = (double)tgt[i].x;
    t[4 * i + 1] = (double)tgt[i].y;
    t[4 * i + 2] = (double)tgt[i].z;
    t[4 * i + 3] = (double)tgt[i].w;
  }
  std::vector<cvm::TermRec64> c64(crec.size());
  for (size_t k = 0; k < crec.size(); ++k) {
    c64[k] = cvm::TermRec64{};
    c64[k].head = crec[k].head;
    for (int q = 0; q < 4; ++q) c64[k].ids[q] = crec[k].ids[q];
    c64[k].f[0] = (double)crec[k].f[0];  // the EE mollifier scale
  }
  g.last_check_max_error =
      check_probe(s, x.data(), t.data(), coff.data(), c64.data(), g.h_scal->origin, set, grab, probe.data());
  std::fprintf(stderr, "GPU/CPU local gradients verified: %zu vertices, maximum absolute error %.3e N\n", n,
               g.last_check_max_error);
}

// carry_in: the previous step of a device-resident run already queued this
// step's start state and prologue reduction and read them back with its own
// final scalars; carry_out: queue them for the next step the same way (one
// host round trip fewer per step).
template <class R>
void run_step_t(cv_gpu& g, cv_solver& s, const cv_settings* set, const cv_grab* grab, StepOut& out, bool diag,
                bool carry_in, bool carry_out) {
  StepCtx<R> c(g, s, set, grab, out);
  c.diag = diag;
  if (!carry_in) {
    c.prologue_reduce();
    CV_HIP(hipStreamSynchronize(g.stream));
  }
  c.begin();
  if (g.check_gradients) check_gradients_t<R>(g, s, c.a, set, grab);
  uint32_t first, end;
  while (true) {
    c.read_back();
    CV_HIP(hipStreamSynchronize(g.stream));
    if (!c.after_sync(first, end)) break;
    c.enqueue(first, end);
    c.drain_timing();
  }
  c.finish();
  c.drain_timing();
  if (carry_out) {
    // The commit is the next step's state: positions into x64 (velocities
    // are in v64), then its origin, motion and scale.
    CV_HIP(hipMemcpyAsync(g.b.x64.p, g.b.xcur64.p, 3 * (size_t)s.n * sizeof(double), hipMemcpyDeviceToDevice,
                          g.stream));
    double dt = set->dt;
    ::prologue_reduce(g, s.n, dt, 1.0 + set->damping * dt, set->gravity);
  }
  read_scalars(g);
  c.check();
}

void run_step(cv_gpu& g, cv_solver& s, const cv_settings* set, const cv_grab* grab, StepOut& out, bool diag,
              bool carry_in = false, bool carry_out = false) {
  if (g.f64)
    run_step_t<double>(g, s, set, grab, out, diag, carry_in, carry_out);
  else
    run_step_t<float>(g, s, set, grab, out, diag, carry_in, carry_out);
}

// Batched step of many solvers (each with its own context g_i and stream):
// per solver the same device work as run_step_t, with the sweeps of one
// colour index of one iteration of every solver in a single launch
// (k_sweep_batch) and the end-of-iteration kernels in another, both on the
// parent's stream.  The per-solver phases (prologue, detection, host
// colouring, refresh, finish) run on host worker threads, each solver on its
// own stream, so detections of different solvers overlap on the device and
// their host colourings on the CPU.  Events order the solver streams against
// the shared launches.  Solvers progress through their own iteration
// indices; a solver waiting for a re-detection sits out.  A solver that
// fails is dropped with its status; the others continue.
template <class R>
void run_batch_t(cv_gpu& P, const std::vector<cv_gpu*>& G, const std::vector<cv_solver*>& S, const cv_settings* set,
                 std::vector<StepOut>& out, std::vector<cv_status>& status, std::vector<std::string>& err) {
  size_t N = G.size();
  hipStream_t st = P.stream;
  std::vector<std::unique_ptr<StepCtx<R>>> C(N);
  std::vector<char> live(N, 1);
  auto drop = [&](size_t i, const CvError& e) {
    live[i] = 0;
    status[i] = e.status;
    err[i] = e.message;
  };
  // Solver streams wait for the parent's work so far.
  auto fork = [&] {
    CV_HIP(hipEventRecord(P.join_ev, st));
  };
  auto wait_parent = [&](size_t i) { CV_HIP(hipStreamWaitEvent(G[i]->stream, P.join_ev, 0)); };
  // The parent stream waits for a solver stream's work so far.
  auto join = [&](size_t i) {
    CV_HIP(hipEventRecord(G[i]->join_ev, G[i]->stream));
    CV_HIP(hipStreamWaitEvent(st, G[i]->join_ev, 0));
  };
  auto on_workers = [&](const std::function<void(size_t)>& f) {
    P.pool->run(N, [&](size_t i) {
      if (!live[i]) return;
      try {
        CV_HIP(hipSetDevice(P.device));
        f(i);
      } catch (const CvError& e) {
        drop(i, e);
      } catch (const std::exception& e) {
        drop(i, CvError{CV_ERR_DEVICE, e.what()});
      }
    });
  };
  fork();
  on_workers([&](size_t i) {
    wait_parent(i);
    C[i] = std::make_unique<StepCtx<R>>(*G[i], *S[i], set, nullptr, out[i]);
    C[i]->prologue_reduce();
    CV_HIP(hipStreamSynchronize(G[i]->stream));
    C[i]->begin();
  });
  struct Range {
    size_t i;
    uint32_t first, end;
  };
  std::vector<SweepArgs<R>> args(N);
  std::vector<BatchSlot> slots;
  std::vector<IterEnd> ends;
  struct Launch {
    size_t slot0, nslots;
    uint32_t total;
  };
  std::vector<char> ready(N);
  std::vector<uint32_t> rf(N), re(N);
  bool swept = false;
  while (true) {
    std::fill(ready.begin(), ready.end(), 0);
    on_workers([&](size_t i) {
      if (C[i]->finished) return;
      if (swept) wait_parent(i);
      C[i]->read_back();
      CV_HIP(hipStreamSynchronize(G[i]->stream));
      uint32_t f = 0, e = 0;
      if (C[i]->after_sync(f, e)) {
        rf[i] = f;
        re[i] = e;
        ready[i] = 1;
      }
    });
    std::vector<Range> rr;
    for (size_t i = 0; i < N; ++i)
      if (live[i] && ready[i]) rr.push_back({i, rf[i], re[i]});
    if (rr.empty()) break;
    for (auto& r : rr) {
      args[r.i] = C[r.i]->a;
      join(r.i);
    }
    uint32_t T = 0;
    for (auto& r : rr) T = std::max(T, r.end - r.first);
    slots.clear();
    ends.clear();
    std::vector<std::vector<Launch>> per_t(T);
    std::vector<std::pair<size_t, size_t>> end_range(T);
    for (uint32_t t = 0; t < T; ++t) {
      size_t maxc = 0;
      for (auto& r : rr)
        if (r.first + t < r.end) maxc = std::max(maxc, G[r.i]->color_off.size() - 1);
      for (size_t c = 0; c < maxc; ++c) {
        size_t s0 = slots.size();
        uint32_t total = 0;
        for (auto& r : rr) {
          const cv_gpu& gi = *G[r.i];
          if (r.first + t >= r.end || c + 1 >= gi.color_off.size()) continue;
          uint32_t off = gi.color_off[c], cnt = gi.color_off[c + 1] - off;
          if (!cnt) continue;
          slots.push_back({total, off, cnt, (uint32_t)r.i});
          total += cnt;
          C[r.i]->launches.push_back({(size_t)-1, r.first + t});
        }
        if (total) per_t[t].push_back({s0, slots.size() - s0, total});
      }
      size_t e0 = ends.size();
      for (auto& r : rr)
        if (r.first + t < r.end)
          ends.push_back({G[r.i]->d_scal, (int)(r.first + t), C[r.i]->threshold,
                          C[r.i]->checks(r.first + t) ? C[r.i]->tol : 0.0, C[r.i]->nfree3});
      end_range[t] = {e0, ends.size() - e0};
    }
    P.batch_args.upload((const unsigned char*)args.data(), N * sizeof(SweepArgs<R>), st);
    P.batch_slots.upload((const unsigned char*)slots.data(), std::max<size_t>(slots.size(), 1) * sizeof(BatchSlot),
                         st);
    P.batch_ends.upload((const unsigned char*)ends.data(), std::max<size_t>(ends.size(), 1) * sizeof(IterEnd), st);
    const SweepArgs<R>* dargs = (const SweepArgs<R>*)P.batch_args.p;
    const BatchSlot* dslots = (const BatchSlot*)P.batch_slots.p;
    const IterEnd* dends = (const IterEnd*)P.batch_ends.p;
    for (uint32_t t = 0; t < T; ++t) {
      for (const Launch& l : per_t[t]) {
        const BatchSlot* sl = dslots + l.slot0;
        uint32_t ns = (uint32_t)l.nslots;
        // Threads per block: the tuned batch block size (64, 128 or 256); the
        // heavy blocks keep kHeavyBlocks * TPB / 64 heavy wavefronts.
        const int B = P.batch_block;
        const uint32_t hb = kHeavyBlocks * (TPB / B);
        switch (P.lanes) {
          case 8: k_sweep_batch<8, R><<<blocks((uint64_t)l.total * 8, B), B, 0, st>>>(dargs, sl, ns, l.total, 0); break;
          case 16:
            k_sweep_batch<16, R><<<blocks((uint64_t)l.total * 16, B), B, 0, st>>>(dargs, sl, ns, l.total, 0);
            break;
          case 64:
            k_sweep_batch<64, R><<<blocks((uint64_t)l.total * 64, B), B, 0, st>>>(dargs, sl, ns, l.total, 0);
            break;
          default:  // every instance context sweeps with 32 lanes and holds its heavy lists
            k_sweep_batch<32, R><<<blocks((uint64_t)l.total * 32, B) + hb, B, 0, st>>>(dargs, sl, ns, l.total, hb);
            break;
        }
      }
      for (auto& r : rr)
        if (r.first + t < r.end && C[r.i]->checks(r.first + t)) true_residual<R>(*G[r.i], C[r.i]->a, S[r.i]->n, st);
      auto [e0, ne] = end_range[t];
      if (ne) k_iter_end_batch<<<blocks(ne), TPB, 0, st>>>(dends + e0, (uint32_t)ne);
      CV_HIP(hipGetLastError());
    }
    fork();
    swept = true;
  }
  on_workers([&](size_t i) {
    if (swept) wait_parent(i);
    C[i]->finish();
    C[i]->read_back();
    CV_HIP(hipStreamSynchronize(G[i]->stream));
    C[i]->check();
  });
}

}  // namespace

void ensure_device_solver(cv_gpu& g, const cv_solver& s) {
  DevSolver& d = g.dev;
  if (d.generation == s.generation) return;
  hipStream_t st = g.stream;
  const OgcTopology& t = s.top;
  d.fc_ready = false;
  g.detected = false;
  d.n = s.n;
  d.n_total = t.n_total;
  d.n_fixed = t.n_total - s.n;
  d.F = (uint32_t)(t.faces.size() / 3);
  d.E = (uint32_t)(t.edges.size() / 2);
  d.term_off.upload(s.term_off, st);
  d.terms.upload(s.terms, st);
  d.terms_d.upload(s.terms64, st);
  d.masses.upload(s.masses, st);
  d.pinned.upload(s.pinned, st);
  std::vector<uint32_t> ft(t.faces);
  ft.insert(ft.end(), t.tri_edges.begin(), t.tri_edges.end());
  d.faces.upload(ft, st);
  d.edges.upload(t.edges, st);
  d.opposites.upload(t.opposites, st);
  d.nb_off.upload(t.nb_off, st);
  std::vector<uint32_t> nb(t.nb);
  if (nb.empty()) nb.push_back(0);
  d.nb.upload(nb, st);
  d.ex_off.upload(t.ex_off, st);
  std::vector<uint32_t> ex(t.ex);
  if (ex.empty()) ex.push_back(0);
  d.ex.upload(ex, st);
  d.face_moving.upload(t.face_moving, st);
  d.edge_moving.upload(t.edge_moving, st);
  d.rest_len2.upload(t.rest_len2, st);
  std::vector<double> fx(s.avatar);
  if (fx.empty()) fx.push_back(0);
  d.fixed_pos.upload(fx, st);
  d.fixed_abs_max = 0;
  for (double v : s.avatar) d.fixed_abs_max = std::max(d.fixed_abs_max, std::fabs(v));
  d.cell = s.face_extent_median;
  {
    // Garment diagnostics inputs (grain rotation per triangle as the host
    // formula computes it).
    d.T = (uint32_t)s.tri_area.size();
    d.P = s.pr_off.empty() ? 0u : (uint32_t)(s.pr_off.size() - 1);
    std::vector<double> cs(2 * (size_t)std::max<uint32_t>(d.T, 1), 0.0);
    for (uint32_t t = 0; t < d.T; ++t) {
      const cv_material& m = s.materials[s.tri_mat[t]];
      cs[2 * t] = std::cos(m.grain_degrees * (M_PI / 180.0));
      cs[2 * t + 1] = std::sin(m.grain_degrees * (M_PI / 180.0));
    }
    auto nz_u = [](std::vector<uint32_t> v) { if (v.empty()) v.push_back(0); return v; };
    auto nz_d = [](std::vector<double> v) { if (v.empty()) v.push_back(0); return v; };
    d.dg_tri.upload(nz_u(s.tri_v), st);
    d.dg_inv.upload(nz_d(s.tri_inv_dm), st);
    d.dg_cs.upload(cs, st);
    d.dg_area.upload(nz_d(s.tri_area), st);
    d.dg_pr_off.upload(nz_u(s.pr_off), st);
    d.dg_pr_v.upload(nz_u(s.pr_v), st);
    d.dg_pr_w.upload(nz_d(s.pr_w), st);
  }
  // Algorithmic bytes of one vertex-iteration (each term's data and the
  // neighbour positions read once, position written once):
  //   f32 minimum (f32 positions and records): 64 fixed + 72 per triangle,
  //     52 per hinge, 68 per stitch (56 per contact incidence);
  //   f32 sweep as stored (f64 positions, f32 records): 120 fixed + 96 per
  //     triangle, 88 per hinge, 104 per stitch (92 per contact);
  //   f64: 120 fixed + 136 per triangle, 92 per hinge, 124 per stitch (100).
  // A pinned vertex reads its position only.
  g.static_bytes32 = g.static_bytes32own = g.static_bytes64 = 0;
  for (uint32_t v = 0; v < s.n; ++v) {
    if (s.pinned[v]) {
      g.static_bytes32 += 16;
      g.static_bytes32own += 32;
      g.static_bytes64 += 32;
      continue;
    }
    double b32 = 64, b32own = 120, b64 = 120;
    for (uint32_t k = s.term_off[v]; k < s.term_off[v + 1]; ++k) {
      uint32_t kind = cvm::rec_kind(s.terms[k].head);
      b32 += kind == cvm::TERM_MEMBRANE ? 72 : (kind == cvm::TERM_HINGE ? 52 : 68);
      b32own += kind == cvm::TERM_MEMBRANE ? 96 : (kind == cvm::TERM_HINGE ? 88 : 104);
      b64 += kind == cvm::TERM_MEMBRANE ? 136 : (kind == cvm::TERM_HINGE ? 92 : 124);
    }
    g.static_bytes32 += b32;
    g.static_bytes32own += b32own;
    g.static_bytes64 += b64;
  }
  d.acc.reserve(3 * (size_t)s.n);
  d.acc_version = 0;
  d.acc_device = false;
  d.generation = s.generation;
  CV_HIP(hipStreamSynchronize(st));
}

void ensure_step_buffers(cv_gpu& g, const cv_solver& s) {
  StepBufs& b = g.b;
  size_t n = s.n, nt = s.top.n_total;
  b.x64.reserve(3 * nt);
  b.x_old64.reserve(3 * n);
  b.v64.reserve(3 * n);
  b.v_old64.reserve(3 * n);
  b.guess64.reserve(4 * n);  // guess (3n) + shrunk radii (n)
  b.xcur64.reserve(3 * n);
  b.radii64.reserve(n);
  b.centers64.reserve(3 * n);
  b.pos_d.reserve(nt);
  b.tgt_d.reserve(n);
  b.ball_d.reserve(n);
  b.gsq.reserve(n);
  b.order.reserve(n);
  size_t cap = std::max<size_t>(1 << 20, 32 * n);
  if (b.contact_cap < cap) {
    b.raw_ids.reserve(5 * cap);
    b.raw_count.reserve(cap);
    b.raw_eps.reserve(cap);
    b.c_ids.reserve(5 * cap);
    b.c_count.reserve(cap);
    b.c_eps.reserve(cap);
    b.contact_cap = cap;
  }
  b.coff.reserve(n + 1);
  if (g.h_cap < 3 * n) {
    if (g.h_x) CV_HIP(hipHostFree(g.h_x));
    if (g.h_v) CV_HIP(hipHostFree(g.h_v));
    if (g.h_acc) CV_HIP(hipHostFree(g.h_acc));
    CV_HIP(hipHostMalloc((void**)&g.h_x, 3 * n * sizeof(double)));
    CV_HIP(hipHostMalloc((void**)&g.h_v, 3 * n * sizeof(double)));
    CV_HIP(hipHostMalloc((void**)&g.h_acc, 3 * n * sizeof(double)));
    g.h_cap = 3 * n;
  }
}

// ------------------------------------------------------------------- C ABI
extern thread_local std::string cv_create_error;  // api.cpp

namespace {


void fill_metrics(cv_metrics* m, const cv_gpu& g, const cv_solver& s, const cv_settings* set,
                  const StepOut& o, double step_ms, double gpu_ms, const double* x_new) {
  if (!m) return;
  *m = cv_metrics{};
  m->dt_s = set->dt;
  m->step_ms = step_ms;
  m->detection_ms = o.detection_ms;
  m->solve_ms = step_ms - o.detection_ms;
  m->gpu_compute_ms = gpu_ms;
  m->refit_ms = 0;
  m->candidates_ms = 0;
  m->narrow_ms = o.narrow_ms;
  m->candidate_pairs = o.candidate_pairs;
  m->aabb_rejections = o.grid_tests - std::min(o.grid_tests, o.candidate_pairs);
  m->iterations = set->iterations;
  m->contacts = g.h_scal->active;
  m->contact_terms = g.h_scal->contacts;
  m->detection_updates = o.detections;
  m->broad_phase_rebuilds = o.detections;
  m->truncated_vertices = o.truncated;
  m->minimum_distance_m = g.h_scal->min_distance;
  m->residual_rms_n = std::sqrt(g.h_scal->residual_sum / (3.0 * s.n));
  m->true_residual_rms_n =
      o.true_residual ? std::sqrt(g.h_scal->true_residual_sum / (3.0 * std::max<uint32_t>(s.n_free, 1))) : NAN;
  m->colors = o.colors;
  m->iterations_run = o.iterations_run;
  m->sweep_ms = g.profile ? o.sweep_ms : NAN;
  m->sweep_timed_launches = o.sweep_timed_launches;
  m->sweep_timed_bytes_min = o.sweep_timed_bytes_min;
  m->sweep_event_overhead_ms = g.profile ? g.event_overhead_ms : NAN;
  m->broad_phase_fallbacks = o.bvh_fallbacks;
  m->f64 = g.f64 ? 1 : 0;
  m->sweep_launches = o.sweep_launches;
  m->sweep_bytes = o.sweep_bytes;
  m->sweep_bytes_min = o.sweep_bytes_min;
  m->steps = 1;
  if (o.diagnostics) {
    // Computed on the device from the committed positions (gpu_diagnostics).
    m->max_seam_gap_m = __builtin_bit_cast(double, (uint64_t)g.h_scal->diag_gap_bits);
    m->extension_p95_percent = g.h_scal->diag_p95;
    m->extension_max_percent = g.h_scal->diag_max;
  } else {
    m->max_seam_gap_m = m->extension_p95_percent = m->extension_max_percent = NAN;
  }
}

// Upload the state and the prediction history for one step.
void stage_state(cv_gpu& g, cv_solver& s, const cv_state* st, const cv_settings* set) {
  size_t n = s.n;
  std::memcpy(g.h_x, st->x, 3 * n * sizeof(double));
  std::memcpy(g.h_v, st->v, 3 * n * sizeof(double));
  CV_HIP(hipMemcpyAsync(g.b.x64.p, g.h_x, 3 * n * sizeof(double), hipMemcpyHostToDevice, g.stream));
  CV_HIP(hipMemcpyAsync(g.b.v64.p, g.h_v, 3 * n * sizeof(double), hipMemcpyHostToDevice, g.stream));
  if (!s.avatar.empty())
    CV_HIP(hipMemcpyAsync(g.b.x64.p + 3 * n, s.avatar.data(), s.avatar.size() * sizeof(double),
                          hipMemcpyHostToDevice, g.stream));
  bool valid = prediction_valid(s, st, set);
  if (valid && g.dev.acc_version != s.acc_version) {
    if (s.acc.size() == 3 * n) {
      CV_HIP(hipMemcpyAsync(g.dev.acc.p, s.acc.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice,
                            g.stream));
      g.dev.acc_version = s.acc_version;
    } else {
      valid = false;
    }
  }
  g.pred_valid = valid;
}

uint64_t next_acc_version() {
  static std::atomic<uint64_t> v{1ull << 40};
  return ++v;
}

template <class R>
void probe_t(cv_gpu& g, cv_solver& s, const cv_state* st, const cv_settings* set, const cv_grab* grab,
             double* out) {
  size_t n = s.n;
  hipStream_t stream = g.stream;
  double dt = set->dt;
  k_prologue<<<blocks(n + g.dev.n_fixed), TPB, 0, stream>>>(
      g.b.x64.p, g.b.v64.p, g.dev.acc.p, g.dev.masses.p, g.dev.pinned.p, (uint32_t)n, g.dev.fixed_pos.p,
      g.dev.n_fixed, dt, set->damping, set->gravity[0], set->gravity[1], set->gravity[2], 0, g.d_scal,
      PBufs<R>::pos(g), PBufs<R>::tgt(g), g.b.guess64.p, g.b.x_old64.p, g.b.v_old64.p);
  double r = set->contact.radius;
  double rq = std::min(std::max(r + g.h_scal->motion, 2.0 * r), 4.0 * r);
  g.b.bvh_rebuild = true;
  gpu_detect(g, s, rq, set->contact, 0);
  // Probe at the state's positions (relative to the origin).
  std::vector<double4> pos(n);
  for (size_t i = 0; i < n; ++i)
    pos[i] = make_double4(st->x[3 * i] - g.h_scal->origin[0], st->x[3 * i + 1] - g.h_scal->origin[1],
                          st->x[3 * i + 2] - g.h_scal->origin[2], s.pinned[i] ? 1.0 : 0.0);
  CV_HIP(hipMemcpyAsync(PBufs<R>::pos(g), pos.data(), n * sizeof(double4), hipMemcpyHostToDevice, stream));
  SweepArgs<R> a = sweep_args<R>(g, set, grab);
  DBuf<double> outd;
  outd.reserve(4 * n);
  k_probe<32, R><<<blocks(n * 32), TPB, 0, stream>>>(a, (uint32_t)n, outd.p);
  CV_HIP(hipGetLastError());
  CV_HIP(hipMemcpyAsync(out, outd.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost, stream));
  CV_HIP(hipStreamSynchronize(stream));
}

// The part of a launch's event interval (hipExtLaunchKernelGGL start/stop
// events on the stream) that is not the kernel: the mean interval of
// back-to-back empty launches on the same stream, after one untimed launch.
__global__ void k_empty() {}
void calibrate_event_overhead(cv_gpu& g) {
  CV_HIP(hipSetDevice(g.device));
  constexpr int K = 32;
  std::vector<hipEvent_t> ev(2 * K);
  for (auto& e : ev) CV_HIP(hipEventCreate(&e));
  hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, g.stream);
  for (int k = 0; k < K; ++k)
    hipExtLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, g.stream, ev[2 * k], ev[2 * k + 1], 0);
  CV_HIP(hipStreamSynchronize(g.stream));
  double sum = 0;
  for (int k = 0; k < K; ++k) {
    float ms = 0;
    CV_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
    sum += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  g.event_overhead_ms = sum / K;
}

}  // namespace

extern "C" {

cv_status cv_gpu_create(int device, cv_gpu** out) {
  if (!out) return CV_ERR_INVALID_ARGUMENT;
  std::string err;
  cv_gpu* g = new cv_gpu();
  cv_status st = guard(&err, [&] {
    int count = 0;
    CV_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(CV_ERR_DEVICE, "no such HIP device");
    g->device = device;
    CV_HIP(hipSetDevice(device));
    warm_scratch_setup(device);

    CV_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    CV_HIP(hipMalloc((void**)&g->d_scal, sizeof(Scalars)));
    CV_HIP(hipHostMalloc((void**)&g->h_scal, sizeof(Scalars)));
    std::memset(g->h_scal, 0, sizeof(Scalars));
    CV_HIP(hipEventCreate(&g->ev0));
    CV_HIP(hipEventCreate(&g->ev1));
  });
  if (st != CV_OK) {
    cv_create_error = err;
    delete g;
    return st;
  }
  *out = g;
  return CV_OK;
}

void cv_gpu_destroy(cv_gpu* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  for (cv_gpu* c : g->subs) cv_gpu_destroy(c);
  g->subs.clear();
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->d_scal) (void)hipFree(g->d_scal);
  if (g->h_scal) (void)hipHostFree(g->h_scal);
  if (g->h_x) (void)hipHostFree(g->h_x);
  if (g->h_v) (void)hipHostFree(g->h_v);
  if (g->h_acc) (void)hipHostFree(g->h_acc);
  if (g->h_cont) (void)hipHostFree(g->h_cont);
  if (g->h_cchg) (void)hipHostFree(g->h_cchg);
  if (g->h_fc) (void)hipHostFree(g->h_fc);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->join_ev) (void)hipEventDestroy(g->join_ev);
  hipStream_t st = g->stream;
  delete g;
  if (st) (void)hipStreamDestroy(st);
}

cv_status cv_gpu_set_check_gradients(cv_gpu* g, int enabled) {
  if (!g) return CV_ERR_INVALID_ARGUMENT;
  g->check_gradients = enabled != 0;
  return CV_OK;
}

cv_status cv_warm_set_smoothing(cv_gpu* g, int passes) {
  if (!g || passes < 0 || passes > 16) return CV_ERR_INVALID_ARGUMENT;
  g->warm_smooth = passes;
  return CV_OK;
}

cv_status cv_warm_set_correction(cv_gpu* g, double scale, int normal_only) {
  if (!g || !(scale >= 0.0 && scale <= 4.0)) return CV_ERR_INVALID_ARGUMENT;
  g->warm_scale = scale;
  g->warm_normal_only = normal_only != 0;
  return CV_OK;
}

cv_status cv_gpu_set_broad_phase(cv_gpu* g, int kind) {
  if (!g || (kind != CV_BROAD_PHASE_GRID && kind != CV_BROAD_PHASE_LBVH)) return CV_ERR_INVALID_ARGUMENT;
  g->broad_phase = kind;
  g->b.bvh_rebuild = true;
  return CV_OK;
}

cv_status cv_gpu_set_lanes(cv_gpu* g, int lanes) {
  if (!g || !(lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64)) return CV_ERR_INVALID_ARGUMENT;
  g->lanes = lanes;
  return CV_OK;
}

cv_status cv_gpu_set_batch_block(cv_gpu* g, int threads) {
  if (!g || !(threads == 64 || threads == 128 || threads == 256)) return CV_ERR_INVALID_ARGUMENT;
  g->batch_block = threads;
  return CV_OK;
}

cv_status cv_gpu_set_precision(cv_gpu* g, int bits) {
  if (!g || (bits != 32 && bits != 64)) return CV_ERR_INVALID_ARGUMENT;
  g->f64 = bits == 64;
  return CV_OK;
}

cv_status cv_gpu_set_profiling(cv_gpu* g, int enabled) {
  if (!g) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    const bool was = g->profile;
    g->profile = enabled != 0;
    if (g->profile) {
      calibrate_event_overhead(*g);
    } else if (was) {
      // An empty launch closes the profiled window, as the calibration's
      // empty launches open it, so a kernel trace can be cut to the window.
      CV_HIP(hipSetDevice(g->device));
      hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, g->stream);
      CV_HIP(hipGetLastError());
      CV_HIP(hipStreamSynchronize(g->stream));
    }
  });
}

// Steps `steps` times; the state crosses PCIe once each way.  Each step has
// the semantics of GpuVbd::step; on error the caller's state is untouched.
static cv_status run_steps(cv_gpu* g, cv_solver* s, cv_state* st, const cv_settings* set, const cv_grab* grab,
                           uint32_t steps, cv_metrics* metrics) {
  if (!g || !s) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    double t0 = now_ms();
    validate_settings(set);
    validate_state(*s, st);
    validate_grab(*s, grab);
    ensure_device_solver(*g, *s);
    ensure_step_buffers(*g, *s);
    stage_state(*g, *s, st, set);
    size_t n = s->n;
    StepOut o, tot;
    float gpu_ms = 0;
    for (uint32_t k = 0; k < steps; ++k) {
      o = StepOut{};
      if (k > 0) {
        // The previous step's commit is the new state (the previous step
        // queued it into x64 with this step's prologue reduction);
        // velocities are in v64, the acceleration history on the device.
        g->pred_valid = true;
        t0 = now_ms();
      } else {
        CV_HIP(hipEventRecord(g->ev0, g->stream));
      }
      run_step(*g, *s, set, grab, o, k + 1 == steps, k > 0, k + 1 < steps);
      tot.detections += o.detections;
      tot.bvh_fallbacks += o.bvh_fallbacks;
      tot.truncated += o.truncated;
      tot.iterations_run += o.iterations_run;
      tot.sweep_launches += o.sweep_launches;
      tot.sweep_ms += o.sweep_ms;
      tot.sweep_timed_launches += o.sweep_timed_launches;
      tot.sweep_timed_bytes_min += o.sweep_timed_bytes_min;
      tot.sweep_bytes += o.sweep_bytes;
      tot.sweep_bytes_min += o.sweep_bytes_min;
    }
    CV_HIP(hipEventRecord(g->ev1, g->stream));
    CV_HIP(hipMemcpyAsync(g->h_x, g->b.xcur64.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipMemcpyAsync(g->h_v, g->b.v64.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    // The acceleration history travels with the solver (Prediction::record,
    // src/vbd/prediction.rs:43-54), so another context can continue it.
    CV_HIP(hipMemcpyAsync(g->h_acc, g->dev.acc.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
    CV_HIP(hipEventElapsedTime(&gpu_ms, g->ev0, g->ev1));
    // Commit (all checks passed).
    s->acc.assign(g->h_acc, g->h_acc + 3 * n);
    std::memcpy(st->x, g->h_x, 3 * n * sizeof(double));
    std::memcpy(st->v, g->h_v, 3 * n * sizeof(double));
    st->time += steps * set->dt;
    s->has_last = true;
    s->last_x.assign(st->x, st->x + 3 * n);
    s->last_v.assign(st->v, st->v + 3 * n);
    s->last_time = st->time;
    for (int k = 0; k < 3; ++k) s->last_gravity[k] = set->gravity[k];
    s->last_damping = set->damping;
    s->acc_version = next_acc_version();
    g->dev.acc_version = s->acc_version;
    double step_ms = now_ms() - t0;
    if (steps > 1) {
      // Run totals (see cv_gpu_run).
      o.detections = tot.detections;
      o.bvh_fallbacks = tot.bvh_fallbacks;
      o.truncated = tot.truncated;
      o.iterations_run = tot.iterations_run;
      o.sweep_launches = tot.sweep_launches;
      o.sweep_ms = tot.sweep_ms;
      o.sweep_timed_launches = tot.sweep_timed_launches;
      o.sweep_timed_bytes_min = tot.sweep_timed_bytes_min;
      o.sweep_bytes = tot.sweep_bytes;
      o.sweep_bytes_min = tot.sweep_bytes_min;
    }
    fill_metrics(metrics, *g, *s, set, o, step_ms, gpu_ms, st->x);
    if (metrics) {
      metrics->time_s = st->time;
      metrics->steps = steps;
    }
  });
}

cv_status cv_gpu_step(cv_gpu* g, cv_solver* s, cv_state* st, const cv_settings* set, const cv_grab* grab,
                      cv_metrics* metrics) {
  return run_steps(g, s, st, set, grab, 1, metrics);
}

cv_status cv_gpu_run(cv_gpu* g, cv_solver* s, cv_state* st, const cv_settings* set, uint32_t steps,
                     cv_metrics* last) {
  if (steps == 0) return CV_OK;
  return run_steps(g, s, st, set, nullptr, steps, last);
}

cv_status cv_gpu_probe(cv_gpu* g, cv_solver* s, const cv_state* st, const cv_settings* set,
                       const cv_grab* grab, double* out) {
  if (!g || !s || !out) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    validate_settings(set);
    validate_state(*s, st);
    validate_grab(*s, grab);
    ensure_device_solver(*g, *s);
    ensure_step_buffers(*g, *s);
    stage_state(*g, *s, st, set);
    g->pred_valid = false;
    size_t n = s->n;
    double dt = set->dt, damp = 1.0 + set->damping * dt;
    prologue_reduce(*g, (uint32_t)n, dt, damp, set->gravity);
    read_scalars(*g);
    if (g->f64)
      probe_t<double>(*g, *s, st, set, grab, out);
    else
      probe_t<float>(*g, *s, st, set, grab, out);
  });
}

cv_status cv_gpu_colors(cv_gpu* g, cv_solver* s, const cv_grab* grab, uint32_t flags, uint32_t* color_of_vertex,
                        uint32_t n, uint32_t* n_colors) {
  if (!g || !s || !color_of_vertex || n != s->n) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    const bool fast = (flags & CV_FLAG_FAST_COLORING) != 0;
    if (g->dev.generation != s->generation || !g->detected || (!fast && !g->h_cont_valid))
      fail(CV_ERR_INVALID_ARGUMENT, "cv_gpu_colors needs a cv_gpu_detect of this solver first");
    validate_grab(*s, grab);
    uint32_t k = fast ? color_on_device(*g, *s, grab) : color_on_host(*g, *s, grab);
    std::vector<uint32_t> order(n);
    CV_HIP(hipMemcpyAsync(order.data(), g->b.order.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
    for (uint32_t c = 0; c < k; ++c)
      for (uint32_t i = g->color_off[c]; i < g->color_off[c + 1]; ++i) color_of_vertex[order[i]] = c;
    if (n_colors) *n_colors = k;
  });
}

cv_status cv_gpu_detect(cv_gpu* g, cv_solver* s, const double* x, double rq, const cv_ogc_settings* ogc,
                        cv_detection* out, uint32_t* vf_pairs, uint64_t vf_capacity, uint32_t* ee_pairs,
                        uint64_t ee_capacity, uint32_t* contact_kind_ids, uint64_t contact_capacity,
                        double* contact_eps, double* radii) {
  if (!g || !s || !x || !ogc || !out) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    const cv_ogc_settings& c = *ogc;
    if (!std::isfinite(c.radius) || c.radius <= 0.0 || !std::isfinite(c.stiffness) || c.stiffness <= 0.0 ||
        !std::isfinite(c.relaxation) || c.relaxation <= 0.0 || c.relaxation >= 0.5 ||
        !std::isfinite(c.refresh_fraction) || !(c.refresh_fraction >= 0.0 && c.refresh_fraction <= 1.0))
      fail(CV_ERR_INVALID_OGC_SETTINGS, "Invalid OGC radius, stiffness or trust-region parameters");
    if (!std::isfinite(rq) || rq < c.radius)
      fail(CV_ERR_QUERY_RADIUS, "OGC query radius must cover the contact radius");
    ensure_device_solver(*g, *s);
    ensure_step_buffers(*g, *s);
    size_t n = s->n;
    for (size_t i = 0; i < 3 * n; ++i)
      if (!std::isfinite(x[i])) fail(CV_ERR_INVALID_POSITIONS, "OGC invalid position array");
    CV_HIP(hipMemcpyAsync(g->b.x64.p, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, g->stream));
    if (!s->avatar.empty())
      CV_HIP(hipMemcpyAsync(g->b.x64.p + 3 * n, s->avatar.data(), s->avatar.size() * sizeof(double),
                            hipMemcpyHostToDevice, g->stream));
    double mabs = 0;
    for (size_t i = 0; i < 3 * n; ++i) mabs = std::max(mabs, std::fabs(x[i]));
    g->h_scal->scale = mabs;
    size_t cap = std::max(vf_capacity, ee_capacity);
    g->b.bvh_rebuild = true;
    gpu_detect(*g, *s, rq, c, (vf_pairs || ee_pairs) ? std::max<size_t>(cap, 1) : 0);
    const Scalars& hs = *g->h_scal;
    out->n_vf_candidates = hs.vf_cand;
    out->n_ee_candidates = hs.ee_cand;
    out->vertex_face_candidates = hs.vf_count;
    out->edge_edge_candidates = hs.ee_count;
    out->n_contacts = hs.contacts;
    out->active_contacts = hs.active;
    out->minimum_distance = hs.min_distance;
    if (vf_pairs && hs.cand_export_vf <= vf_capacity)
      CV_HIP(hipMemcpyAsync(vf_pairs, g->b.cand_vf.p, 2 * hs.cand_export_vf * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, g->stream));
    if (ee_pairs && hs.cand_export_ee <= ee_capacity)
      CV_HIP(hipMemcpyAsync(ee_pairs, g->b.cand_ee.p, 2 * hs.cand_export_ee * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, g->stream));
    if (contact_kind_ids && hs.contacts <= contact_capacity && hs.contacts)
      CV_HIP(hipMemcpyAsync(contact_kind_ids, g->b.c_ids.p, 5 * hs.contacts * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, g->stream));
    if (contact_eps && hs.contacts <= contact_capacity && hs.contacts)
      CV_HIP(hipMemcpyAsync(contact_eps, g->b.c_eps.p, hs.contacts * sizeof(double), hipMemcpyDeviceToHost,
                            g->stream));
    if (radii)
      CV_HIP(hipMemcpyAsync(radii, g->b.radii64.p, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}

cv_status cv_gpu_step_batch(cv_gpu* g, cv_solver* const* solvers, cv_state* const* states, uint32_t count,
                            const cv_settings* settings, cv_metrics* metrics) {
  if (!g || (count && (!solvers || !states))) return CV_ERR_INVALID_ARGUMENT;
  if (count == 0) return CV_OK;
  std::vector<cv_status> status(count, CV_OK);
  std::vector<std::string> err(count);
  cv_status outer = guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    validate_settings(settings);
    if (settings->flags & CV_FLAG_WARM_START)
      fail(CV_ERR_INVALID_SETTINGS, "the learned warm start is not available in batched steps");
    for (uint32_t i = 0; i < count; ++i) {
      if (!solvers[i] || !states[i]) fail(CV_ERR_INVALID_ARGUMENT, "null solver or state in batch");
      for (uint32_t j = 0; j < i; ++j)
        if (solvers[j] == solvers[i]) fail(CV_ERR_INVALID_ARGUMENT, "a solver appears twice in one batch");
    }
    // One context per batch slot, on this context's device, each with its
    // own stream; host worker threads for the per-solver phases.
    if (!g->join_ev) CV_HIP(hipEventCreateWithFlags(&g->join_ev, hipEventDisableTiming));
    if (!g->pool) {
      // Host threads for the per-solver phases: CLOTHVBD_HOST_THREADS, else
      // min(8, hardware threads); the calling thread is one of them.
      unsigned hw = std::max(1u, std::thread::hardware_concurrency());
      unsigned want = std::min(8u, hw);
      if (const char* e = std::getenv("CLOTHVBD_HOST_THREADS")) {
        long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 256) want = (unsigned)v;
      }
      g->pool = std::make_unique<WorkPool>(want - 1);
    }
    while (g->subs.size() < count) {
      cv_gpu* c = new cv_gpu();
      c->device = g->device;
      g->subs.push_back(c);
      CV_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      CV_HIP(hipMalloc((void**)&c->d_scal, sizeof(Scalars)));
      CV_HIP(hipHostMalloc((void**)&c->h_scal, sizeof(Scalars)));
      std::memset(c->h_scal, 0, sizeof(Scalars));
      CV_HIP(hipEventCreate(&c->ev0));
      CV_HIP(hipEventCreate(&c->ev1));
      CV_HIP(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
    }
    std::vector<double> t0(count);
    auto each_slot = [&](const std::function<void(uint32_t)>& f) {
      g->pool->run(count, [&](size_t i) {
        if (status[i] != CV_OK) return;
        status[i] = guard(&err[i], [&] {
          CV_HIP(hipSetDevice(g->device));
          f((uint32_t)i);
        });
      });
    };
    // Stage every state on its solver's stream (parallel host copies).
    for (uint32_t i = 0; i < count; ++i) {
      cv_gpu& c = *g->subs[i];
      c.lanes = g->lanes;
      c.f64 = g->f64;
      c.profile = false;
    }
    each_slot([&](uint32_t i) {
      cv_gpu& c = *g->subs[i];
      t0[i] = now_ms();
      validate_state(*solvers[i], states[i]);
      ensure_device_solver(c, *solvers[i]);
      ensure_step_buffers(c, *solvers[i]);
      stage_state(c, *solvers[i], states[i], settings);
      CV_HIP(hipStreamSynchronize(c.stream));
    });
    std::vector<cv_gpu*> G;
    std::vector<cv_solver*> S;
    std::vector<uint32_t> idx;
    for (uint32_t i = 0; i < count; ++i)
      if (status[i] == CV_OK) {
        G.push_back(g->subs[i]);
        S.push_back(solvers[i]);
        idx.push_back(i);
      }
    std::vector<StepOut> out(G.size());
    std::vector<cv_status> st2(G.size(), CV_OK);
    std::vector<std::string> er2(G.size());
    if (!G.empty()) {
      CV_HIP(hipEventRecord(g->ev0, g->stream));
      if (g->f64)
        run_batch_t<double>(*g, G, S, settings, out, st2, er2);
      else
        run_batch_t<float>(*g, G, S, settings, out, st2, er2);
      // Every solver stream's work precedes the end event.
      for (cv_gpu* c : G) {
        CV_HIP(hipEventRecord(c->join_ev, c->stream));
        CV_HIP(hipStreamWaitEvent(g->stream, c->join_ev, 0));
      }
      CV_HIP(hipEventRecord(g->ev1, g->stream));
      CV_HIP(hipEventSynchronize(g->ev1));
    }
    float gpu_ms = 0;
    if (!G.empty()) CV_HIP(hipEventElapsedTime(&gpu_ms, g->ev0, g->ev1));
    std::vector<uint32_t> slot_of(count, UINT32_MAX);
    for (size_t k = 0; k < G.size(); ++k) {
      slot_of[idx[k]] = (uint32_t)k;
      if (st2[k] != CV_OK) {
        status[idx[k]] = st2[k];
        err[idx[k]] = er2[k];
      }
    }
    // Commit the solvers whose step succeeded.
    each_slot([&](uint32_t i) {
      size_t k = slot_of[i];
      cv_gpu& c = *G[k];
      cv_solver& sv = *S[k];
      cv_state* stt = states[i];
      size_t n = sv.n;
      CV_HIP(hipMemcpyAsync(c.h_x, c.b.xcur64.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      CV_HIP(hipMemcpyAsync(c.h_v, c.b.v64.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      CV_HIP(hipMemcpyAsync(c.h_acc, c.dev.acc.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      CV_HIP(hipStreamSynchronize(c.stream));
      sv.acc.assign(c.h_acc, c.h_acc + 3 * n);
      std::memcpy(stt->x, c.h_x, 3 * n * sizeof(double));
      std::memcpy(stt->v, c.h_v, 3 * n * sizeof(double));
      stt->time += settings->dt;
      sv.has_last = true;
      sv.last_x.assign(stt->x, stt->x + 3 * n);
      sv.last_v.assign(stt->v, stt->v + 3 * n);
      sv.last_time = stt->time;
      for (int q = 0; q < 3; ++q) sv.last_gravity[q] = settings->gravity[q];
      sv.last_damping = settings->damping;
      sv.acc_version = next_acc_version();
      c.dev.acc_version = sv.acc_version;
      if (metrics) {
        fill_metrics(&metrics[i], c, sv, settings, out[k], now_ms() - t0[i], gpu_ms, stt->x);
        metrics[i].time_s = stt->time;
        metrics[i].gpu_compute_ms = gpu_ms;  // the whole batch
      }
    });
  });
  if (outer != CV_OK) return outer;
  for (uint32_t i = 0; i < count; ++i)
    if (status[i] != CV_OK) {
      g->last_error = "batch slot " + std::to_string(i) + ": " + err[i];
      return status[i];
    }
  return CV_OK;
}

// ------------------------------------------------------- learned warm start
cv_status cv_warm_set_params(cv_gpu* g, const float* params) {
  if (!g || !params) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    for (int i = 0; i < cvmlp::NPARAM; ++i)
      if (!std::isfinite(params[i])) fail(CV_ERR_INVALID_ARGUMENT, "non-finite warm-start parameter");
    mlp_set_params(*g, params);
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}

cv_status cv_warm_get_params(cv_gpu* g, float* params) {
  if (!g || !params) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    mlp_get_params(*g, params);
  });
}

cv_status cv_warm_infer(cv_gpu* g, const float* features, uint64_t n, float* out) {
  if (!g || (n && (!features || !out)) || n > 0xFFFFFFF0ull) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    if (!g->mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
    if (!n) return;
    DBuf<float> X, Y;
    X.upload(features, 12 * n, g->stream);
    Y.reserve(3 * n);
    mlp_infer_device(*g, X.p, (uint32_t)n, Y.p);
    CV_HIP(hipMemcpyAsync(out, Y.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}

cv_status cv_warm_gradient(cv_gpu* g, const float* features, const float* targets, uint64_t n, float* grad,
                           float* loss) {
  if (!g || !features || !targets || !grad || n == 0 || n > 0xFFFFFFF0ull) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    if (!g->mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
    DBuf<float> X, T;
    X.upload(features, 12 * n, g->stream);
    T.upload(targets, 3 * n, g->stream);
    mlp_grad_device(*g, X.p, T.p, (uint32_t)n, 0, (uint32_t)n, false);
    CV_HIP(hipMemcpyAsync(grad, g->mlp.grad.p, cvmlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    float l = 0;
    CV_HIP(hipMemcpyAsync(&l, g->mlp.loss.p, sizeof(float), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
    if (loss) *loss = l / (3.0f * (float)n);
  });
}

cv_status cv_warm_train(cv_gpu* g, const float* features, const float* targets, uint64_t n, uint32_t epochs,
                        uint32_t batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!g || !features || !targets || n == 0 || n > 0xFFFFFFF0ull || batch == 0 || !(lr > 0.0f))
    return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    if (!g->mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
    Mlp& m = g->mlp;
    m.X.upload(features, 12 * n, g->stream);
    m.T.upload(targets, 3 * n, g->stream);
    m.Xb.reserve(12 * n);
    m.Tb.reserve(3 * n);
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint64_t state = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto next = [&] {  // splitmix64
      uint64_t z = (state += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    for (uint32_t e = 0; e < epochs; ++e) {
      for (uint64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
      m.perm.upload(perm, g->stream);
      mlp_gather(*g, m.X.p, m.perm.p, (uint32_t)n, 12, m.Xb.p);
      mlp_gather(*g, m.T.p, m.perm.p, (uint32_t)n, 3, m.Tb.p);
      for (uint64_t r = 0; r < n; r += batch) {
        mlp_grad_device(*g, m.Xb.p, m.Tb.p, (uint32_t)n, (uint32_t)r, batch, r > 0);
        mlp_adam_device(*g, lr);
      }
      if (epoch_loss) {
        float l = 0;
        CV_HIP(hipMemcpyAsync(&l, m.loss.p, sizeof(float), hipMemcpyDeviceToHost, g->stream));
        CV_HIP(hipStreamSynchronize(g->stream));
        epoch_loss[e] = l / (3.0f * (float)n);
      }
    }
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}

cv_status cv_warm_features(cv_gpu* g, cv_solver* s, const cv_state* st, const cv_settings* set, float* out) {
  if (!g || !s || !out) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    CV_HIP(hipSetDevice(g->device));
    validate_settings(set);
    validate_state(*s, st);
    ensure_device_solver(*g, *s);
    ensure_step_buffers(*g, *s);
    stage_state(*g, *s, st, set);  // the solver's prediction history, as a step
    uint32_t n = s->n;
    double dt = set->dt, damp = 1.0 + set->damping * dt;
    prologue_reduce(*g, n, dt, damp, set->gravity);
    read_scalars(*g);
    DevSolver& d = g->dev;
    StepBufs& b = g->b;
    auto prologue = [&](auto tag) {
      using R = decltype(tag);
      k_prologue<<<blocks(n + d.n_fixed), TPB, 0, g->stream>>>(
          b.x64.p, b.v64.p, d.acc.p, d.masses.p, d.pinned.p, n, d.fixed_pos.p, d.n_fixed, dt, set->damping,
          set->gravity[0], set->gravity[1], set->gravity[2], g->pred_valid ? 1 : 0, g->d_scal, PBufs<R>::pos(*g),
          PBufs<R>::tgt(*g), b.guess64.p, b.x_old64.p, b.v_old64.p);
    };
    if (g->f64)
      prologue(0.0);
    else
      prologue(0.0f);
    CV_HIP(hipGetLastError());
    double r = set->contact.radius;
    double rq = std::min(std::max(r + g->h_scal->motion, 2.0 * r), 4.0 * r);
    g->b.bvh_rebuild = true;
    gpu_detect(*g, *s, rq, set->contact, 0);
    warm_features(*g, *s, set, rq);
    CV_HIP(hipMemcpyAsync(out, g->feat.p, 12 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}
}  // extern "C"

const std::string* cv_gpu_error_string(const cv_gpu* g) { return &g->last_error; }
