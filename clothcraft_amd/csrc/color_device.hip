// color_device.hip — the fast colouring (CV_FLAG_FAST_COLORING) on the device.
//
// The static conflict graph is coloured once per solver on the host, greedily
// in smallest-last order (colours col0; fewer than index order gives, so fewer
// sweep launches per iteration).  After every detection the device colours
// the graph with its contacts and grab by Jones-Plassmann rounds: in a round,
// every vertex whose uncoloured neighbours all come later in the priority
// order takes the smallest colour that no coloured neighbour holds.  The
// order is col0 ascending, ties by the lowbias32 hash of the vertex index
// (descending).  A vertex is coloured after all its earlier neighbours and
// before all its later ones, so the result is the sequential greedy pass in
// that order, however the rounds interleave: without contacts it is col0
// itself, and a greedy pass in the class order of a colouring never needs
// more colours than that colouring, so contacts are absorbed by the whole
// graph.  The rounds number about the static colour count.  The vertices of
// each colour are listed in index order (stable radix sort by colour).  The
// contact list never leaves the device; the host reads back one colour
// histogram per batch of rounds.
//
// This is not the reference's colouring (greedy in index order,
// src/vbd/coloring.rs:5-43): the Gauss-Seidel order differs, so the steps are
// valid VBD steps but not the reference's.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "gpu.hpp"
#include "solver.hpp"

namespace {

constexpr int TPB = 256;
constexpr uint32_t kMaxColors = 1024;  // colour classes a fast colouring may use
constexpr uint32_t kMaxRounds = 4096;  // Jones-Plassmann rounds per colouring
constexpr uint32_t kRoundBatch = 8;    // rounds queued between read-backs
constexpr uint32_t kNone = 0xFFFFFFFFu;

inline unsigned blocks(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + TPB - 1) / TPB); }

// lowbias32 (C. Wellons, hash-prospector): xorshift-multiply rounds, a
// bijection on 32-bit integers.
__device__ __forceinline__ uint32_t hash32(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

struct Grab3 {
  uint32_t v[3];  // kNone without a grab
};

// Priority of each vertex: static colour ascending, then hash descending
// (larger = earlier).
__global__ void k_fc_prio(uint32_t n, const uint32_t* col0, unsigned long long* prio) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) prio[v] = ((unsigned long long)(0xFFFFFFFFu - col0[v]) << 32) | hash32(v);
}
// Contact-neighbour counts: per contact, the other moving vertices of each of
// its moving vertices.  Contacts list at most 4 vertices; ids >= n are avatar
// vertices, which take no colour.
__global__ void k_fc_count(const uint32_t* cids, const uint32_t* ccount, uint64_t m, uint32_t n, uint32_t* cnt) {
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const uint32_t k = min(ccount[c], 4u);
  uint32_t ids[4];
  for (uint32_t q = 0; q < 4; ++q) ids[q] = q < k ? cids[5 * c + 1 + q] : kNone;
  for (uint32_t a = 0; a < k; ++a) {
    const uint32_t va = ids[a];
    if (va >= n) continue;
    uint32_t others = 0;
    for (uint32_t b = 0; b < k; ++b)
      if (b != a && ids[b] < n && ids[b] != va) ++others;
    if (others) atomicAdd(&cnt[va], others);
  }
}
// Contact-neighbour lists (unordered within a vertex: the colouring reads
// them as sets).
__global__ void k_fc_fill(const uint32_t* cids, const uint32_t* ccount, uint64_t m, uint32_t n,
                          const uint32_t* off, uint32_t* cur, uint32_t* nb) {
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const uint32_t k = min(ccount[c], 4u);
  uint32_t ids[4];
  for (uint32_t q = 0; q < 4; ++q) ids[q] = q < k ? cids[5 * c + 1 + q] : kNone;
  for (uint32_t a = 0; a < k; ++a) {
    const uint32_t va = ids[a];
    if (va >= n) continue;
    for (uint32_t b = 0; b < k; ++b) {
      const uint32_t vb = ids[b];
      if (b == a || vb >= n || vb == va) continue;
      nb[off[va] + atomicAdd(&cur[va], 1u)] = vb;
    }
  }
}
// Jones-Plassmann round r >= 1 over the states (r' << 16) | (colour + 1),
// r' the round that coloured the vertex, 0 = no colour yet.  A neighbour
// coloured in this same round is read as uncoloured (it may or may not have
// been written yet): it then comes earlier, so either reading holds v back,
// and the outcome does not depend on timing.  A vertex left uncoloured raises
// pend[r].
__global__ void k_fc_round(uint32_t r, uint32_t n, const uint32_t* goff, const uint32_t* gnb, const uint32_t* coff,
                           const uint32_t* cnb, const unsigned long long* prio, Grab3 gr, uint32_t* state,
                           uint32_t* pend) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || state[v] != 0) return;
  const unsigned long long pv = prio[v];
  const bool grabbed = gr.v[0] == v || gr.v[1] == v || gr.v[2] == v;
  unsigned long long used = 0;
  // Records the colour of a coloured neighbour within the window [base,
  // base + 64); true when the neighbour is uncoloured and comes earlier.
  auto waits_on = [&](uint32_t u, uint32_t base) {
    if (u >= n || u == v) return false;
    const uint32_t su = state[u];
    if (su == 0 || (su >> 16) == r) return prio[u] > pv;
    const uint32_t c = (su & 0xFFFFu) - 1u;
    if (c >= base && c - base < 64) used |= 1ull << (c - base);
    return false;
  };
  for (uint32_t base = 0;; base += 64) {  // colour windows of 64
    used = 0;
    bool wait = false;
    for (uint32_t k = goff[v]; k < goff[v + 1] && !wait; ++k) wait = waits_on(gnb[k], base);
    for (uint32_t k = coff[v]; k < coff[v + 1] && !wait; ++k) wait = waits_on(cnb[k], base);
    if (grabbed)
      for (int q = 0; q < 3 && !wait; ++q) wait = waits_on(gr.v[q], base);
    if (wait) {
      pend[r] = 1;
      return;
    }
    if (~used) {
      state[v] = (r << 16) | (base + (uint32_t)__builtin_ctzll(~used) + 1u);
      return;
    }
  }
}
// Colour histogram and radix-sort inputs; colours past the table (and, should
// the host read it before the rounds end, vertices still without a colour)
// fall in the overflow bin kMaxColors.
__global__ __launch_bounds__(TPB) void k_fc_keys(uint32_t n, const uint32_t* state, uint32_t* hist, uint32_t* key,
                                                 uint32_t* val) {
  __shared__ uint32_t h[kMaxColors + 1];  // block histogram, one global add per bin
  for (uint32_t i = threadIdx.x; i <= kMaxColors; i += TPB) h[i] = 0;
  __syncthreads();
  for (uint32_t v = blockIdx.x * TPB + threadIdx.x; v < n; v += gridDim.x * TPB) {
    const uint32_t s = state[v];
    const uint32_t c = s ? min((s & 0xFFFFu) - 1u, kMaxColors) : kMaxColors;
    atomicAdd(&h[c], 1u);
    key[v] = c;
    val[v] = v;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i <= kMaxColors; i += TPB)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

}  // namespace

uint32_t gpu_fast_colors(cv_gpu& g, const cv_solver& s, const cv_grab* grab, std::vector<uint32_t>& color_off) {
  DevSolver& d = g.dev;
  StepBufs& b = g.b;
  hipStream_t st = g.stream;
  const uint32_t n = s.n;
  if (!d.fc_ready) {
    std::vector<uint32_t> col, order;
    greedy_colors(s, nullptr, nullptr, nullptr, col, order, d.fc_off0, true);
    d.fc_col0.upload(col, st);
    d.fc_order0.upload(order, st);
    d.graph_off.upload(s.graph_off, st);
    std::vector<uint32_t> gnb(s.graph);
    if (gnb.empty()) gnb.push_back(0);
    d.graph.upload(gnb, st);
    d.fc_prio.reserve(n);
    k_fc_prio<<<blocks(n), TPB, 0, st>>>(n, d.fc_col0.p, d.fc_prio.p);
    CV_HIP(hipGetLastError());
    CV_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    d.fc_ready = true;
  }
  if (!g.h_fc) CV_HIP(hipHostMalloc((void**)&g.h_fc, (kMaxColors + 2) * sizeof(uint32_t)));
  b.order.reserve(n);
  const uint64_t m = g.h_scal->contacts;
  if (m == 0 && !grab) {  // without contacts or grab the colouring is the static one
    CV_HIP(hipMemcpyAsync(b.order.p, d.fc_order0.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    color_off = d.fc_off0;
    return (uint32_t)color_off.size() - 1;
  }
  Grab3 gr{{kNone, kNone, kNone}};
  if (grab)
    for (int q = 0; q < 3; ++q) gr.v[q] = grab->vertices[q];
  b.fc_cnt.reserve(n + 1);
  b.fc_off.reserve(n + 1);
  b.fc_cur.reserve(n);
  b.fc_nb.reserve(std::max<uint64_t>(12 * m, 1));  // at most 4 x 3 per contact
  b.fc_state.reserve(n);
  b.fc_pend.reserve(kMaxRounds);
  b.fc_hist.reserve(kMaxColors + 1);
  b.fc_key.reserve(n);
  b.fc_key_sorted.reserve(n);
  b.fc_val.reserve(n);
  CV_HIP(hipMemsetAsync(b.fc_cnt.p, 0, (n + 1) * sizeof(uint32_t), st));
  CV_HIP(hipMemsetAsync(b.fc_cur.p, 0, n * sizeof(uint32_t), st));
  CV_HIP(hipMemsetAsync(b.fc_state.p, 0, n * sizeof(uint32_t), st));
  CV_HIP(hipMemsetAsync(b.fc_pend.p, 0, kMaxRounds * sizeof(uint32_t), st));
  if (m) {
    k_fc_count<<<blocks(m), TPB, 0, st>>>(b.c_ids.p, b.c_count.p, m, n, b.fc_cnt.p);
    CV_HIP(hipGetLastError());
  }
  device_scan_u32(g, b.fc_cnt.p, b.fc_off.p, (size_t)n + 1);
  if (m) {
    k_fc_fill<<<blocks(m), TPB, 0, st>>>(b.c_ids.p, b.c_count.p, m, n, b.fc_off.p, b.fc_cur.p, b.fc_nb.p);
    CV_HIP(hipGetLastError());
  }
  // Rounds in batches; each batch ends with the histogram and one read-back
  // that also says whether the last round left a vertex uncoloured.
  // The first batch covers the static colour count (the rounds a colouring
  // without conflicts needs) and a batch more.
  uint32_t r = 0, batch = (uint32_t)d.fc_off0.size() - 1 + kRoundBatch;
  for (;; batch = kRoundBatch) {
    for (uint32_t i = 0; i < batch; ++i) {
      if (++r >= kMaxRounds) fail(CV_ERR_CAPACITY, "fast colouring did not finish within 4096 rounds");
      k_fc_round<<<blocks(n), TPB, 0, st>>>(r, n, d.graph_off.p, d.graph.p, b.fc_off.p, b.fc_nb.p, d.fc_prio.p,
                                            gr, b.fc_state.p, b.fc_pend.p);
    }
    CV_HIP(hipGetLastError());
    CV_HIP(hipMemsetAsync(b.fc_hist.p, 0, (kMaxColors + 1) * sizeof(uint32_t), st));
    k_fc_keys<<<std::min(blocks(n), 64u), TPB, 0, st>>>(n, b.fc_state.p, b.fc_hist.p, b.fc_key.p, b.fc_val.p);
    CV_HIP(hipGetLastError());
    CV_HIP(hipMemcpyAsync(g.h_fc, b.fc_hist.p, (kMaxColors + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CV_HIP(hipMemcpyAsync(g.h_fc + kMaxColors + 1, b.fc_pend.p + r, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CV_HIP(hipStreamSynchronize(st));
    if (!g.h_fc[kMaxColors + 1]) break;
  }
  if (g.h_fc[kMaxColors]) fail(CV_ERR_CAPACITY, "fast colouring needs more than 1024 colours");
  // Colour-grouped order, index order within a colour (stable sort).
  size_t bytes = 0;
  CV_HIP(rocprim::radix_sort_pairs(nullptr, bytes, b.fc_key.p, b.fc_key_sorted.p, b.fc_val.p, b.order.p,
                                   (size_t)n, 0, 11, st));
  b.temp.reserve(bytes + 256);
  bytes = b.temp.n;
  CV_HIP(rocprim::radix_sort_pairs((void*)b.temp.p, bytes, b.fc_key.p, b.fc_key_sorted.p, b.fc_val.p, b.order.p,
                                   (size_t)n, 0, 11, st));
  uint32_t k = 0;
  for (uint32_t c = 0; c < kMaxColors; ++c)
    if (g.h_fc[c]) k = c + 1;
  color_off.assign(k + 1, 0);
  for (uint32_t c = 0; c < k; ++c) color_off[c + 1] = color_off[c] + g.h_fc[c];
  return k;
}
