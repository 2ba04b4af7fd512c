// gpu.hpp — device-side state of libclothvbd: the cv_gpu handle (stream,
// scratch, per-solver device cache) and the layout of every device array.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mlp.hpp"
#include "solver.hpp"

#define CV_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      fail(CV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

// Grow-only device buffer.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  // Grows by at least half the current capacity, so sizes that creep up
  // step by step (contact and pair counts) reallocate rarely: hipFree waits
  // for the device.
  void reserve(size_t count) {
    if (count <= n) return;
    if (p) CV_HIP(hipFree(p));
    p = nullptr;
    size_t cap = std::max(count, n + n / 2);
    CV_HIP(hipMalloc(&p, std::max<size_t>(cap, 1) * sizeof(T)));
    n = cap;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    reserve(count);
    if (count) CV_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// Step scalars shared by kernels and read back by the host in one copy.
// One hash-grid registration in bucket order: the cell key (hash
// collisions are filtered by it), the primitive and its unpadded box, one
// 64-byte line, so the counting sort's scatter writes whole lines and the
// broad pass reads each entry with one contiguous load.  Queries are
// registered the same way (a point in its own cell, a moving edge in the
// cells of its box), so the queries of one cell sit together.
struct alignas(64) GridEntry {
  unsigned long long key;
  uint32_t prim, pad;
  double box[6];
};
static_assert(sizeof(GridEntry) == 64, "one grid entry per 64-byte line");

struct Scalars {
  double origin[3];
  double rq;
  double padded_radius2;
  double motion;
  double min_distance;
  double residual_sum;
  double true_residual_sum;
  double scale;
  unsigned long long vf_count, ee_count;        // pairs with distance < rq
  unsigned long long vf_cand, ee_cand;          // pairs passing the padded AABB test
  unsigned long long grid_tests;                // pairs examined by the grid
  unsigned long long raw_contacts, contacts, active, incidences;
  unsigned long long clipped;
  unsigned long long truncated_total;  // clipped vertices summed over the step's iterations
  unsigned int stop, stop_it;          // refresh needed after iteration stop_it (device decides)
  unsigned int iters_run;
  unsigned int sweep_blocks;           // blocks of the iteration's last sweep launch that have finished
  unsigned long long cand_export_vf, cand_export_ee;
  unsigned long long vf_list_n, ee_list_n;      // candidates queued for the narrow phase (all regions)
  unsigned long long vf_list_max, ee_list_max;  // the most queued in one region
  unsigned long long grid_entries;          // hash-grid registrations of the last detection
  double diag_p95, diag_max;                // extension percent of the committed positions
  unsigned long long diag_gap_bits;         // max seam gap (bit pattern of a non-negative double)
  unsigned int err_degenerate, err_overflow, err_nonfinite, err_trust;
};

// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin
// (block b runs on XCD b % 8), and each XCD has its own L2.  Kernels whose
// neighbouring blocks share data remap the block index so every XCD receives
// one contiguous range of logical blocks: a bijection of [0, nb).
#ifdef __HIPCC__
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  constexpr uint32_t X = 8;
  uint32_t x = b % X, k = b / X, q = nb / X, r = nb % X;
  return x * q + (x < r ? x : r) + k;
}
#endif

// Per-solver device cache (the role of Prepared, src/gpu_vbd.rs:478-547):
// static arrays are uploaded once per solver generation.
struct DevSolver {
  uint64_t generation = 0;
  uint32_t n = 0, n_fixed = 0, n_total = 0, F = 0, E = 0;
  DBuf<uint32_t> term_off;
  DBuf<cvm::TermRec> terms;
  DBuf<cvm::TermRec64> terms_d;
  DBuf<double> masses;
  DBuf<uint8_t> pinned;
  DBuf<uint32_t> faces, edges, nb_off, nb, ex_off, ex;
  DBuf<int32_t> opposites;
  DBuf<uint8_t> face_moving, edge_moving;
  DBuf<double> rest_len2;
  DBuf<double> fixed_pos;  // avatar f64
  double fixed_abs_max = 0;
  double cell = 0;         // grid cell size for detection
  // garment diagnostics: triangles, inverse rest shapes, grain cos/sin,
  // areas, seam probes
  uint32_t T = 0, P = 0;
  DBuf<uint32_t> dg_tri, dg_pr_off, dg_pr_v;
  DBuf<double> dg_inv, dg_cs, dg_area, dg_pr_w;
  // fast colouring: static conflict graph, its smallest-last colouring (per
  // vertex and as a colour-grouped order with host offsets) and the colouring
  // priority; built on the first fast colouring of this solver
  bool fc_ready = false;
  DBuf<uint32_t> graph_off, graph, fc_col0, fc_order0;
  DBuf<unsigned long long> fc_prio;  // colouring priority (static colour, then hash)
  std::vector<uint32_t> fc_off0;
  // prediction
  DBuf<double> acc;
  uint64_t acc_version = 0;
  bool acc_device = false;
};

// Device working set of one step (sized to the largest solver seen).
// One linear BVH (ogc_device.hip): Morton keys, leaf order, Karras tree, boxes.
struct BvhBufs {
  DBuf<unsigned long long> key, key_sorted;
  DBuf<uint32_t> idx, prim, flag, top;
  DBuf<int32_t> child, parent, range;
  DBuf<double> box;
  uint32_t N = 0;
};

struct StepBufs {
  DBuf<double> x64, x_old64, v64, v_old64, guess64, xcur64, radii64, centers64;
  DBuf<unsigned long long> dist_bits;
  DBuf<double4> pos_d, tgt_d, ball_d;  // sweep positions, targets, balls (both precisions)
  DBuf<double> gsq;
  DBuf<double> sum_parts;  // partials of the deterministic sums
  DBuf<double> gtrue;  // per-vertex |g|^2 of the true residual
  DBuf<uint32_t> order;
  DBuf<uint32_t> color_dev, color_chg, color_keys, color_iota;  // colour of each vertex; changes; sort scratch
  DBuf<uint4> vinfo;  // colour-ordered (vertex, term offset, term count, contact offset)
  DBuf<uint32_t> vnc;  // colour-ordered contact-record counts
  // 32-lane sweep: slots whose vertex has more than 32 records (flag, its
  // exclusive scan over the slots, the slots in order), swept by 64 lanes
  DBuf<uint32_t> hflag, hpos, hlist;
  // fast colouring: contact-neighbour CSR, round states, pending-round
  // flags, colour histogram, sort keys
  DBuf<uint32_t> fc_cnt, fc_off, fc_cur, fc_nb, fc_state, fc_pend, fc_hist, fc_key, fc_key_sorted,
      fc_val;
  // detection
  DBuf<double> fbox, ebox;  // [6F], [6E]
  BvhBufs bvh_faces, bvh_edges;          // LBVH broad phase (cv_gpu_set_broad_phase 1)
  BvhBufs bvh_points;                    // points in Morton order (the VF walk's query order)
  uint64_t cand_cap_vf = 0, cand_cap_ee = 0;  // candidate list regions of the broad pass (grid or LBVH)
  bool bvh_rebuild = true;               // rebuild (not refit) at the next LBVH detection
  bool bvh_fallback = false;             // this detection runs the grid (an LBVH walk overflowed)
  uint64_t bvh_fallbacks = 0;
  DBuf<uint32_t> cell_count, cell_off;
  DBuf<unsigned long long> cell_key;
  DBuf<uint32_t> cell_val, cell_rank, cell_owner;
  DBuf<GridEntry> cell_entry;  // bucketed registrations
  DBuf<uint32_t> bcount, bstart;  // hash-grid buckets [4T+1]: faces, edges, points, moving-edge cells
  DBuf<uint2> pair_list;  // candidates for the narrow phase: VF region, then EE region
  DBuf<unsigned long long> list_ctr;  // broad-pass counters, one per 128-byte line (see ogc_device.hip)
  DBuf<uint32_t> scan_sums;  // tile totals of the two-level scan
  DBuf<double> red_parts;    // per-block partials of the prologue reduction
  DBuf<unsigned long long> ext_key, ext_key_sorted;  // diagnostics: per-triangle extension
  DBuf<double> ext_area, ext_area_sorted, ext_prefix;
  uint64_t grid_cap = 0, grid_last = 0;  // registration capacity; last detection's count
  DBuf<uint32_t> raw_ids;   // [5*cap] kind, ids
  DBuf<uint32_t> raw_count;
  DBuf<double> raw_eps;
  DBuf<unsigned long long> key_hi, key_lo, key_tmp, key_hi_sorted;
  DBuf<uint32_t> perm, perm_tmp;
  DBuf<uint32_t> uniq_flag, uniq_pos;
  DBuf<uint32_t> c_ids;     // unique contacts [5*cap]
  DBuf<uint32_t> c_count;
  DBuf<double> c_eps;
  DBuf<unsigned long long> inc_key, inc_key_sorted;
  DBuf<uint32_t> coff, vcount;
  DBuf<cvm::TermRec> crec;
  DBuf<cvm::TermRec64> crec_d;
  DBuf<uint32_t> cand_vf, cand_ee;  // optional candidate export
  DBuf<unsigned char> temp;         // rocPRIM temporary storage
  size_t contact_cap = 0;
};

// Warm-start network state on the device (FP32 master parameters, Adam
// moments, BF16 operand tiles, gradient scratch, training data).
// Workspace of the device Newton reference (cv_gpu_newton).
struct NtScal {  // PCG scalars kept on the device between read-backs
  double rz, tol, h, inv2h, alpha, beta;
  unsigned int state;  // 0 running, 1 converged, 2 non-positive curvature
  unsigned int its;    // iterations taken when state became non-zero
};
struct NewtonBufs {
  DBuf<double4> X, Xt;  // iterate; trial and difference points (garment, then avatar)
  DBuf<double> x0, v0, g, gp, gm, diag, diag0, sol, r, z, p, ap, epart, parts;
  DBuf<NtScal> sc;
};

struct Mlp {
  DBuf<float> master, grad, m1, m2, partial, loss_partial, loss;
  DBuf<uint16_t> fw, bw;
  DBuf<float> X, T, Xb, Tb;  // training set and the current shuffled copy
  DBuf<uint32_t> perm;
  int adam_t = 0;
  bool ready = false;
};

// Fixed worker threads for the per-solver host phases of a batched step
// (detection read-backs, host colouring, staging).  run() calls f(i) for
// every i in [0, count) on the workers and the calling thread and returns
// when all calls are done; f must not throw.
class WorkPool {
 public:
  explicit WorkPool(unsigned threads) {
    for (unsigned t = 0; t < threads; ++t) workers_.emplace_back([this] { loop(); });
  }
  ~WorkPool() {
    {
      std::lock_guard<std::mutex> l(m_);
      quit_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  void run(size_t count, const std::function<void(size_t)>& f) {
    if (count == 0) return;
    {
      std::lock_guard<std::mutex> l(m_);
      f_ = &f;
      count_ = count;
      next_.store(0);
      pending_ = count;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [this] { return pending_ == 0 && active_ == 0; });
    f_ = nullptr;
  }

 private:
  // A thread that picked up the function pointer counts as active until it
  // leaves, so run() cannot return (and the function go out of scope) while a
  // late worker still holds it.
  void work() {
    const std::function<void(size_t)>* f;
    size_t count, done = 0;
    {
      std::lock_guard<std::mutex> l(m_);
      f = f_;
      count = count_;
      if (!f) return;
      ++active_;
    }
    for (size_t i; (i = next_.fetch_add(1)) < count;) {
      (*f)(i);
      ++done;
    }
    std::lock_guard<std::mutex> l(m_);
    pending_ -= done;
    --active_;
    if (pending_ == 0 && active_ == 0) done_.notify_all();
  }
  void loop() {
    uint64_t seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return quit_ || gen_ != seen; });
        if (quit_) return;
        seen = gen_;
      }
      work();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)>* f_ = nullptr;
  size_t count_ = 0, pending_ = 0, active_ = 0;
  std::atomic<size_t> next_{0};
  uint64_t gen_ = 0;
  bool quit_ = false;
};

struct cv_gpu {
  int device = 0;
  hipStream_t stream = nullptr;
  bool check_gradients = false;
  int broad_phase = 0;  // 0: hash grid, 1: LBVH
  int warm_smooth = 0;  // one-ring averaging passes of the warm-start correction
  double warm_scale = 1.0;        // factor of the warm-start correction
  bool warm_normal_only = false;  // keep only the correction's component along the vertex normal
  double last_check_max_error = 0;  // largest gradient error of the last check (N)
  DBuf<double> probe_out;           // [4n] probe of the gradient check
  DevSolver dev;
  StepBufs b;
  Scalars* d_scal = nullptr;  // device
  Scalars* h_scal = nullptr;  // pinned host mirror
  double* h_x = nullptr;      // pinned staging [3n]
  double* h_v = nullptr;
  double* h_acc = nullptr;    // pinned staging of the acceleration history [3n]
  size_t h_cap = 0;
  std::string last_error;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // host-side contact incidence for the greedy colouring
  // Pinned copy of the contact list for the host colouring: kind and ids
  // (5 per contact) then counts (1 per contact), capacity h_cont_cap contacts.
  uint32_t* h_fc = nullptr;     // pinned read-back of the fast colouring's histogram
  bool host_contacts = true;    // detection copies the contact list to h_cont
  bool h_cont_valid = false;    // h_cont holds the last detection's contacts
  bool detected = false;        // a detection of the current device solver ran
  uint32_t* h_cont = nullptr;
  size_t h_cont_cap = 0;
  uint32_t* h_cchg = nullptr;  // pinned staging of an incremental colouring's (vertex, colour) changes
  bool order_host = false;     // b.order holds the host colouring's order for b.color_dev
  size_t h_cchg_cap = 0;
  std::vector<uint32_t> color_off;
  ColoringState coloring;  // incremental greedy colouring across detections
  // Inputs of the last colouring, to reuse it when they repeat.
  bool color_valid = false;
  uint64_t color_generation = 0, color_m = 0;
  uint32_t color_key[4] = {0, 0, 0, 0};
  std::vector<uint32_t> color_cids;  // contact kinds/ids then counts
  // Sweep launch shape (lanes per vertex), chosen by the autotuner.
  int lanes = 32;
  int batch_block = 256;  // threads per block of the batched sweep, chosen by the autotuner
  uint32_t tol_iters = 0;  // iterations the last residual-tolerance step ran (sizes the sweep queue)
  bool pred_valid = false;
  // Live sweep-kernel timing (events around every sweep launch).
  bool profile = false;
  std::vector<hipEvent_t> sweep_events;
  double event_overhead_ms = 0;  // event interval of an empty launch (profiling calibration)
  uint64_t timed_counter = 0;  // iterations queued with profiling on
  // Algorithmic bytes of one sweep over the static terms: the f32 minimum,
  // the f32 sweep as stored, the f64 sweep.
  double static_bytes32 = 0, static_bytes32own = 0, static_bytes64 = 0;
  bool f64 = true;  // sweep precision (f64 reproduces the f64 reference step)
  NewtonBufs nt;
  Mlp mlp;
  DBuf<float> feat, wpred;  // warm-start features [12n] and predictions [3n]
  // Batched steps: one context per solver of the batch, each with its own
  // stream for its detection and bookkeeping; the shared sweeps run on this
  // context's stream, ordered against the solvers' streams by events.
  // Launch tables of the shared sweeps; host worker threads.
  std::vector<cv_gpu*> subs;
  hipEvent_t join_ev = nullptr;  // no timing: stream-order handoffs
  DBuf<unsigned char> batch_args, batch_slots, batch_ends;
  std::unique_ptr<WorkPool> pool;
};

// Runs f, mapping exceptions to a status and message (no exception crosses
// the C ABI).
template <class F>
cv_status guard(std::string* err, F&& f) {
  try {
    f();
    return CV_OK;
  } catch (const CvError& e) {
    if (err) *err = e.message;
    return e.status;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return CV_ERR_DEVICE;
  }
}

// Warm-start network (mlp_device.hip).
void mlp_ensure(cv_gpu& g);
// Takes the device address of cv_warm_infer_noalloc's staging array on `device`
// (current), once per device.
void warm_scratch_setup(int device);
void mlp_set_params(cv_gpu& g, const float* params);
void mlp_get_params(cv_gpu& g, float* params);
void mlp_infer_device(cv_gpu& g, const float* dX, uint32_t n, float* dY);
void mlp_grad_device(cv_gpu& g, const float* dX, const float* dT, uint32_t n, uint32_t row0, uint32_t count,
                     bool accumulate_loss);
void mlp_gather(cv_gpu& g, const float* src, const uint32_t* perm, uint32_t n, int width, float* dst);
void mlp_adam_device(cv_gpu& g, float lr);

// Detection on device positions b.x64 (all n_total vertices) with query
// radius rq; fills distance, radii, unique contacts and contact records.
// Candidate pairs are also exported when cand_cap > 0.
// Garment diagnostics of the positions dx (moving vertices, f64) on the
// device into the scalars: max seam gap and the area-weighted 95th percentile
// and maximum of the positive axial extension.
void gpu_diagnostics(cv_gpu& g, const double* dx);

void gpu_detect(cv_gpu& g, const cv_solver& s, double rq, const cv_ogc_settings& ogc,
                size_t cand_cap);
// Exclusive scan of count uint32 values on the context's stream.
void device_scan_u32(cv_gpu& g, const uint32_t* in, uint32_t* out, size_t count);
// b.order = the vertices 0..n-1 stably sorted by b.color_dev (colours below ncolors).
void device_color_order(cv_gpu& g, uint32_t n, uint32_t ncolors);
// Device Newton reference of one step (newton_device.hip).
void newton_reference(cv_gpu& g, const cv_solver& s, const cv_state* st, const cv_settings* set,
                      const cv_newton_settings* ns, double* x_out, cv_newton_metrics* out);
// Fast colouring (CV_FLAG_FAST_COLORING) of the last detection's contacts and
// the grab, on the device: the colour-grouped vertex order into b.order and
// the colour offsets into color_off; returns the colour count.
uint32_t gpu_fast_colors(cv_gpu& g, const cv_solver& s, const cv_grab* grab, std::vector<uint32_t>& color_off);
void ensure_device_solver(cv_gpu& g, const cv_solver& s);
void ensure_step_buffers(cv_gpu& g, const cv_solver& s);
