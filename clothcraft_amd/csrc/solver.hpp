// solver.hpp — host-side VbdSolver state of libclothvbd (the handle behind
// cv_solver).  Mirrors what VbdSolver::new builds (src/vbd.rs:101-192): the
// OGC topology over garment + avatar faces, seam exclusions, per-vertex term
// incidence and the symmetric conflict graph; plus the packed per-vertex
// term records the device sweep reads.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/clothvbd.h"
#include "vbd_math.hpp"

struct CvError {
  cv_status status;
  std::string message;
};

struct OgcTopology {  // src/ogc/topology.rs:13-52 over all faces (garment, then avatar)
  uint32_t n_total = 0;                 // garment + avatar vertices
  uint32_t moving = 0;                  // garment vertex count
  std::vector<uint32_t> faces;          // [3F]
  std::vector<uint32_t> edges;          // [2E] (min, max)
  std::vector<uint32_t> tri_edges;      // [3F]
  std::vector<int32_t> opposites;       // [2E], -1 when absent
  std::vector<uint32_t> nb_off, nb;     // neighbour CSR (sorted)
  std::vector<uint32_t> ex_off, ex;     // seam-excluded neighbour CSR (sorted)
  std::vector<double> rest_len2;        // [E]
  uint32_t moving_faces = 0;            // faces touching a garment vertex come first? (count only)
  std::vector<uint8_t> face_moving;     // [F]
  std::vector<uint8_t> edge_moving;     // [E]
};

struct cv_solver {
  uint32_t n = 0;
  uint32_t n_free = 0;                  // vertices without a pin
  std::vector<double> masses;
  std::vector<uint8_t> pinned;
  std::vector<double> pin_targets;
  std::vector<uint32_t> tri_v;          // [3T]
  std::vector<double> tri_inv_dm;       // [4T]
  std::vector<double> tri_area;
  std::vector<uint32_t> tri_mat;
  std::vector<uint32_t> hinge_v;        // [4H]
  std::vector<double> hinge_weight, hinge_across;
  std::vector<uint32_t> hinge_mat;
  std::vector<cv_material> materials;
  std::vector<uint32_t> st_off, st_v;   // stitch samples
  std::vector<double> st_w, st_k;
  std::vector<uint32_t> pr_off, pr_v;   // seam probes
  std::vector<double> pr_w;
  std::vector<double> avatar;           // [3*n_fixed] fixed positions
  OgcTopology top;
  std::vector<uint32_t> graph_off, graph;  // static conflict graph (sorted)
  std::vector<uint32_t> sl_order;          // smallest-last order of that graph (CV_FLAG_FAST_COLORING)
  std::vector<uint32_t> term_off;          // [n+1] static TermRec per vertex
  std::vector<cvm::TermRec> terms;      // f32 coefficients (f32 sweep)
  std::vector<cvm::TermRec64> terms64;  // f64 coefficients (f64 sweep)
  uint64_t generation = 0;                 // changes when a new solver is built
  double face_extent_median = 0;           // grid cell hint for device detection

  // Prediction history (src/vbd/prediction.rs:8-13); the acceleration lives
  // where it was last computed: `acc` on the host after record_prediction,
  // the device cache after a device step (acc_on_device).
  bool has_last = false;
  std::vector<double> last_x, last_v;
  double last_time = 0;
  double last_gravity[3] = {0, 0, 0};
  double last_damping = 0;
  std::vector<double> acc;
  uint64_t acc_version = 0;

  std::string last_error;
};

// Builders (solver.cpp).
void build_solver(cv_solver& s, const cv_garment* g, const cv_surface* avatar, const cv_state* st);
void validate_settings(const cv_settings* s);
void validate_state(const cv_solver& s, const cv_state* st);
void validate_grab(const cv_solver& s, const cv_grab* g);
// Host half of GpuVbd.check_gradients (src/gpu_vbd.rs:336-372): every
// vertex's local objective in f64 (closed-form term derivatives over the
// solver's f64 records, the contact records, the grab, inertia) at the given
// positions x ([3 (n + fixed)], relative to `origin`) and targets t ([4n]:
// target relative to `origin`, inertia), compared with the device probe
// ([4n]: g, E).  Throws CV_ERR_GRADIENT_CHECK with the reference's message on
// a gradient error above 3 mN + 0.3 % or an energy error above 1e-5 + 0.1 %;
// returns the largest gradient error.
double check_probe(const cv_solver& s, const double* x, const double* t, const uint32_t* coff,
                   const cvm::TermRec64* crec, const double* origin, const cv_settings* set, const cv_grab* grab,
                   const double* probe);
bool prediction_valid(const cv_solver& s, const cv_state* st, const cv_settings* set);

// Greedy colouring in vertex order (src/vbd/coloring.rs:5-43), or with
// smallest_last in the smallest-last order of the static conflict graph
// (s.sl_order: fewer colours, a different Gauss-Seidel order than the
// reference; without contacts these are the static classes of the device's
// fast colouring).  cnb_off/cnb is the per-vertex contact incidence: for moving
// vertex v, the vertices of each incident contact.  Returns colour of each
// vertex and the colour count; `order` lists vertices colour by colour, in
// index order within a colour.
uint32_t greedy_colors(const cv_solver& s, const uint32_t* cnb_off, const uint32_t* cnb,
                       const cv_grab* grab, std::vector<uint32_t>& color_of,
                       std::vector<uint32_t>& order, std::vector<uint32_t>& color_off,
                       bool smallest_last = false);


// The contacts incident to the vertices < n (indices into the contact rows),
// as (vertex, row) pairs grouped by blocks of 64 vertices: a counting sort
// whose counters stay in cache, and a lookup scans one block's pairs.
struct ContactNb {
  std::vector<uint32_t> off;      // [n / 64 + 2] start of each block's pairs
  std::vector<uint64_t> pairs;    // vertex << 32 | row
  void build(uint32_t n, const uint32_t* kind_ids, const uint32_t* counts, uint64_t m);
  template <class F>
  void for_rows(uint32_t v, F&& f) const {
    for (uint32_t k = off[v >> 6]; k < off[(v >> 6) + 1]; ++k)
      if ((uint32_t)(pairs[k] >> 32) == v) f((uint32_t)pairs[k]);
  }
};

// Greedy colouring kept across detections of one solver: a vertex is
// recoloured only when its contacts or grab membership changed or an earlier
// vertex (in colouring order) among its neighbours changed colour, so the
// result equals greedy_colors from scratch while the work follows the change.
// The changed contacts are found by merging the old and new contact rows
// (both in the canonical sorted order; a row met in only one list marks its
// vertices, so any order is safe, sorted is fast); the vertices to recolour
// are taken in colouring order from a bitmap, and only contact vertices and
// their cascade are visited.  A new solver generation or order mode starts
// over.
struct ColoringState {
  bool valid = false, smallest_last = false;
  uint64_t generation = 0;
  uint32_t grab_key[4] = {0, 0, 0, 0};
  std::vector<uint32_t> color_of, pos, count;  // colour and order position of each vertex; vertices per colour
  std::vector<uint32_t> prev;                  // contact rows of the last colouring: kind, ids[4], count
  ContactNb nb;                                // incident contacts of this colouring's rows
  std::vector<uint64_t> dirty;                 // bitmap over colouring positions, all zero between calls
  std::vector<uint64_t> seen;                  // [n + 2] colour marks, stamped per visited vertex
  uint64_t stamp = 0;
  bool full = false;              // the last update coloured from scratch
  std::vector<uint32_t> changed;  // vertices whose colour the last incremental update changed
};
// Contact rows are (kind, ids[4]) with `counts` vertices each.  Writes the
// colour offsets (vertices per colour, prefix-summed) and returns the colour
// count; cs.color_of holds every vertex's colour.
uint32_t update_colors(const cv_solver& s, ColoringState& cs, const uint32_t* kind_ids, const uint32_t* counts,
                       uint64_t m, const cv_grab* grab, bool smallest_last, std::vector<uint32_t>& color_off);

[[noreturn]] void fail(cv_status st, const std::string& msg);
