/*
 * clothvbd.h — C ABI of libclothvbd.so, the MI355X Vertex Block Descent step.
 *
 * Every entry point is extern "C", takes plain pointers and sizes and never
 * lets an exception cross the boundary.  Status codes map one-to-one to the
 * error strings of the Rust API this library replaces; cv_last_error() returns
 * the exact message.  On any status other than CV_OK the caller's state arrays
 * are left untouched (the reference steps are transactional: src/vbd.rs:256-274,
 * 399-405 and src/gpu_vbd.rs:132-152,170-172,435-439).
 *
 * Ownership: the caller owns every array passed in; create calls copy what
 * they keep.  Threading: a handle is not thread-safe (the reference GpuVbd is
 * !Sync, src/gpu_vbd.rs:23-24); one HIP stream per cv_gpu.
 *
 * Layouts follow the reference types: positions and velocities are
 * xyz-interleaved f64 (glam DVec3, i.e. what Vec<DVec3>::as_ptr() yields),
 * Dm^-1 is glam DMat2 column-major (x_axis.x, x_axis.y, y_axis.x, y_axis.y).
 */
#ifndef CLOTHVBD_H
#define CLOTHVBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum cv_status {
  CV_OK = 0,
  CV_ERR_INVALID_ARGUMENT = 1,     /* null pointer, size mismatch in the call itself        */
  CV_ERR_INVALID_SETTINGS = 2,     /* "Invalid VBD settings" (src/vbd.rs:38-51)             */
  CV_ERR_INVALID_OGC_SETTINGS = 3, /* "Invalid OGC radius, stiffness ..." (ogc/world.rs:56) */
  CV_ERR_INVALID_STATE = 4,        /* "Invalid VBD state" (src/vbd.rs:258-264)              */
  CV_ERR_INVALID_GRAB = 5,         /* "Invalid mouse attachment" (src/vbd.rs:265-274)       */
  CV_ERR_DIMENSIONS = 6,           /* "VBD state dimensions do not match garment"           */
  CV_ERR_PIN_MISMATCH = 7,         /* "VBD pinned positions must match ..." (vbd.rs:110)    */
  CV_ERR_TOPOLOGY = 8,             /* invalid triangle indices / non-manifold edge          */
  CV_ERR_DEGENERATE_FACE = 9,      /* "OGC degenerate (initial) face"                       */
  CV_ERR_INVALID_AVATAR = 10,      /* SurfaceMesh::validate (src/avatar.rs:78-97)           */
  CV_ERR_INITIAL_INTERSECTION = 11,/* "OGC initial intersection between faces i and j"      */
  CV_ERR_INITIAL_TOUCHING = 12,    /* "OGC initial touching/coplanar features"              */
  CV_ERR_ZERO_SEPARATION = 13,     /* "OGC zero separation; step rejected"                  */
  CV_ERR_SEPARATION_F32 = 14,      /* "GPU OGC separation below f32 safety margin"          */
  CV_ERR_NONFINITE_RESULT = 15,    /* "Non-finite VBD result; state unchanged"              */
  CV_ERR_TRUST_REGION = 16,        /* "OGC trust-region invariant failed"                   */
  CV_ERR_FIXED_CHANGED = 17,       /* "OGC fixed geometry changed; rebuild the world"       */
  CV_ERR_QUERY_RADIUS = 18,        /* "OGC query radius must cover the contact radius"      */
  CV_ERR_INVALID_POSITIONS = 19,   /* "Invalid OGC positions" / invalid position array      */
  CV_ERR_DEVICE = 20,              /* a HIP runtime call failed (message has the HIP text)  */
  CV_ERR_CAPACITY = 21,            /* a request exceeds a documented size limit             */
  CV_ERR_INVALID_SEAM = 22,        /* stitch weights invalid                                */
  CV_ERR_NOT_BUILT = 23,           /* a feature (e.g. warm-start network) is not set up     */
  CV_ERR_INVALID_MATERIAL = 24,    /* Material::validate (src/model.rs:40-63)               */
  CV_ERR_LINE_SEARCH = 25,         /* "Newton line search failed; ..." (src/solver.rs:547)   */
  CV_ERR_GRADIENT_CHECK = 26       /* "GPU gradient/energy mismatch at i" (gpu_vbd.rs:361-372) */
} cv_status;

/* ------------------------------------------------------------- inputs */
/* Material — src/model.rs:7-22 (all fields, same units). */
typedef struct cv_material {
  double warp, weft, shear;
  double compression_ratio, compression_transition;
  double bend_warp, bend_weft;
  double density, thickness, grain_degrees;
} cv_material;

/* CompiledGarment — src/garment/template.rs:170-180, flattened.  Stitch
 * samples are the seam Gauss samples (src/garment/seam.rs:269-272) with
 * stitch_stiffness[s] = seam.stiffness * sample.weight_m, exactly the
 * coefficient VbdSolver::new forms (src/vbd.rs:139-142).  Probes are the
 * seam probe points used by max_seam_gap (src/garment/seam.rs:418-423). */
typedef struct cv_garment {
  uint32_t n_vertices;
  const double* masses;           /* [n]                                        */
  const uint8_t* pinned;          /* [n]  Mesh::pins[i].is_some()                */
  const double* pin_targets;      /* [3n] Mesh::pins[i] value where pinned       */
  uint32_t n_triangles;
  const uint32_t* tri_vertices;   /* [3T] */
  const double* tri_inv_dm;       /* [4T] */
  const double* tri_area;         /* [T]  */
  const uint32_t* tri_material;   /* [T]  */
  uint32_t n_hinges;
  const uint32_t* hinge_vertices; /* [4H] shared edge a,b then opposite c,d      */
  const double* hinge_weight;     /* [H]  */
  const double* hinge_across;     /* [2H] */
  const uint32_t* hinge_material; /* [H]  */
  uint32_t n_materials;
  const cv_material* materials;   /* [M]  */
  uint32_t n_stitches;
  const uint32_t* stitch_offsets; /* [S+1] */
  const uint32_t* stitch_vertices;
  const double* stitch_weights;
  const double* stitch_stiffness; /* [S] */
  uint32_t n_probes;
  const uint32_t* probe_offsets;  /* [P+1] */
  const uint32_t* probe_vertices;
  const double* probe_weights;
} cv_garment;

/* SurfaceMesh — src/avatar.rs:9-12: a static collider (fixed OGC geometry). */
typedef struct cv_surface {
  uint32_t n_vertices;
  const double* vertices;         /* [3V] */
  uint32_t n_triangles;
  const uint32_t* triangles;      /* [3F] */
} cv_surface;

/* State — src/model.rs:116-121.  x and v are [3n] f64, updated in place on OK. */
typedef struct cv_state {
  double* x;
  double* v;
  uint32_t n;
  double time;
} cv_state;

/* OgcSettings — src/ogc/world.rs:33-54. */
typedef struct cv_ogc_settings {
  double radius, stiffness, relaxation, refresh_fraction;
} cv_ogc_settings;

/* VbdSettings — src/vbd.rs:18-36.  Fields past `contact` are extensions with
 * no reference counterpart; zero selects the reference behaviour. */
typedef struct cv_settings {
  double dt;
  uint32_t iterations;
  double damping;
  double gravity[3];
  cv_ogc_settings contact;
  /* extensions */
  double residual_tolerance;  /* >0: stop once the post-sweep residual <= tol (max `iterations`) */
  uint32_t flags;             /* CV_FLAG_* */
  uint32_t residual_interval; /* with a tolerance: the residual is evaluated after every k-th
                                 sweep (sweeps k, 2k, ...) and after the last; 0 or 1: every
                                 sweep.  k > 1 trades up to k - 1 extra sweeps for 1/k of the
                                 residual passes (a residual pass costs about a quarter of a
                                 sweep over all colours). */
} cv_settings;

enum {
  CV_FLAG_DIAGNOSTICS_OFF = 1u << 0, /* skip max_seam_gap / extension_percent metrics      */
  CV_FLAG_TRUE_RESIDUAL = 1u << 1,   /* compute the post-sweep residual every iteration     */
  CV_FLAG_FAST_COLORING = 1u << 2,   /* colouring on the device: the static conflict graph
                                        coloured once in smallest-last order, and after
                                        each detection the conflicts that contacts and the
                                        grab add repaired by Jones-Plassmann rounds (hash
                                        priority).  Fewer colours and launches and no host
                                        round trip, but a different Gauss-Seidel order
                                        than the reference's vertex-order colouring
                                        (src/vbd/coloring.rs:5-43), so steps are not the
                                        reference's steps                                 */
  CV_FLAG_WARM_START = 1u << 3       /* learned correction added to the adaptive guess      */
};

/* Grab — src/vbd.rs:54-60. */
typedef struct cv_grab {
  uint32_t vertices[3];
  double weights[3];
  double target[3];
  double stiffness;
} cv_grab;

/* VbdMetrics — src/vbd.rs:61-83 + DetectionProfile (src/ogc/world.rs:14-22).
 * Optional reference fields are NaN when absent. */
typedef struct cv_metrics {
  double time_s, dt_s, step_ms, detection_ms, solve_ms, gpu_compute_ms;
  double refit_ms, candidates_ms, narrow_ms;
  uint64_t candidate_pairs, aabb_rejections;
  uint32_t iterations;
  uint64_t contacts, contact_terms, detection_updates, broad_phase_rebuilds, truncated_vertices;
  double minimum_distance_m, max_seam_gap_m, extension_p95_percent, extension_max_percent;
  double residual_rms_n;
  /* extensions */
  double true_residual_rms_n; /* sqrt(sum_free |g|^2 / 3 n_free) after the last sweep, or NaN */
  uint32_t colors;            /* colour count of the last colouring                    */
  uint32_t iterations_run;    /* sweeps actually run (< iterations with a tolerance)   */
  double sweep_ms;            /* profiling on: summed device time of the sweep launches
                                 stamped with events (sweep_timed_launches of them), else NaN */
  uint64_t sweep_launches;    /* sweep kernel launches in the step                     */
  double sweep_bytes;         /* algorithmic bytes those launches move as stored (f64
                                 positions; f32 or f64 term records)                */
  uint32_t f64;               /* 1 when the sweep ran in f64, 0 for f32                */
  double sweep_bytes_min;     /* the same launches priced at the f32 record sizes:
                                 64 B per free vertex + 72 per triangle, 52 per hinge,
                                 68 per stitch, 56 per contact incidence             */
  uint32_t steps;             /* steps these metrics cover (1 for cv_gpu_step)         */
  uint64_t sweep_timed_launches;  /* launches in sweep_ms: every launch of every 4th
                                     iteration queued while profiling is on           */
  double sweep_timed_bytes_min;   /* sweep_bytes_min of those launches                */
  double sweep_event_overhead_ms; /* profiling on: event interval of an empty launch on
                                     the same stream (the dispatch latency inside each
                                     timed launch's interval), else NaN              */
  uint64_t broad_phase_fallbacks; /* LBVH detections redone with the grid          */
} cv_metrics;

/* ------------------------------------------------------------- solver */
typedef struct cv_solver cv_solver;
typedef struct cv_gpu cv_gpu;

/* VbdSolver::new — src/vbd.rs:101-192.  `avatar` may be NULL. */
cv_status cv_solver_create(const cv_garment* garment, const cv_surface* avatar,
                           const cv_state* state, cv_solver** out);
void cv_solver_destroy(cv_solver* solver);

/* Prediction::record — src/vbd/prediction.rs:43-54.  Installs the adaptive
 * guess history as if `state` had just been committed from `old_velocity`. */
cv_status cv_solver_record_prediction(cv_solver* solver, const double* old_velocity,
                                      const cv_state* state, const cv_settings* settings);

/* Greedy colouring of the last detection — src/vbd/coloring.rs:5-43.  Writes
 * the colour of each moving vertex; returns the colour count in *n_colors. */
cv_status cv_solver_colors(const cv_solver* solver, uint32_t* color_of_vertex, uint32_t n,
                           uint32_t* n_colors);

/* ---------------------------------------------------------------- GPU */
/* GpuVbd::new — src/gpu_vbd.rs:56-124. */
cv_status cv_gpu_create(int device, cv_gpu** out);
void cv_gpu_destroy(cv_gpu* gpu);

/* GpuVbd.check_gradients — src/gpu_vbd.rs:22.  While set, every cv_gpu_step
 * and every step of cv_gpu_run checks, after the detection and the truncated
 * guess, the device probe of each vertex's local objective (gradient and
 * energy) against the same objective evaluated in f64 on the host at the
 * identical rounded positions and targets (src/gpu_vbd.rs:305-379): a
 * gradient error above 3 mN + 0.3 % or an energy error above 1e-5 J + 0.1 %
 * returns CV_ERR_GRADIENT_CHECK ("GPU gradient mismatch at i: ..." / "GPU
 * energy mismatch at i") with the state unchanged; a passing check prints the
 * reference's "GPU/CPU local gradients verified" line to stderr.  The batched
 * step does not check. */
cv_status cv_gpu_set_check_gradients(cv_gpu* gpu, int enabled);

/* GpuVbd::step — src/gpu_vbd.rs:125-472.  `grab` and `metrics` may be NULL. */
cv_status cv_gpu_step(cv_gpu* gpu, cv_solver* solver, cv_state* state,
                      const cv_settings* settings, const cv_grab* grab, cv_metrics* metrics);

/* Device-resident stepping (no reference counterpart): upload `state`, run
 * `steps` steps without host round trips of the state, download once.  Each
 * step has GpuVbd::step semantics; the run is transactional as a whole.
 * `last` describes the last step, except gpu_compute_ms (device time from the
 * first step's first kernel to the last step's last kernel, state already in
 * HBM), iterations_run, sweep_ms, sweep_launches, sweep_bytes,
 * sweep_bytes_min, detection_updates and truncated_vertices, which are totals
 * over the run, and steps = `steps`. */
cv_status cv_gpu_run(cv_gpu* gpu, cv_solver* solver, cv_state* state,
                     const cv_settings* settings, uint32_t steps, cv_metrics* last_metrics);

/* Device probe: per moving vertex (g.x, g.y, g.z, E) of the local objective at
 * the state's geometry, in the sweep's precision — the WGSL `probe` entry
 * (src/vbd.wgsl:142-147).
 * `out` is [4n] f64. */
cv_status cv_gpu_probe(cv_gpu* gpu, cv_solver* solver, const cv_state* state,
                       const cv_settings* settings, const cv_grab* grab, double* out);

/* Detection only, for candidate-set parity: runs the device OGC detection on
 * the given f64 positions (moving part; the avatar is appended) with query
 * radius rq and returns counts.  When the arrays are non-NULL they receive the
 * canonical sorted VF / EE candidate pairs and the contacts. */
typedef struct cv_detection {
  uint64_t n_vf_candidates, n_ee_candidates; /* pairs surviving the padded AABB test */
  uint64_t vertex_face_candidates, edge_edge_candidates; /* pairs with distance < rq */
  uint64_t n_contacts, active_contacts;
  double minimum_distance;
} cv_detection;
cv_status cv_gpu_detect(cv_gpu* gpu, cv_solver* solver, const double* x, double rq,
                        const cv_ogc_settings* settings, cv_detection* out,
                        uint32_t* vf_pairs, uint64_t vf_capacity,
                        uint32_t* ee_pairs, uint64_t ee_capacity,
                        uint32_t* contact_kind_ids, uint64_t contact_capacity,
                        double* contact_eps, double* radii);

/* Colouring of the contact set that the last cv_gpu_detect of this solver
 * found, with the grab (NULL for none), exactly as a step colours it: the
 * reference's greedy in vertex order (src/vbd/coloring.rs:5-43), or with
 * CV_FLAG_FAST_COLORING in `flags` the fast colouring on the device.  Writes
 * the colour of each moving vertex; returns the colour count in *n_colors. */
cv_status cv_gpu_colors(cv_gpu* gpu, cv_solver* solver, const cv_grab* grab, uint32_t flags,
                        uint32_t* color_of_vertex, uint32_t n, uint32_t* n_colors);

/* Newton energy reference on the device — ReferenceSolver::step
 * (src/solver.rs:422-647, SolverSettings of src/solver.rs:15-58 split
 * between cv_newton_settings and the step fields of cv_settings): the f64 Newton-PCG minimiser of one
 * step's objective from `start` (inertial form, membrane, bending, seams)
 * plus the OGC contacts that the last cv_gpu_step or cv_gpu_detect of this
 * solver found, frozen (the reference Newton has no OGC term; with them the
 * objective is the one a contact step minimises).  dt, damping, gravity and
 * the contact radius and stiffness come from `settings`.  Writes the
 * minimiser to x_out [3n] and the solver's statistics to `metrics`; the
 * handle's step state is not touched.  Needs the f64 sweep.  Errors and
 * messages are the reference's ("Invalid solver settings", "Non-finite FEM
 * assembly; step rolled back", "Newton line search failed; ..."). */
typedef struct cv_newton_settings {
  uint32_t newton_iterations;  /* 1..200  */
  uint32_t pcg_iterations;     /* 1..5000 */
  double absolute_tolerance;   /* > 0, RMS gradient (N)  */
  double relative_tolerance;   /* [0, 1), of the initial RMS gradient */
} cv_newton_settings;
typedef struct cv_newton_metrics {  /* the solver fields of Metrics, src/solver.rs:60-75 */
  uint32_t newton_iterations;
  uint32_t pcg_iterations;
  uint32_t line_search_backtracks;
  uint32_t converged;
  double residual_rms_n;
  double tolerance_n;
  double objective_initial_j;
  double objective_final_j;
  double max_regularization;
} cv_newton_metrics;
cv_status cv_gpu_newton(cv_gpu* gpu, cv_solver* solver, const cv_state* start,
                        const cv_settings* settings, const cv_newton_settings* newton,
                        double* x_out, cv_newton_metrics* metrics);

/* Batched uneven scenes (no reference counterpart): one step of `count`
 * independent solvers of any sizes.  Each solver gets exactly the device work
 * of its own cv_gpu_step (same results, bit for bit); the sweeps of one
 * colour index of one iteration of all solvers share one launch.  Each solver
 * commits independently: on failure the status of the first failing solver
 * is returned (cv_last_error names its slot) and only the failing solvers'
 * states are left untouched.  metrics[i] (optional) describes solver i, with
 * gpu_compute_ms the whole batch.  The learned warm start is not available
 * here.  Per-solver host phases (detection read-backs, colouring, staging)
 * run on min(8, hardware threads) worker threads, or CLOTHVBD_HOST_THREADS
 * if set, read when the first batch runs on this context. */
cv_status cv_gpu_step_batch(cv_gpu* gpu, cv_solver* const* solvers, cv_state* const* states,
                            uint32_t count, const cv_settings* settings, cv_metrics* metrics);

/* ------------------------------------------------- learned warm start */
/* No reference counterpart.  A per-vertex network (12 -> 128 -> 128 -> 128
 * -> 3, GELU, BF16 weights and activations, FP32 accumulation on the matrix
 * cores) predicts each step's displacement as a correction to the inertial
 * prediction of Prediction::guess (src/vbd/prediction.rs:15-41), the adaptive
 * x + (v dt + f g dt^2) / (1 + c dt) (f = 1 without a valid history); with
 * CV_FLAG_WARM_START in the settings the step starts from that prediction
 * plus L * correction (the correction optionally averaged over each vertex's
 * one-ring, cv_warm_set_smoothing) and clips it to the OGC balls as it clips
 * the plain guess (src/vbd.rs:316-324); iterations, residuals and energies
 * are unchanged, so the solver itself judges the network, and a network
 * that outputs zero is the plain start.  L = |g| dt^2 / (1 + c dt).
 * Features (12 floats per vertex): offset of the inertial prediction / L (3),
 * v dt / L (3), mean Green strain of the incident triangles along warp and
 * weft x 100 (2), mean one-ring neighbour velocity v dt / L (3), nearest
 * contact distance / query radius (1).
 * Parameters: FP32 vector of CV_WARM_PARAMS values, W0 [12 x 128] row-major,
 * b0, W1 [128 x 128], b1, W2, b2, W3 [128 x 3], b3. */
#define CV_WARM_PARAMS 35075
cv_status cv_warm_set_params(cv_gpu* gpu, const float* params);
/* Passes of one-ring averaging of the correction before it is added (0-16,
 * default 0). */
cv_status cv_warm_set_smoothing(cv_gpu* gpu, int passes);
/* Factor of the correction (0-4, default 1) and whether only its component
 * along the area-weighted vertex normal is added (default 0: all of it). */
cv_status cv_warm_set_correction(cv_gpu* gpu, double scale, int normal_only);
cv_status cv_warm_get_params(cv_gpu* gpu, float* params);
/* Predictions [3n] for features [12n] (host arrays). */
cv_status cv_warm_infer(cv_gpu* gpu, const float* features, uint64_t n, float* out);
/* Gradient [CV_WARM_PARAMS] and mean squared error of the n samples as one
 * batch (the device training kernels' backward pass). */
cv_status cv_warm_gradient(cv_gpu* gpu, const float* features, const float* targets, uint64_t n,
                           float* grad, float* loss);
/* Adam training on the device: `epochs` passes over the n samples in seeded
 * shuffled order, mini-batches of `batch`; epoch_loss[e] (optional) is the
 * epoch's mean squared error. */
cv_status cv_warm_train(cv_gpu* gpu, const float* features, const float* targets, uint64_t n,
                        uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* epoch_loss);
/* Allocation-free variant of cv_warm_infer; it has no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the features and predictions are staged in a static device array of
 * CV_WARM_NOALLOC_MAX_ROWS rows (15 MiB, one per GPU, held until the process
 * ends) instead of buffers allocated per call, and no call allocates or frees
 * device memory.  Same results as cv_warm_infer; n above the limit returns
 * CV_ERR_CAPACITY.  The array is shared by every call on a GPU: calls that
 * overlap on one GPU (any handles, any threads) race. */
#define CV_WARM_NOALLOC_MAX_ROWS 262144
cv_status cv_warm_infer_noalloc(cv_gpu* gpu, const float* features, uint64_t n, float* out);
/* The features [12n] the step would feed the network at `state`. */
cv_status cv_warm_features(cv_gpu* gpu, cv_solver* solver, const cv_state* state,
                           const cv_settings* settings, float* features);

/* ------------------------------------------------------------ errors */
/* Copies the last error message of `gpu` (or, for NULL, the calling thread's
 * last create error) into buf; returns the full message length. */
size_t cv_last_error(const cv_gpu* gpu, char* buf, size_t n);
/* Last error of a solver call (cv_solver_record_prediction). */
size_t cv_solver_last_error(const cv_solver* solver, char* buf, size_t n);

/* Sweep precision: 64 (default; reproduces the f64 reference step to
 * round-off) or 32 (faster: positions stay f64, each vertex solve runs in f32
 * on coordinates relative to that vertex, so f32 rounds one-ring lengths;
 * the step then differs from the f64 step by about what the reference step
 * itself moves when its input moves by 1e-10 m). */
cv_status cv_gpu_set_precision(cv_gpu* gpu, int bits);

/* Begin/end events around the sweep launches of every 4th iteration;
 * metrics.sweep_ms reports their summed duration, sweep_timed_launches their
 * count and sweep_timed_bytes_min their algorithmic bytes.  Enabling it
 * measures the event interval of an empty launch on the same stream once
 * (metrics.sweep_event_overhead_ms), the part of each timed interval that is
 * dispatch rather than kernel.  Those empty launches (kernel k_empty) open
 * the profiled window in a kernel trace and disabling it launches one more to
 * close it, so a trace can be cut to the window. */
cv_status cv_gpu_set_profiling(cv_gpu* gpu, int enabled);

/* Broad phase of the OGC detection: CV_BROAD_PHASE_GRID (default; a hashed
 * uniform grid of cell size max(cell, 2 padded radii), the candidate walk of
 * src/ogc/world.rs:407-470) or CV_BROAD_PHASE_LBVH (a linear BVH over the
 * same f64 boxes, rebuilt at each step's first detection and refitted at its
 * refreshes; a walk deeper than its stack redoes that detection with the
 * grid, counted in metrics.broad_phase_fallbacks).  Both produce the same
 * candidate pairs; only their cost differs. */
#define CV_BROAD_PHASE_GRID 0
#define CV_BROAD_PHASE_LBVH 1
cv_status cv_gpu_set_broad_phase(cv_gpu* gpu, int kind);

/* Lanes per vertex of the sweep kernel (8, 16, 32 or 64); the autotuner sets it. */
cv_status cv_gpu_set_lanes(cv_gpu* gpu, int lanes);

/* Threads per block of the batched step's shared sweep launches (64, 128 or
 * 256; default 256); the autotuner sets it with the lanes.  Results do not
 * depend on it.  No counterpart in the reference: its batched dispatch
 * (src/gpu_vbd.rs:383-405) has a fixed workgroup size. */
cv_status cv_gpu_set_batch_block(cv_gpu* gpu, int threads);

/* Library version string. */
const char* cv_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CLOTHVBD_H */
