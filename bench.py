#!/usr/bin/env python3
"""Throughput of the device VBD step on the sewn-garment workload at a fixed residual.

Workload (default): two 0.6 m x 0.7 m panels of seeded Poisson-disk meshes
(99,287 vertices, 196,402 triangles) sewn by side and shoulder seams (1,368
penalty stitches, 1e4 N/m), wrapped on half-ellipses 1 cm outside a static
capsule-union mannequin (torso r 0.15 m and two arms, 19,488 triangles);
StVK mu 2e4 / lambda 8e3 N/m, bending 2e-5 N m, 0.2 kg/m^2; OGC contact
radius 1 mm (below the 1.2 mm rest distance of non-adjacent primitives of the
mesh), stiffness 1e5 N/m, self- and body contact on; dt 1/60 s.  Synthetic,
seeded geometry (no external data).

One step is the complete VbdSolver::step of the reference (src/vbd.rs:250-422):
prediction (the adaptive guess, src/vbd/prediction.rs:15-41), OGC detection
and colouring, coloured Gauss-Seidel sweeps with the ball-truncated line
search, mid-step re-detection, commit.  The sweep computes in f32
(`--precision 32`, the default here: positions stay f64 and each vertex's
local solve runs in f32 on coordinates relative to that vertex; `--precision
64` is the f64 sweep that reproduces the f64 reference step to round-off).
Iterations run to a fixed true
residual: after every `--residual-interval`-th sweep (4: sweeps 4, 8, ...,
and the last) the post-sweep residual sqrt(sum |g|^2 / 3 n) of the step's
objective (src/solver.rs:229-231 form) is computed, and the step ends once
it is at most `--tol` (0.02 N, about what the configuration's 40 iterations
reach on this garment in contact) or after the configuration's 40
iterations, whichever comes first.  (A residual pass costs about a quarter
of a sweep over all colours; checking every 4th sweep costs at most 3 extra
sweeps and saves three passes in four.)  (Beyond that the residual stops
falling: vertices drift onto their trust-region boundaries and it rises
again, so a larger cap only degrades the trajectory.)

Phases: `--settle` (30) untimed steps of the configuration's own 40-iteration
solve bring the garment from its start onto the body (the panels start 1 cm
out; contacts appear within ~10 steps), then `--warmup` untimed steps and the
`--steps` timed steps continue the trajectory at the fixed residual, so the
timed window is in contact whatever `--steps` is.  `value` is steps/s from
HIP events on the solver's stream around the device work, state resident in
HBM; the rate including the state's upload/download is reported beside it.

Warm start vs plain: after the timed window, from `--compare` successive
states of the same trajectory, the iterations each start needs to reach the
tolerance (at most `--compare-cap`, 100): the plain start is the reference's
adaptive guess with its prediction history, the warm start that guess plus
the committed network's correction (CV_FLAG_WARM_START;
clothcraft_amd/warm_params.npy, trained by clothcraft_amd/train_warm.py).

roofline: the sweep kernel (k_sweep), algorithmic bytes per launch at the f32
minimum record sizes (64 B per free vertex + 72 per triangle, 52 per hinge,
68 per stitch, 56 per contact incidence), divided by its mean duration over
the sampled launches (every launch of every 4th iteration of the timed
window): the interval between the HIP events recorded around each of those
launches on the solver's stream (hipExtLaunchKernelGGL start/stop events).
The interval of an empty launch on the same stream is reported beside it.
Switching profiling on and off brackets the timed window with empty kernels
(k_empty), so a rocprofv3 trace or counter run of this command can be cut to
the window (`--window-stats`); `traffic` is the measured HBM bytes per
launch (rocprofv3 FETCH_SIZE + WRITE_SIZE over the window, calibrated by a
copy kernel) from the committed counter runs of this command (same workload,
tolerance, precision, lanes and broad phase), null for other arguments.

cpu_baseline: the f64 CPU restatement of VbdSolver::step (OpenMP within a
colour) on the host, from the settled state with the same settings, on the
process's CPU share (16 threads where the machine gives each GPU 16) and on one thread: steps
capped at two iteration counts give each a per-step fixed cost and a cost per
iteration (true residual included), evaluated at the GPU's mean iterations
per step (a full step to the residual takes minutes on the CPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
WARM_PARAMS = os.path.join(ROOT, "clothcraft_amd", "warm_params.npy")


def _traffic(kernel, workload, argv_key):
    """HBM bytes per sweep launch from the committed rocprofv3 counter runs of
    this command (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known
    size), or None when the work-defining arguments differ from the profiled
    ones."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return d.get("bytes_per_launch") if d.get("args") == argv_key else None


def window_stats(path, real=0):
    """Per-kernel mean duration (a kernel trace) or per-kernel, per-counter
    mean (a counter collection) over the launches of the timed window: the
    launches after the burst of empty kernels that switching profiling on
    issues (its event calibration) and before the empty kernel that switching
    it off issues."""
    import csv
    with open(path) as f:
        rows = list(csv.DictReader(f))
    trace = "Start_Timestamp" in rows[0] and "Counter_Name" not in rows[0]
    key = (lambda r: int(r["Start_Timestamp"])) if trace else (lambda r: int(r["Dispatch_Id"]))
    rows.sort(key=key)
    empty = [i for i, r in enumerate(rows) if "k_empty" in r["Kernel_Name"]]
    if not empty:
        raise SystemExit("no profiling window (k_empty launches) in " + path)
    # The window opens after the first burst of consecutive empty launches and
    # closes at the next empty launch.
    first = empty[0]
    while first + 1 < len(rows) and "k_empty" in rows[first + 1]["Kernel_Name"]:
        first += 1
    later = [i for i in empty if i > first]
    last = later[0] if later else len(rows)
    out = {}
    for r in rows[first + 1:last]:
        name = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        if trace:
            d = out.setdefault(name, {"launches": 0, "total_ns": 0, "durations_ns": []})
            d["launches"] += 1
            d["total_ns"] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            d["durations_ns"].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        else:
            d = out.setdefault(name, {}).setdefault(r["Counter_Name"], {"dispatches": 0, "sum": 0.0})
            d["dispatches"] += 1
            d["sum"] += float(r["Counter_Value"])
    res = {"source": path, "window_rows": last - first - 1, "kernels": out}
    if trace and last - first > 1:
        # Device time between the window's dispatches: the gaps between one
        # kernel's end and the next one's start (dependent launches; host
        # waits show as the long ones).
        iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows[first + 1:last])
        busy, end, gaps = 0, iv[0][0], []
        for s0, e0 in iv:
            if s0 > end:
                gaps.append(s0 - end)
            busy += max(0, e0 - max(s0, end))
            end = max(end, e0)
        g = np.array(gaps, np.float64)
        res["span_ms"] = (end - iv[0][0]) / 1e6
        res["busy_ms"] = busy / 1e6
        res["gaps_ms"] = {"under_10us": float(g[g < 1e4].sum() / 1e6),
                          "10_to_100us": float(g[(g >= 1e4) & (g < 1e5)].sum() / 1e6),
                          "over_100us": float(g[g >= 1e5].sum() / 1e6),
                          "count_over_100us": int((g >= 1e5).sum())}
    for name, d in out.items():
        if trace:
            d["mean_us"] = d["total_ns"] / d["launches"] / 1e3
            dur = np.sort(np.array(d.pop("durations_ns"), np.float64))
            d["p10_p50_p90_us"] = [float(v) for v in np.percentile(dur, [10, 50, 90]) / 1e3]
            if real and name.startswith("k_sweep") and 0 < real <= len(dur):
                d["mean_us_of_longest"] = {"launches": real, "mean_us": float(dur[-real:].mean() / 1e3)}
        else:
            for c in d.values():
                c["mean"] = c["sum"] / c["dispatches"]
                if real and name.startswith("k_sweep") and 0 < real <= c["dispatches"]:
                    # the no-op launches read the iteration flag and exit, so
                    # the window's counts belong to the real launches
                    c["per_real_launch"] = c["sum"] / real
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_share():
    """Threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS
    (16 where the machine gives each GPU 16 threads)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


def cpu_step_seconds(sc, state, settings, threads, steps=1):
    """Seconds per oracle f64 step on the host from `state`, and the mean
    iterations those steps ran."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU restatement: baseline only, never the product
    st = state.copy()
    ref = O.Solver(sc.garment, sc.avatar, st, threads=threads)
    its, t0 = 0, time.perf_counter()
    for _ in range(steps):
        m = ref.step(st, settings)
        its += m["iterations_run"]
    return (time.perf_counter() - t0) / steps, its / steps


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=30,
                    help="untimed steps of the configuration's own solve before the warm-up")
    ap.add_argument("--tol", type=float, default=0.02, help="fixed true residual (N, RMS per coordinate)")
    ap.add_argument("--max-iterations", type=int, default=0,
                    help="iteration cap of a step (0: the configuration's own count, 40 for the garment)")
    ap.add_argument("--compare-cap", type=int, default=100,
                    help="iteration cap of the warm-vs-plain comparison steps")
    ap.add_argument("--residual-interval", type=int, default=4,
                    help="evaluate the true residual after every k-th sweep (and the last)")
    ap.add_argument("--fixed-iterations", action="store_true",
                    help="run the configuration's fixed iteration count instead of the fixed residual")
    ap.add_argument("--compare", type=int, default=6, help="states of the warm-vs-plain comparison (0: off)")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64),
                    help="sweep arithmetic (32: f32 solves in each vertex's frame over f64 positions, the "
                         "default here; 64: f64, the library's default)")
    ap.add_argument("--lanes", type=int, default=32, choices=(8, 16, 32, 64),
                    help="lanes per vertex of the sweep kernel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nx", type=int, default=128, help="sheet_on_sphere: quads per side")
    ap.add_argument("--workload", default="sewn_garment",
                    choices=("sewn_garment", "sheet_on_sphere", "batch64", "warm_infer", "newton_sheet"),
                    help="sheet_on_sphere: the 128x128 sheet dropped on an icosphere (20 iterations, OGC "
                         "r 2 mm); batch64: 64 independent sheets and sewn tubes of 1k-50k vertices in "
                         "batched steps; warm_infer: the warm-start network on 4 Mi rows; newton_sheet: "
                         "the device Newton energy reference of one sheet step in contact")
    ap.add_argument("--fast-coloring", action="store_true",
                    help="fast colouring on the device (fewer colours; not the reference's colouring order)")
    ap.add_argument("--broad-phase", default="grid", choices=("grid", "lbvh"),
                    help="OGC broad phase: hashed grid or linear BVH (same candidate sets)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch64: use --lanes instead of timing the lanes per vertex first")
    ap.add_argument("--window-real", type=int, default=0, metavar="N",
                    help="with --window-stats: also the mean of the N longest k_sweep launches (a "
                         "kernel trace) or each counter's window sum over N launches (a counter run); N "
                         "is the bench's real launches, the others are no-ops queued behind a converged "
                         "step")
    ap.add_argument("--window-stats", metavar="CSV",
                    help="only summarise a rocprofv3 kernel_trace.csv or counter_collection.csv of this "
                         "command over the timed window (the launches between the profiling switch's "
                         "empty kernels)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.window_stats:
        print(json.dumps(window_stats(args.window_stats, args.window_real)))
        return
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the step has no multi-GPU path in this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "warm_infer":
        return bench_warm_infer(args)
    if args.workload == "newton_sheet":
        return bench_newton(args)
    return bench_step(args)


def bench_step(args):
    from clothcraft_amd import GpuVbd, VbdSolver, garments, scenes  # fails loudly if not built
    from clothcraft_amd.abi import CV_FLAG_FAST_COLORING, CV_FLAG_TRUE_RESIDUAL, CV_FLAG_WARM_START

    if args.workload == "sewn_garment":
        sc = garments.sewn_garment_on_body()
        workload = (f"sewn garment on a body: {sc.garment.n} vertices, {len(sc.garment.tri_vertices)} "
                    f"triangles, {len(sc.garment.stitch_stiffness)} seam stitches (1e4 N/m), panels 1 cm "
                    f"outside a capsule-union body ({len(sc.avatar.triangles)} triangles), OGC r 1 mm k 1e5, "
                    f"dt 1/60, self- and body contact on")
    else:
        sc = scenes.sheet_on_sphere(nx=args.nx)
        workload = (f"sheet on sphere {args.nx}x{args.nx} quads: {sc.garment.n} vertices, "
                    f"{len(sc.garment.tri_vertices)} triangles, icosphere r 0.25 m "
                    f"({len(sc.avatar.triangles)} triangles), OGC r 2 mm k 1e5, dt 1/60, self-contact on")
    base = dict(sc.settings)
    if args.fast_coloring:
        base["flags"] = base.get("flags", 0) | CV_FLAG_FAST_COLORING
    if args.fixed_iterations:
        run_s = dict(base, flags=base.get("flags", 0) | CV_FLAG_TRUE_RESIDUAL)
    else:
        cap = args.max_iterations or base["iterations"]
        run_s = dict(base, iterations=cap, residual_tolerance=args.tol,
                     residual_interval=args.residual_interval,
                     flags=base.get("flags", 0) | CV_FLAG_TRUE_RESIDUAL)
    gpu = GpuVbd(0)
    gpu.set_precision(args.precision)
    gpu.set_lanes(args.lanes)
    gpu.set_broad_phase(args.broad_phase)
    st = sc.state.copy()
    sol = VbdSolver(sc.garment, sc.avatar, st)
    n_free = int(np.count_nonzero(sc.garment.pinned == 0))
    log(f"{sc.name}: {sc.garment.n} vertices; settle {args.settle} steps at "
        f"{base['iterations']} iterations, warm-up {args.warmup} steps")
    if args.settle:
        m = gpu.run(sol, st, base, args.settle)
        log(f"settled: {m['contact_terms']} contact terms, minimum distance {m['minimum_distance_m']:.2e} m")
    if args.warmup:
        gpu.run(sol, st, run_s, args.warmup)
    settled = st.copy()

    log(f"timed: {args.steps} steps")
    gpu.set_profiling(True)
    tot = dict(gpu_compute_ms=0.0, iterations_run=0, sweep_launches=0, sweep_ms=0.0, sweep_bytes=0.0,
               sweep_bytes_min=0.0, detection_updates=0, sweep_timed_launches=0, sweep_timed_bytes_min=0.0)
    done, m, contacts, reached = 0, None, [], 0
    t0 = time.perf_counter()
    while done < args.steps:
        # Chunks of device-resident steps (run() keeps the prediction history
        # across calls, so results equal one long run).
        k = min(5, args.steps - done)
        m = gpu.run(sol, st, run_s, k)
        for key in tot:
            tot[key] += m[key]
        contacts.append(m["contact_terms"])
        reached += int(m["true_residual_rms_n"] <= args.tol)
        done += k
        log(f"{done}/{args.steps} steps, last step {m['iterations_run']} iterations" if k == 1 else
            f"{done}/{args.steps} steps")
    wall_s = time.perf_counter() - t0
    gpu.set_profiling(False)
    if not np.all(np.isfinite(st.x)):
        raise RuntimeError("non-finite state after the timed steps")
    last = m
    m = dict(m, **tot)
    K = args.steps
    dev_s = m["gpu_compute_ms"] / 1e3
    vi = n_free * m["iterations_run"]
    launches = max(m["sweep_launches"], 1)
    timed = max(m["sweep_timed_launches"], 1)
    # Kernel duration: the HIP event interval of each sampled launch (the
    # start and stop events of hipExtLaunchKernelGGL).  It agrees with a
    # rocprofv3 kernel trace of the same launches; the interval of an empty
    # launch on the same stream is reported beside it but not subtracted (a
    # queued launch's interval does not contain that much dispatch time).
    overhead_s = m["sweep_event_overhead_ms"] / 1e3
    avg_interval_s = m["sweep_ms"] / 1e3 / timed
    avg_launch_s = avg_interval_s
    bytes_per_launch = m["sweep_timed_bytes_min"] / timed
    achieved = bytes_per_launch / avg_launch_s / 1e9
    sweep_s = avg_interval_s * launches

    # The reference's call pattern: one GpuVbd::step per frame, state in and
    # out through host memory and the metrics each step (wall time per call).
    api_st = st.copy()
    api_sol = VbdSolver(sc.garment, sc.avatar, api_st)
    api_ms = []
    for _ in range(12):
        t1 = time.perf_counter()
        gpu.step(api_sol, api_st, run_s)
        api_ms.append(1e3 * (time.perf_counter() - t1))
    api_ms = np.array(api_ms[2:])

    compare = None
    if args.compare and not args.fixed_iterations:
        # One more step of the trajectory's own solver gives a state with its
        # previous velocity (the adaptive guess's history).
        pv = st.v.copy()
        gpu.step(sol, st, run_s)
        compare = compare_starts(gpu, sc, st, pv, dict(run_s, iterations=args.compare_cap), args.compare)
        if compare is not None:
            compare["trajectory"] = learned_trajectory(gpu, sc, settled, run_s, args, K)
            compare["trajectory"]["plain_steps_per_s"] = K / dev_s
            compare["trajectory"]["plain_iterations_per_step"] = m["iterations_run"] / K
            compare["trajectory"]["plain_detections_per_step"] = m["detection_updates"] / K

    kernel = "k_sweep"
    # The arguments that define the per-launch work of the sweep (the step
    # count and warm-up only change how many steps are timed).
    argv_key = {"workload": args.workload, "settle": args.settle, "tol": args.tol,
                "residual_interval": args.residual_interval,
                "max_iterations": args.max_iterations, "precision": args.precision, "lanes": args.lanes,
                "fixed_iterations": args.fixed_iterations, "fast_coloring": args.fast_coloring,
                "broad_phase": args.broad_phase}
    out = {
        "metric": "steps/s + vertex-iters/s at fixed residual; warm-start vs plain iters-to-tol",
        "value": K / dev_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dev_s / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f32",
        "data": "synthetic (seeded Poisson-disk panels, capsule body)" if args.workload == "sewn_garment"
                else "synthetic (seeded sheet and icosphere geometry)",
        "config": {
            "workload": workload,
            "steps_timed": f"{args.settle + args.warmup + 1}..{args.settle + args.warmup + K} of the trajectory "
                           f"(in contact: {min(contacts)}-{max(contacts)} contact terms at the chunk ends)",
            "iteration_mode": ("fixed iterations" if args.fixed_iterations else
                               f"to true residual <= {args.tol} N (evaluated after every "
                               f"{args.residual_interval}th sweep and the last) or the configuration's "
                               f"{run_s['iterations']} iterations, whichever comes first"),
            "coloring": ("fast: static smallest-last, device repair (CV_FLAG_FAST_COLORING)"
                         if args.fast_coloring else "reference vertex-order greedy"),
            "lanes_per_vertex": args.lanes,
            "broad_phase": args.broad_phase,
        },
        "iterations_per_step": m["iterations_run"] / K,
        "chunk_end_steps_at_the_residual": f"{reached} of {len(contacts)}",
        "vertex_iters_per_s": vi / dev_s,
        "vertex_iters_per_s_sweep_only": vi / sweep_s if sweep_s > 0 else None,
        "steps_per_s_wall_incl_transfers": K / wall_s,
        "step_api_wall_ms": {"median": float(np.median(api_ms)), "p95": float(np.percentile(api_ms, 95)),
                             "calls": int(len(api_ms)), "warmup_calls": 2},
        "true_residual_rms_n_last_step": last["true_residual_rms_n"],
        "residual_rms_n_last_step": last["residual_rms_n"],
        "contacts_last_step": last["contact_terms"],
        "colors_last_step": last["colors"],
        "candidate_pairs_last_step": last["candidate_pairs"],
        "broad_phase_tests_last_step": last["candidate_pairs"] + last["aabb_rejections"],
        "detections": m["detection_updates"],
        "sweep_share_of_step": sweep_s / dev_s,
        "warm_vs_plain": compare,
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": _traffic(kernel, args.workload, argv_key),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_us": 1e6 * avg_launch_s,
            "avg_launch_interval_us_hip_events": 1e6 * avg_interval_s,
            "empty_launch_interval_us_hip_events": 1e6 * overhead_s,
            "launches": m["sweep_launches"],
            "launches_timed": m["sweep_timed_launches"],
            "bytes_per_launch_in_sweep_precision": m["sweep_bytes"] / launches,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, settled, run_s, n_free, m["iterations_run"] / K)
    print(json.dumps(out), flush=True)


def compare_starts(gpu, sc, st, prev_v, run_s, count):
    """From `count` successive states of the plain trajectory: iterations to
    the fixed residual with the plain adaptive start (prediction history
    installed) and with the learned start, their device times and detections."""
    from clothcraft_amd import VbdSolver
    from clothcraft_amd.abi import CV_FLAG_WARM_START
    if not os.path.exists(WARM_PARAMS):
        log("no trained warm-start network; comparison skipped")
        return None
    gpu.warm.set_params(np.load(WARM_PARAMS))
    gpu.warm.set_smoothing(1)  # the correction averaged over each one-ring once
    run_s = dict(run_s, residual_interval=1)  # exact iteration counts
    warm_s = dict(run_s, flags=run_s["flags"] | CV_FLAG_WARM_START)
    rows = []
    st, pv = st.copy(), prev_v.copy()
    for k in range(count):
        plain = st.copy()
        ps = VbdSolver(sc.garment, sc.avatar, plain)
        ps.record_prediction(pv, plain, run_s)
        mp = gpu.step(ps, plain, run_s)
        warm = st.copy()
        mw = gpu.step(VbdSolver(sc.garment, sc.avatar, warm), warm, warm_s)
        rows.append((mp["iterations_run"], mw["iterations_run"], mp["gpu_compute_ms"], mw["gpu_compute_ms"],
                     mp["true_residual_rms_n"], mw["true_residual_rms_n"], mp["detection_updates"],
                     mw["detection_updates"]))
        log(f"compare {k + 1}/{count}: iterations plain {mp['iterations_run']} warm {mw['iterations_run']}")
        pv, st = st.v.copy(), plain
    r = np.array(rows, dtype=np.float64)
    return {"states": count, "plain_iterations_mean": float(r[:, 0].mean()),
            "warm_iterations_mean": float(r[:, 1].mean()),
            "iterations_gain": float(r[:, 0].mean() / max(r[:, 1].mean(), 1e-9)),
            "plain_ms_mean": float(r[:, 2].mean()), "warm_ms_mean": float(r[:, 3].mean()),
            "plain_final_residual_mean": float(r[:, 4].mean()), "warm_final_residual_mean": float(r[:, 5].mean()),
            "plain_detections_mean": float(r[:, 6].mean()), "warm_detections_mean": float(r[:, 7].mean()),
            "network": "clothcraft_amd/warm_params.npy (clothcraft_amd/train_warm.py), correction "
                       "averaged over the one-ring once"}


def learned_trajectory(gpu, sc, settled, run_s, args, K):
    """The timed steps again from the settled state with the learned start
    (the network compare_starts set up): steps/s, iterations and detections
    per step, timed like the plain run."""
    from clothcraft_amd import VbdSolver
    from clothcraft_amd.abi import CV_FLAG_WARM_START
    warm_s = dict(run_s, flags=run_s["flags"] | CV_FLAG_WARM_START)
    st = settled.copy()
    sol = VbdSolver(sc.garment, sc.avatar, st)
    gpu.run(sol, st, warm_s, max(args.warmup, 1))
    gpu.set_profiling(True)
    ms = its = dets = 0.0
    done = 0
    while done < K:
        k = min(5, K - done)
        m = gpu.run(sol, st, warm_s, k)
        ms += m["gpu_compute_ms"]
        its += m["iterations_run"]
        dets += m["detection_updates"]
        done += k
    gpu.set_profiling(False)
    return {"steps": f"{K} steps from the settled state ({max(args.warmup, 1)} warm-up) with the learned start",
            "learned_steps_per_s": K / (ms / 1e3), "learned_iterations_per_step": its / K,
            "learned_detections_per_step": dets / K}


def cpu_fit(sc, state, run_s, threads, lo, hi):
    """Seconds of one oracle step from `state` capped at lo and at hi
    iterations (with a tolerance it never reaches, the true residual computed
    after the last sweep only): the step's fixed cost (one residual pass
    included) and its cost per sweep.  The model counts one residual pass per
    step where the device runs one per `residual_interval` sweeps, so it
    slightly favours the CPU."""
    cap = lambda k: dict(run_s, iterations=k, residual_tolerance=1e-30, residual_interval=k)
    t_lo, _ = cpu_step_seconds(sc, state, cap(lo), threads)
    t_hi, _ = cpu_step_seconds(sc, state, cap(hi), threads)
    per_it = max((t_hi - t_lo) / (hi - lo), 1e-9)
    return max(t_lo - lo * per_it, 0.0), per_it, t_lo, t_hi


def cpu_baseline(sc, state, run_s, n_free, gpu_iters_per_step):
    share = cpu_share()
    log(f"cpu baseline: oracle f64 step, {share} threads and 1 thread")
    f_all, it_all, a_lo, a_hi = cpu_fit(sc, state, run_s, share, 2, 4)
    f_one, it_one, o_lo, o_hi = cpu_fit(sc, state, run_s, 1, 1, 2)
    t_all = f_all + it_all * gpu_iters_per_step
    t_one = f_one + it_one * gpu_iters_per_step
    return {
        "value": 1.0 / t_all,
        "unit": "steps/s",
        "cores": share,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"f64 C++ restatement of VbdSolver::step (OpenMP within a colour) from the settled state, "
                  f"same settings, on {share} threads (the process's CPU share): steps capped at 2 and 4 "
                  f"iterations took {a_lo:.2f} s and {a_hi:.2f} s, i.e. {f_all:.2f} s per step + {it_all:.2f} s "
                  f"per iteration, evaluated at the GPU's {gpu_iters_per_step:.1f} iterations per step",
        "vertex_iters_per_s": n_free / it_all,
        "single_core": {"value": 1.0 / t_one, "unit": "steps/s", "cores": 1,
                        "sample": f"one thread: steps capped at 1 and 2 iterations took {o_lo:.2f} s and "
                                  f"{o_hi:.2f} s ({f_one:.2f} s per step + {it_one:.2f} s per iteration), "
                                  f"evaluated at {gpu_iters_per_step:.1f} iterations per step",
                        "vertex_iters_per_s": n_free / it_one},
    }


def bench_warm_infer(args):
    """Warm-start network inference (BF16 MFMA) on 4 Mi feature rows per
    call.  The line's rate includes the host transfers of cv_warm_infer; the
    kernel's own duration comes from a kernel trace of this command."""
    from clothcraft_amd import GpuVbd, warmstart
    n = 1 << 22
    gpu = GpuVbd(0)
    gpu.warm.set_params(warmstart.init_params(0))
    X = np.random.default_rng(0).normal(size=(n, 12)).astype(np.float32)
    for _ in range(max(args.warmup, 1)):
        gpu.warm.infer(X)
    K = args.steps
    t0 = time.perf_counter()
    for k in range(K):
        gpu.warm.infer(X)
    wall = time.perf_counter() - t0
    flop = 69376.0 * n  # 2 (12*128 + 128*128 + 128*128 + 128*3) per row
    print(json.dumps({
        "metric": "warm-start inference rows/s (host transfers included)", "value": K * n / wall,
        "unit": "rows/s", "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic normal features", "config": {"workload": f"MLP 12-128-128-128-3 GELU, {n} rows per call"},
        "algorithmic_flop_per_call": flop}), flush=True)


def bench_newton(args):
    """The Newton energy reference (cv_gpu_newton, the reference's Newton
    settings: 30 iterations, 250 PCG iterations) of one step of the
    128x128 sheet on the sphere, in contact after 140 device steps, with the
    step's contacts frozen.  A "step" of this line is one Newton solve."""
    from clothcraft_amd import GpuVbd, VbdSolver, scenes
    sc = scenes.sheet_on_sphere(nx=args.nx)
    s = sc.settings
    gpu = GpuVbd(0)
    st = sc.state.copy()
    gpu.run(VbdSolver(sc.garment, sc.avatar, st), st, s, 140)
    sol = VbdSolver(sc.garment, sc.avatar, st)
    gpu.step(sol, st.copy(), s)
    for _ in range(args.warmup):
        gpu.newton(sol, st, s)
    K = args.steps
    t0 = time.perf_counter()
    for _ in range(K):
        _, m = gpu.newton(sol, st, s)
    wall = time.perf_counter() - t0
    print(json.dumps({
        "metric": "device Newton energy reference solves/s", "value": K / wall, "unit": "solves/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (procedural sheet and sphere)",
        "config": {"workload": f"Newton-PCG of one step of the {args.nx}x{args.nx} sheet on the sphere "
                               f"({sc.garment.n} vertices) in contact, contacts frozen"},
        "newton_iterations": m["newton_iterations"], "pcg_iterations": m["pcg_iterations"],
        "gradient_passes_per_s": K * (2 * m["pcg_iterations"] + m["newton_iterations"] + 1) / wall,
        "converged": bool(m["converged"])}), flush=True)


def bench_batch(args):
    """64 independent instances stepped together (cv_gpu_step_batch)."""
    from clothcraft_amd import GpuVbd, VbdSolver, autotune, garments
    scs = garments.batched_scenes(64, seed=3)
    settings = scs[0].settings
    gpu = GpuVbd(0)
    gpu.set_precision(args.precision)
    tuned = None
    if not args.no_autotune:
        tuned = autotune.tune_lanes(gpu, scs, settings, repeats=1)
    else:
        gpu.set_lanes(args.lanes)
    states = [sc.state.copy() for sc in scs]
    sols = [VbdSolver(sc.garment, sc.avatar, st) for sc, st in zip(scs, states)]
    print(f"[bench] batch of {len(scs)}: {sum(sc.garment.n for sc in scs)} vertices", file=sys.stderr, flush=True)
    if args.warmup:
        ws = [st.copy() for st in states]
        wsol = [VbdSolver(sc.garment, sc.avatar, st) for sc, st in zip(scs, ws)]
        for _ in range(args.warmup):
            gpu.step_batch(wsol, ws, settings)
    K = args.steps
    dev_ms, iters = 0.0, 0
    t0 = time.perf_counter()
    det = {"detection_updates": 0, "candidate_pairs": 0, "contact_terms": 0, "colors_max": 0}
    for k in range(K):
        ms = gpu.step_batch(sols, states, settings)
        dev_ms += ms[0]["gpu_compute_ms"]
        iters += sum(m["iterations_run"] * sc.garment.n for m, sc in zip(ms, scs))
        det["detection_updates"] += sum(m["detection_updates"] for m in ms)
        det["candidate_pairs"] += sum(m["candidate_pairs"] for m in ms)
        det["contact_terms"] += sum(m["contact_terms"] for m in ms)
        det["colors_max"] = max(det["colors_max"], max(m["colors"] for m in ms))
        print(f"[bench] batch step {k + 1}/{K}: {ms[0]['gpu_compute_ms']:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    n_total = sum(sc.garment.n for sc in scs)
    out = {
        "metric": "steps/s + vertex-iters/s at fixed residual; warm-start vs plain iters-to-tol",
        "value": K / (dev_ms / 1e3), "unit": "batch steps/s (64 instances each)", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": dev_ms / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == 64 else "f32",
        "data": "synthetic (seeded sheets and sewn tubes)",
        "config": {"workload": f"64 independent instances, {n_total} vertices in all (log-uniform 1k-50k), "
                               f"sheets over icospheres or capsules and sewn tubes around capsules, "
                               f"{settings['iterations']} iterations, dt 1/60, OGC r 2 mm k 1e5"},
        "instance_steps_per_s": 64 * K / (dev_ms / 1e3),
        "vertex_iters_per_s": iters / (dev_ms / 1e3),
        "steps_per_s_wall_incl_transfers": K / wall,
        "lanes": tuned if tuned else args.lanes,
        "per_batch_step": {k: (v / K if k != "colors_max" else v) for k, v in det.items()},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
