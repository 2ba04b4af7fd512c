#!/usr/bin/env python3
"""Throughput of the device VBD step on the sheet-on-sphere workload.

Workload: a 1 m x 1 m sheet of 128 x 128 quads (16,641 vertices, 32,768
triangles, 32,258 hinges) dropped from 0.3 m onto a static icosphere collider
(radius 0.25 m, subdivision 4, 5,120 triangles); StVK mu 2e4 N/m scaled per
triangle by a seeded U[0.8, 1.2], lambda 8e3 N/m, bending 2e-5 N m, 0.2 kg/m^2;
OGC contact radius 2 mm, stiffness 1e5 N/m, self-contact on; 20 VBD iterations
per step, dt 1/60 s.  Synthetic, seeded geometry (no external data).

One step is the complete VbdSolver::step of the reference (src/vbd.rs:194-436):
prediction, OGC detection and colouring, 20 coloured Gauss-Seidel sweeps with
the ball-truncated line search, mid-step re-detection, commit.  `--warmup`
steps run untimed on a throw-away copy of the start state (buffers, code
objects); the timed region then runs the workload's `--steps` (300) steps of
the drop from the start state, device-resident.  The OGC trust regions
(0.45 x 5.5 mm self-distance) hold the fall to about 2.5 mm per step, so the
sheet reaches the sphere near step 120 and the rest of the drop is in
contact.  `value` is steps/s from HIP events on the solver's stream around
the device work, state already in HBM; the wall-clock rate including the one
upload/download of the state is reported beside it.

roofline: the sweep kernel (k_sweep), algorithmic bytes per launch at the f32
minimum record sizes (64 B per free vertex + 72 per triangle, 52 per hinge,
68 per stitch, 56 per contact incidence), divided by its mean launch time from
HIP events recorded around every launch on the solver's stream.

cpu_baseline: the f64 CPU restatement of VbdSolver::step (OpenMP within a
colour) on the host, on a bounded sample of the same drop: steps from the
start state (free fall) and from the end state (in contact), half the time
budget each, combined in the drop's proportions of the two phases.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def _traffic(kernel):
    """HBM bytes per sweep launch from the committed rocprofv3 counter runs
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_sheet_on_sphere_128_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_time_per_step(sc, state, settings, budget_s, threads):
    """Seconds per oracle f64 step on the host from `state` (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU restatement: baseline only, never the product
    st = state.copy()
    ref = O.Solver(sc.garment, sc.avatar, st, threads=threads)
    steps, t0 = 0, time.perf_counter()
    while True:
        ref.step(st, settings)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    return steps, el


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", type=int, default=64, choices=(32, 64),
                    help="sweep arithmetic (64: f64 default; 32: f32 opt-in)")
    ap.add_argument("--lanes", type=int, default=32, choices=(8, 16, 32, 64),
                    help="lanes per vertex of the sweep kernel")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, host CPUs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nx", type=int, default=128, help="quads per side (128 is the workload)")
    ap.add_argument("--workload", default="sheet_on_sphere",
                    choices=("sheet_on_sphere", "sewn_garment", "batch64", "warm_infer", "newton_sheet"),
                    help="sewn_garment: two ~50k-vertex Poisson-disk panels with side and shoulder "
                         "seams on a capsule body, 40 iterations; batch64: 64 independent sheets and "
                         "sewn tubes of 1k-50k vertices in batched steps; newton_sheet: the device "
                         "Newton energy reference of one headline step in contact, one solve per "
                         "--steps (not the headline line)")
    ap.add_argument("--fast-coloring", action="store_true",
                    help="fast colouring on the device (fewer colours; not the reference's colouring order)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch64: use --lanes instead of timing the lanes per vertex first")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
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

    from clothcraft_amd import GpuVbd, VbdSolver, scenes  # fails loudly if not built

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "warm_infer":
        return bench_warm_infer(args)
    if args.workload == "newton_sheet":
        return bench_newton(args)

    if args.workload == "sewn_garment":
        from clothcraft_amd import garments
        sc = garments.sewn_garment_on_body()
    else:
        sc = scenes.sheet_on_sphere(nx=args.nx)
    settings = sc.settings
    if args.fast_coloring:
        from clothcraft_amd.abi import CV_FLAG_FAST_COLORING
        settings = dict(settings, flags=settings.get("flags", 0) | CV_FLAG_FAST_COLORING)
    st = sc.state.copy()
    gpu = GpuVbd(0)
    gpu.set_precision(args.precision)
    gpu.set_lanes(args.lanes)
    sol = VbdSolver(sc.garment, sc.avatar, st)
    print(f"[bench] {sc.name}: {sc.garment.n} vertices, warmup {args.warmup} steps", file=sys.stderr,
          flush=True)
    if args.warmup:
        gpu.run(sol, st.copy(), settings, args.warmup)
    start_state = st.copy()

    print(f"[bench] timed: {args.steps} steps", file=sys.stderr, flush=True)
    # The timed steps run as device-resident chunks of 10 (run() keeps the
    # prediction history across calls, so results equal one long run); the
    # chunk ends show when the drop reaches the sphere.
    gpu.set_profiling(True)
    tot = dict(gpu_compute_ms=0.0, iterations_run=0, sweep_launches=0, sweep_ms=0.0, sweep_bytes=0.0,
               sweep_bytes_min=0.0, detection_updates=0, sweep_timed_launches=0, sweep_timed_bytes_min=0.0)
    first_contact, done, m = None, 0, None
    t0 = time.perf_counter()
    while done < args.steps:
        k = min(10, args.steps - done)
        m = gpu.run(sol, st, settings, k)
        for key in tot:
            tot[key] += m[key]
        done += k
        if first_contact is None and m["contact_terms"] > 0:
            first_contact = done - k
        if done % 50 == 0:
            print(f"[bench] {done}/{args.steps} steps", file=sys.stderr, flush=True)
    wall_s = time.perf_counter() - t0
    gpu.set_profiling(False)
    if not np.all(np.isfinite(st.x)):
        raise RuntimeError("non-finite state after the timed steps")
    # The reference's call pattern: one GpuVbd::step per frame, state in and
    # out through host memory and the metrics (diagnostics included) each
    # step; wall time per call, continuing from the final state on a copy.
    api_st = st.copy()
    api_sol = VbdSolver(sc.garment, sc.avatar, api_st)
    api_ms = []
    for _ in range(20):
        t1 = time.perf_counter()
        gpu.step(api_sol, api_st, settings)
        api_ms.append(1e3 * (time.perf_counter() - t1))
    api_ms = np.array(api_ms[2:])
    m = dict(m, **tot)

    K = args.steps
    dev_s = m["gpu_compute_ms"] / 1e3
    n_free = int(np.count_nonzero(sc.garment.pinned == 0))
    vi = n_free * m["iterations_run"]  # vertex-iterations in the timed window
    # The sweep launches of every 4th iteration carry begin/end events: their
    # mean duration and algorithmic bytes give the achieved bandwidth; the
    # sweep's share of the step extrapolates to all launches.
    launches = max(m["sweep_launches"], 1)
    timed = max(m["sweep_timed_launches"], 1)
    avg_launch_s = m["sweep_ms"] / 1e3 / timed
    bytes_per_launch = m["sweep_timed_bytes_min"] / timed
    achieved = bytes_per_launch / avg_launch_s / 1e9
    sweep_s = avg_launch_s * launches
    kernel = "k_sweep"

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
        "data": "synthetic (seeded sheet and icosphere geometry)",
        "config": {
            "workload": (f"sheet on sphere {args.nx}x{args.nx} quads: {sc.garment.n} vertices, "
                         f"{len(sc.garment.tri_vertices)} triangles, icosphere r 0.25 m "
                         f"({len(sc.avatar.triangles)} triangles), OGC r 2 mm k 1e5, "
                         f"{settings['iterations']} iterations, dt 1/60, self-contact on")
            if args.workload == "sheet_on_sphere" else
            (f"sewn garment: {sc.garment.n} vertices, {len(sc.garment.tri_vertices)} triangles, "
             f"{len(sc.garment.stitch_stiffness)} seam stitches, capsule body "
             f"({len(sc.avatar.triangles)} triangles), OGC r 1 mm k 1e5, "
             f"{settings['iterations']} iterations, dt 1/60, self-contact on"),
            "steps_timed": f"1..{K} of the drop (warmup on a discarded copy)",
            "iteration_mode": "fixed iterations (residual tolerance 0)",
            "coloring": ("fast: static smallest-last, device repair (CV_FLAG_FAST_COLORING)"
                         if args.fast_coloring else "reference vertex-order greedy"),
            "lanes_per_vertex": args.lanes,
        },
        "vertex_iters_per_s": vi / dev_s,
        "vertex_iters_per_s_sweep_only": vi / sweep_s if sweep_s > 0 else None,
        "steps_per_s_wall_incl_transfers": K / wall_s,
        "step_api_wall_ms": {"median": float(np.median(api_ms)), "p95": float(np.percentile(api_ms, 95)),
                             "calls": int(len(api_ms)), "warmup_calls": 2},
        "residual_rms_n_last_step": m["residual_rms_n"],
        "contacts_last_step": m["contact_terms"],
        "colors_last_step": m["colors"],
        "candidate_pairs_last_step": m["candidate_pairs"],
        "grid_pairs_examined_last_step": m["candidate_pairs"] + m["aabb_rejections"],
        "detections": m["detection_updates"],
        "first_contact_by_step": None if first_contact is None else first_contact + 10,
        "sweep_share_of_step": sweep_s / dev_s,
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": _traffic(kernel) if (args.workload == "sheet_on_sphere" and args.nx == 128
                                             and args.precision == 64
                                             and args.lanes == 32) else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_us": 1e6 * avg_launch_s,
            "launches": m["sweep_launches"],
            "launches_timed": m["sweep_timed_launches"],
            "bytes_per_launch_in_sweep_precision": m["sweep_bytes"] / launches,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline: oracle f64 step, {threads} threads, ~{args.cpu_seconds:.0f} s",
              file=sys.stderr, flush=True)
        half = args.cpu_seconds / 2
        fs, fe = cpu_time_per_step(sc, start_state, settings, half, threads)
        cs, ce = cpu_time_per_step(sc, st, settings, half, threads)
        # The drop's two phases: free fall until the first contact, then contact.
        k_fall = K if first_contact is None else first_contact
        per_step = (k_fall * fe / fs + (K - k_fall) * ce / cs) / K
        out["cpu_baseline"] = {
            "value": 1.0 / per_step,
            "unit": "steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"f64 C++ restatement of VbdSolver::step (OpenMP within a colour): {fs} steps "
                      f"from the start state ({fe:.1f} s) and {cs} steps from the state after step {K} "
                      f"({ce:.1f} s), weighted {k_fall}:{K - k_fall} as the drop's free-fall and "
                      f"contact steps",
            "vertex_iters_per_s": n_free * settings["iterations"] / per_step,
        }
    print(json.dumps(out), flush=True)


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
