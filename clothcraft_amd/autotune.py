"""Launch-shape autotuner: lanes per vertex of the sweep kernels.

The sweep gives each vertex a group of L lanes (8, 16, 32 or 64) that split
its term records and reduce with cross-lane shuffles; the best L depends on
the terms per vertex (triangles, hinges, stitches, contacts) and on how many
vertices a colour holds.  The tuner times one step per candidate from scratch
copies of the given states (the caller's states are not touched), keeps the
fastest and leaves it set on the GpuVbd.  Timing is device time: the sweep
launches' mean event-stamped duration times their count for one solver, the
batch's device span for a batch.
"""
from __future__ import annotations

from .vbd import GpuVbd, VbdSolver

CANDIDATES = (8, 16, 32, 64)


def tune_lanes(gpu: GpuVbd, scenes, settings, candidates=CANDIDATES, repeats: int = 2) -> dict:
    """scenes: one Scene or a list (then timed through the batched step).
    Returns {"lanes": best, "ms": {L: ms}}."""
    batch = isinstance(scenes, (list, tuple))
    items = list(scenes) if batch else [scenes]
    times = {}
    for L in candidates:
        gpu.set_lanes(L)
        best = float("inf")
        for _ in range(repeats):
            states = [sc.state.copy() for sc in items]
            sols = [VbdSolver(sc.garment, sc.avatar, st) for sc, st in zip(items, states)]
            if batch:
                m = gpu.step_batch(sols, states, settings)
                t = m[0]["gpu_compute_ms"]
            else:
                gpu.set_profiling(True)
                try:
                    m = gpu.step(sols[0], states[0], settings)
                finally:
                    gpu.set_profiling(False)
                # Sampled launches' mean duration times all launches.
                t = m["sweep_ms"] / max(m["sweep_timed_launches"], 1) * m["sweep_launches"]
            best = min(best, t)
        times[L] = best
    lanes = min(times, key=times.get)
    gpu.set_lanes(lanes)
    return {"lanes": lanes, "ms": times}
