"""Launch-shape autotuner: lanes per vertex of the sweep kernels and, for the
batched step, threads per block of its shared sweep launches.

The sweep gives each vertex a group of L lanes (8, 16, 32 or 64) that split
its term records and reduce with cross-lane shuffles; the best L depends on
the terms per vertex (triangles, hinges, stitches, contacts) and on how many
vertices a colour holds.  The batched step's shared launches also take a
block size (64, 128 or 256 threads: 64/L to 256/L vertices per block), which
sets how finely the launch's many instances are spread over the compute
units.  The tuner times one step per candidate from scratch copies of the
given states (the caller's states are not touched), keeps the fastest and
leaves it set on the GpuVbd.  Timing is device time: the sweep launches'
mean event-stamped duration times their count for one solver, the batch's
device span for a batch.  Neither choice changes results.
"""
from __future__ import annotations

from .vbd import GpuVbd, VbdSolver

CANDIDATES = (8, 16, 32, 64)
BLOCKS = (64, 128, 256)


def tune_lanes(gpu: GpuVbd, scenes, settings, candidates=CANDIDATES, repeats: int = 2,
               blocks=BLOCKS) -> dict:
    """scenes: one Scene or a list (then timed through the batched step, over
    lanes x block sizes).  Returns {"lanes": best, "ms": {L: ms}} and, for a
    batch, "block" and "ms_by_block": {B: {L: ms}}."""
    batch = isinstance(scenes, (list, tuple))
    _warm_up(gpu, scenes, settings, batch)
    if batch:
        by_block = {}
        for B in blocks:
            gpu.set_batch_block(B)
            by_block[B] = _tune(gpu, scenes, settings, candidates, repeats, True)
        B, L = min(((B, L) for B in by_block for L in by_block[B]), key=lambda k: by_block[k[0]][k[1]])
        gpu.set_batch_block(B)
        gpu.set_lanes(L)
        return {"lanes": L, "block": B, "ms": by_block[B], "ms_by_block": by_block}
    times = _tune(gpu, scenes, settings, candidates, repeats, False)
    lanes = min(times, key=times.get)
    gpu.set_lanes(lanes)
    return {"lanes": lanes, "ms": times}


def _warm_up(gpu: GpuVbd, scenes, settings, batch) -> None:
    """One untimed step, so the first timed candidate does not pay the
    context's first allocations and code loads."""
    items = list(scenes) if batch else [scenes]
    states = [sc.state.copy() for sc in items]
    sols = [VbdSolver(sc.garment, sc.avatar, st) for sc, st in zip(items, states)]
    if batch:
        gpu.step_batch(sols, states, settings)
    else:
        gpu.step(sols[0], states[0], settings)


def _tune(gpu: GpuVbd, scenes, settings, candidates, repeats, batch) -> dict:
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
    return times
