"""Training data, training and evaluation of the learned warm start.

Samples are (features, displacement / L) pairs of moving vertices, taken from
trajectories the plain solver itself generates: at each state the device
computes the features the step would use (cv_warm_features), the plain step
runs in the residual-tolerance mode to a tight tolerance (a converged step),
and the label is the step's displacement in units of
L = |g| dt^2 / (1 + c dt).  Seeded velocity perturbations diversify the
states.  Evaluation compares iterations-to-tolerance of the adaptive start and
the learned start from the same states, with the true residual and the step's
objective value reported for both.
"""
from __future__ import annotations

import numpy as np

from .vbd import GpuVbd, VbdSolver


def step_length(settings) -> float:
    g = np.linalg.norm(settings["gravity"])
    return max(g * settings["dt"] ** 2 / (1.0 + settings["damping"] * settings["dt"]), 1e-9)


def generate(gpu: GpuVbd, scene, steps: int, tol: float, max_iterations: int = 200, seed: int = 0,
             perturb: float = 0.0, keep_every: int = 1):
    """Roll the plain solver `steps` steps; returns (X [m,12], Y [m,3], states)."""
    rng = np.random.default_rng(seed)
    st = scene.state.copy()
    sol = VbdSolver(scene.garment, scene.avatar, st)
    s = dict(scene.settings, residual_tolerance=tol, iterations=max_iterations)
    L = step_length(s)
    free = scene.garment.pinned == 0
    Xs, Ys, states = [], [], []
    for k in range(steps):
        if perturb > 0.0:
            st.v[free] += rng.normal(scale=perturb, size=(int(free.sum()), 3))
        X = gpu.warm.features(sol, st, s)
        x0 = st.x.copy()
        states.append(st.copy())
        prev_v = st.v.copy()
        gpu.step(sol, st, s)
        if perturb > 0.0:
            sol.record_prediction(prev_v, st, s)
        if k % keep_every == 0:
            Xs.append(X[free])
            Ys.append(((st.x - x0) / L)[free].astype(np.float32))
    return np.vstack(Xs), np.vstack(Ys), states


def evaluate(gpu: GpuVbd, scene, states, tol: float, max_iterations: int = 200):
    """Iterations-to-tolerance of the adaptive and the learned start from each
    state (each from a fresh solver with the same prediction history: none)."""
    out = []
    for st0 in states:
        row = {}
        for name, flags in (("plain", 0), ("warm", 8)):
            st = st0.copy()
            sol = VbdSolver(scene.garment, scene.avatar, st)
            s = dict(scene.settings, residual_tolerance=tol, iterations=max_iterations, flags=flags)
            m = gpu.step(sol, st, s)
            row[name] = {"iterations": m["iterations_run"], "true_residual": m["true_residual_rms_n"],
                         "x": st.x.copy()}
        out.append(row)
    return out
