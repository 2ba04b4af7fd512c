"""Training data, training and evaluation of the learned warm start.

Samples are (features, correction / L) pairs of moving vertices, taken from
trajectories the plain solver itself generates: at each state, with its
prediction history, the device computes the features the step would use
(cv_warm_features), a converged step (many sweeps) is solved from it, and
the label is that step's displacement minus the offset of the step's
adaptive inertial prediction (Prediction::guess), in units of
L = |g| dt^2 / (1 + c dt): the network learns a correction, so a network that
outputs zero is the plain start.  Seeded velocity perturbations of the
labelled states diversify them.  Evaluation measures, from held-out states,
how many sweeps each start needs to reach the true residual the plain start
reaches after K sweeps (equal final accuracy).
"""
from __future__ import annotations

import numpy as np

from .vbd import GpuVbd, VbdSolver


# Layer shapes in the order of the FP32 parameter vector (mlp.hpp):
# W0 [12 x 128] (input-major), b0, W1 [128 x 128], b1, W2, b2, W3 [128 x 3], b3.
SHAPES = [("W0", (12, 128)), ("b0", (128,)), ("W1", (128, 128)), ("b1", (128,)), ("W2", (128, 128)),
          ("b2", (128,)), ("W3", (128, 3)), ("b3", (3,))]


def init_params(seed: int = 0, out_scale: float = 0.1) -> np.ndarray:
    """Fresh network parameters: He-normal weights, zero biases, the output
    layer scaled by out_scale so an untrained network starts close to the
    plain inertial guess."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            w = rng.normal(0.0, np.sqrt(2.0 / shape[0]), size=shape)
            parts.append(w * out_scale if name == "W3" else w)
        else:
            parts.append(np.zeros(shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def step_length(settings) -> float:
    g = np.linalg.norm(settings["gravity"])
    return max(g * settings["dt"] ** 2 / (1.0 + settings["damping"] * settings["dt"]), 1e-9)


def generate(gpu: GpuVbd, scene, steps: int, converge_iterations: int = 200, seed: int = 0,
             perturb: float = 0.0, start: int = 0, every: int = 1):
    """Roll the plain solver (the workload's own settings) `start + steps`
    steps; from every `every`-th of the last `steps` states also solve one
    converged step (`converge_iterations` sweeps) whose displacement / L,
    minus the offset of the step's adaptive inertial prediction, is the label.
    Each labelled state carries its prediction history (the velocity before
    its step), so features and label refer to the guess the step would use.
    `perturb` adds seeded N(0, perturb) velocity noise to each labelled state
    (not carried into the trajectory).  Returns (X [m,12], Y [m,3], states),
    states as (state, previous velocity) pairs."""
    rng = np.random.default_rng(seed)
    st = scene.state.copy()
    sol = VbdSolver(scene.garment, scene.avatar, st)
    s = scene.settings
    conv = dict(s, iterations=converge_iterations)
    L = step_length(s)
    free = scene.garment.pinned == 0
    Xs, Ys, states = [], [], []
    prev_v = st.v.copy()
    for k in range(start + steps):
        if k >= start and (k - start) % every == 0:
            probe = st.copy()
            if perturb > 0.0:
                probe.v[free] += rng.normal(scale=perturb, size=(int(free.sum()), 3))
            psol = with_history(scene, probe, prev_v, s)
            X = gpu.warm.features(psol, probe, conv)
            x0 = probe.x.copy()
            states.append((probe.copy(), prev_v.copy()))
            gpu.step(psol, probe, conv)
            Xs.append(X[free])
            # features 0..2 are the offset of the adaptive prediction / L
            Ys.append(((probe.x - x0) / L)[free].astype(np.float32) - X[free, 0:3])
        prev_v = st.v.copy()
        gpu.step(sol, st, s)
    return np.vstack(Xs), np.vstack(Ys), states


def with_history(scene, state, prev_v, settings) -> VbdSolver:
    """A solver for `state` whose prediction history is the step that ended
    at `state` from velocity prev_v (Prediction::record, src/vbd/prediction.rs:43-54)."""
    sol = VbdSolver(scene.garment, scene.avatar, state)
    sol.record_prediction(prev_v, state, settings)
    return sol


def evaluate(gpu: GpuVbd, scene, states, K: int, max_iterations: int = 200, tol=None):
    """From each (state, previous velocity): the plain (adaptive) start's true
    residual after K sweeps, and the sweeps the plain and the learned start
    need to reach it (or the absolute `tol`), at most max_iterations; also the
    learned start's residual after K sweeps."""
    s = scene.settings
    rows = []
    for st0, pv in states:
        a = st0.copy()
        mp = gpu.step(with_history(scene, a, pv, s), a, dict(s, iterations=K, flags=2))
        t = tol if tol is not None else mp["true_residual_rms_n"]
        ts = dict(s, iterations=max_iterations, residual_tolerance=t, flags=2)
        b = st0.copy()
        mpt = gpu.step(with_history(scene, b, pv, s), b, ts)
        c = st0.copy()
        mw = gpu.step(with_history(scene, c, pv, s), c, dict(ts, flags=2 | 8))
        d = st0.copy()
        mk = gpu.step(with_history(scene, d, pv, s), d, dict(s, iterations=K, flags=2 | 8))
        rows.append({"tolerance": t, "plain_residual_after_K": mp["true_residual_rms_n"],
                     "plain_iterations": mpt["iterations_run"], "warm_iterations": mw["iterations_run"],
                     "warm_residual_after_K": mk["true_residual_rms_n"]})
    pi = np.array([r["plain_iterations"] for r in rows], np.float64)
    wi = np.array([r["warm_iterations"] for r in rows], np.float64)
    return {"K": K, "states": len(rows), "plain_iterations_mean": float(pi.mean()),
            "warm_iterations_mean": float(wi.mean()), "gain": float(pi.mean() / max(wi.mean(), 1e-9)),
            "reached": int((wi < max_iterations).sum()),
            "plain_residual_after_K_mean": float(np.mean([r["plain_residual_after_K"] for r in rows])),
            "warm_residual_after_K_mean": float(np.mean([r["warm_residual_after_K"] for r in rows])),
            "rows": rows}
