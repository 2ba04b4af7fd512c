"""Training data, training and evaluation of the learned warm start.

Samples are (features, displacement / L) pairs of moving vertices, taken from
trajectories the plain solver itself generates: at each state the device
computes the features the step would use (cv_warm_features), a converged step
(many sweeps) is solved from it, and the label is that step's displacement
minus the inertial prediction (v dt + g dt^2) / (1 + c dt), in units of
L = |g| dt^2 / (1 + c dt) (the network learns the correction, so a network
that outputs zero is the plain inertial start).  Seeded velocity perturbations of the
labelled states diversify them.  Evaluation measures, from held-out states,
how many sweeps the learned start needs to reach the true residual the plain
start reaches after K sweeps (equal final accuracy).
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
             perturb: float = 0.0, start: int = 0):
    """Roll the plain solver (the workload's own settings) `start + steps`
    steps; from each of the last `steps` states also solve one converged step
    (`converge_iterations` sweeps) whose displacement / L is the label.
    `perturb` adds seeded N(0, perturb) velocity noise to each labelled state
    (not carried into the trajectory).  Returns (X [m,12], Y [m,3], states)."""
    rng = np.random.default_rng(seed)
    st = scene.state.copy()
    sol = VbdSolver(scene.garment, scene.avatar, st)
    s = scene.settings
    conv = dict(s, iterations=converge_iterations)
    L = step_length(s)
    free = scene.garment.pinned == 0
    Xs, Ys, states = [], [], []
    for k in range(start + steps):
        if k >= start:
            probe = st.copy()
            if perturb > 0.0:
                probe.v[free] += rng.normal(scale=perturb, size=(int(free.sum()), 3))
            psol = VbdSolver(scene.garment, scene.avatar, probe)
            X = gpu.warm.features(psol, probe, conv)
            x0 = probe.x.copy()
            states.append(probe.copy())
            gpu.step(psol, probe, conv)
            Xs.append(X[free])
            # features 0..2 are the inertial prediction offset / L
            Ys.append(((probe.x - x0) / L)[free].astype(np.float32) - X[free, 0:3])
        gpu.step(sol, st, s)
    return np.vstack(Xs), np.vstack(Ys), states


def evaluate(gpu: GpuVbd, scene, states, K: int, max_iterations: int = 200):
    """From each state: the plain start's true residual after K sweeps, and the
    sweeps the learned start needs to reach that residual (at most
    max_iterations); also the learned start's residual after K sweeps.  Each
    run uses a fresh solver (no prediction history on either side)."""
    rows = []
    for st0 in states:
        a = st0.copy()
        mp = gpu.step(VbdSolver(scene.garment, scene.avatar, a), a, dict(scene.settings, iterations=K, flags=2))
        tol = mp["true_residual_rms_n"]
        b = st0.copy()
        mw = gpu.step(VbdSolver(scene.garment, scene.avatar, b), b,
                      dict(scene.settings, iterations=max_iterations, residual_tolerance=tol, flags=8))
        c = st0.copy()
        mk = gpu.step(VbdSolver(scene.garment, scene.avatar, c), c, dict(scene.settings, iterations=K, flags=8 | 2))
        rows.append({"plain_residual": tol, "warm_iterations": mw["iterations_run"],
                     "warm_residual_at_tol": mw["true_residual_rms_n"],
                     "warm_residual_after_K": mk["true_residual_rms_n"]})
    its = np.array([r["warm_iterations"] for r in rows])
    return {"K": K, "states": len(rows), "median_warm_iterations": float(np.median(its)),
            "mean_warm_iterations": float(its.mean()), "reached": int((its < max_iterations).sum()),
            "gain_median": K / float(np.median(its)),
            "plain_residual_mean": float(np.mean([r["plain_residual"] for r in rows])),
            "warm_residual_after_K_mean": float(np.mean([r["warm_residual_after_K"] for r in rows])),
            "rows": rows}
