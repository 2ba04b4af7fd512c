"""Trains the learned warm start on the GPU and writes warm_params.npy.

Data: (features, correction) pairs of moving vertices from trajectories the
plain solver generates (warmstart.generate): the config-2 sewn garment
(99,287 vertices, seed 2) every 6th of 66 steps after 4, once as rolled and
once with seeded N(0, 0.02 m/s) velocity perturbations, plus six config-3
instances (5k-40k vertices, seeded sheets and sewn tubes) every 4th of 60
steps: about 3.3 million pairs.  Labels are 200-sweep solves from the
adaptive guess.  Training: the device's BF16-MFMA forward / exact backward /
Adam kernels, 30 epochs, mini-batches of 4,096, learning rate 1e-3, seed 1.

Run on a machine with an MI355X from the repository root:  python -m clothcraft_amd.train_warm
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import garments, warmstart
from .vbd import GpuVbd

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warm_params.npy")


def training_set(gpu: GpuVbd, log=print):
    Xs, Ys = [], []
    sc = garments.sewn_garment_on_body(seed=2)
    for pert, seed in ((0.0, 1), (0.02, 2)):
        X, Y, _ = warmstart.generate(gpu, sc, 66, seed=seed, perturb=pert, start=4, every=6)
        Xs.append(X)
        Ys.append(Y)
        log(f"garment (perturbation {pert}): {len(X)} pairs")
    batch = garments.batched_scenes(64, seed=3)
    for i, b in enumerate([b for b in batch if 5000 < b.garment.n < 40000][:6]):
        X, Y, _ = warmstart.generate(gpu, b, 60, seed=20 + i, start=0, every=4)
        Xs.append(X)
        Ys.append(Y)
        log(f"{b.name}: {len(X)} pairs")
    return np.vstack(Xs), np.vstack(Ys)


def main(out: str = OUT) -> np.ndarray:
    t0 = time.time()

    def log(msg):
        print(f"[train_warm {time.time() - t0:6.1f}s] {msg}", file=sys.stderr, flush=True)

    gpu = GpuVbd(0)
    X, Y = training_set(gpu, log)
    gpu.warm.set_params(warmstart.init_params(0))
    losses = gpu.warm.train(X, Y, epochs=30, batch=4096, lr=1e-3, seed=1)
    log(f"{len(X)} pairs; loss {losses[0]:.6f} -> {losses[-1]:.6f} (zero network {float((Y ** 2).mean()):.6f})")
    p = gpu.warm.params()
    np.save(out, p)
    log(f"wrote {out}")
    return p


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else OUT)
