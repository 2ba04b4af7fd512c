"""Replays the device step on the CPU: the same prologue, ball rule,
truncation, refresh and colour order as vbd_device.hip, with each sweep
launch executed by the host build of the kernel math (host_mirror.cpp) and
the contact-free detection radii taken from the oracle.  Test helper only."""
import numpy as np

import oracle as O
from clothcraft_amd import host


def emulate_step(sc, state, f32=True):
    """f32: the device's f32 sweep (f64 positions, each vertex solved in f32
    in its own frame); otherwise the f64 sweep."""
    g, s = sc.garment, sc.settings
    n = g.n
    dt, c = s["dt"], s["damping"]
    damp = 1 + c * dt
    grav = np.array(s["gravity"])
    dtype = np.float64
    origin = state.x.mean(axis=0)
    target = state.x + (state.v * dt + grav * dt * dt) / damp
    inertia = g.masses * damp / (dt * dt)
    guess = state.x + (state.v * dt + 1.0 * grav * (dt * dt)) / (1 + c * dt)
    motion = np.linalg.norm(target - state.x, axis=1).max()
    r = s["radius"]
    rq = min(max(r + motion, 2 * r), 4 * r)
    hs = host.HostSolver(g, sc.avatar, state)
    off, recs = hs.terms()
    ref = O.Solver(g, sc.avatar, state)
    pos = np.zeros((n, 4), dtype)
    tgt = np.zeros((n, 4), dtype)
    ball = np.zeros((n, 4), dtype)
    tgt[:, :3] = target - origin
    tgt[:, 3] = inertia
    pinned = g.pinned.astype(bool)

    def detect(x64):
        cols, k, cs = ref.detect_and_color(x64, rq, O.ogc_c(radius=r, stiffness=s["stiffness"]))
        assert len(cs["kind_ids"]) == 0, "contact-free scenes only"
        radii = cs["radii"].copy()
        ball[:, :3] = x64 - origin
        if f32:  # the f32 sweep's radius shrink by four f32 ulps of itself (k_balls)
            radii = radii - 4 * 1.1920928955078125e-07 * radii
        ball[:, 3] = radii
        return [np.nonzero(cols == cc)[0] for cc in range(k)], radii, x64.copy()

    groups, rr, centers = detect(state.x.copy())
    clipped = 0
    for i in range(n):
        if pinned[i]:
            pos[i, :3] = state.x[i] - origin
            pos[i, 3] = 1
            continue
        d = guess[i] - centers[i]
        L = np.linalg.norm(d)
        p = centers[i] + d * (rr[i] / L) if L > rr[i] else guess[i]
        clipped += L > rr[i]
        pos[i, :3] = p - origin
    gsq = np.zeros(n, dtype)
    coff = np.zeros(n + 1, np.uint32)
    crec = np.zeros(0, host.TERM_DTYPE if f32 else host.TERM64_DTYPE)
    terms = recs if f32 else hs.terms64()
    for _ in range(s["iterations"]):
        if clipped > 0 and clipped >= s["refresh_fraction"] * n:
            x64 = origin + pos[:, :3].astype(np.float64)
            x64[pinned] = state.x[pinned]
            groups, rr, centers = detect(x64)
        clipped = 0
        for grp in groups:
            cl, _ = host.sweep_color(grp, pos, tgt, ball, off, terms, coff, crec, r, s["stiffness"], gsq,
                                     f32=f32)
            clipped += cl
    x = origin + pos[:, :3].astype(np.float64)
    x[pinned] = state.x[pinned]
    return x
