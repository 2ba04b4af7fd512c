"""The learned warm start (no reference counterpart): the numpy checker is
pinned by finite differences; the device forward pass, backward pass,
training, features and the warm-started step are checked against it and
against the oracle step started from the same guess."""
import numpy as np
import pytest

import mlp_ref as M
import oracle as O
from clothcraft_amd import scenes


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 3], X[:, 1] * X[:, 2] * 0.3, np.tanh(X[:, 11] - X[:, 6])], 1)
    return X, T.astype(np.float32)


def test_checker_backward_matches_finite_differences():
    p = M.init_params(3).astype(np.float64)
    X, T = _data(9)
    loss, g = M.loss_and_grad(p, X, T, rounding=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.NPARAM, 60, replace=False), np.arange(M.NPARAM - 3, M.NPARAM)])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (M.loss_and_grad(p + e, X, T, False)[0] - M.loss_and_grad(p - e, X, T, False)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e-3], np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 + 2 ** -7 and r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


@pytest.mark.gpu
def test_device_inference_matches_checker(gpu):
    p = M.init_params(5, out_scale=1.0)
    gpu.warm.set_params(p)
    assert np.array_equal(gpu.warm.params(), p)
    X, _ = _data(1000, 2)
    y = gpu.warm.infer(X)
    ref, _, _ = M.forward(p, X, rounding=True)
    exact, _, _ = M.forward(p, X, rounding=False)
    scale = np.abs(ref).max()
    print(f"inference: max |gpu - checker(bf16)| {np.abs(y - ref).max():.3e}, "
          f"vs exact f64 {np.abs(y - exact).max():.3e}, scale {scale:.3e}")
    assert np.abs(y - ref).max() <= 1e-2 * scale
    assert np.abs(y - exact).max() <= 5e-2 * scale


@pytest.mark.gpu
def test_device_gradient_matches_checker(gpu):
    """The backward kernels (matrix-core products for dA and dW, column sums
    for the biases) against the checker with the same BF16 rounding points."""
    p = M.init_params(7, out_scale=1.0)
    gpu.warm.set_params(p)
    for n in (32, 1000):
        X, T = _data(n, 4)
        g, loss = gpu.warm.gradient(X, T)
        lr, gr = M.loss_and_grad(p, X, T, rounding=True)
        le, ge = M.loss_and_grad(p, X, T, rounding=False)
        rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        rel_exact = np.linalg.norm(g - ge) / np.linalg.norm(ge)
        print(f"n={n}: loss {loss:.6f} / {lr:.6f}, gradient rel err vs checker {rel:.2e}, vs exact {rel_exact:.2e}")
        assert abs(loss - lr) <= 1e-3 * lr
        assert rel <= 1e-2
        assert rel_exact <= 3e-2
        # every block of parameters is exercised
        o = 0
        for name, shape in M.SHAPES:
            k = int(np.prod(shape))
            blk = slice(o, o + k)
            assert np.linalg.norm(g[blk] - gr[blk]) <= 2e-2 * np.linalg.norm(gr[blk]) + 1e-9, name
            o += k


@pytest.mark.gpu
def test_device_training_fits_a_smooth_map(gpu):
    p = M.init_params(11)
    gpu.warm.set_params(p)
    X, T = _data(20000, 6)
    losses = gpu.warm.train(X, T, epochs=12, batch=1024, lr=2e-3, seed=3)
    print("epoch losses", np.round(losses, 5))
    assert losses[-1] < 0.1 * losses[0]
    y = gpu.warm.infer(X[:2000])
    assert np.mean((y - T[:2000]) ** 2) <= 1.5 * losses[-1] + 1e-3


def _numpy_features(sc, state, settings, radii, relaxation, rq):
    g = sc.garment
    x, v = state.x, state.v
    dt, c = settings["dt"], settings["damping"]
    grav = np.array(settings["gravity"])
    L = np.linalg.norm(grav) * dt * dt / (1 + c * dt)
    F = np.zeros((g.n, 12))
    F[:, 0:3] = (v * dt + grav * dt * dt) / (1 + c * dt) / L
    F[:, 3:6] = v * dt / L
    tv = g.tri_vertices.astype(np.int64)
    inv = g.tri_inv_dm.reshape(-1, 2, 2).transpose(0, 2, 1)  # column-major -> [row][col]
    e1, e2 = x[tv[:, 1]] - x[tv[:, 0]], x[tv[:, 2]] - x[tv[:, 0]]
    D = np.stack([e1, e2], 2)  # [T,3,2]
    Fd = D @ inv
    ew = 0.5 * ((Fd[:, :, 0] ** 2).sum(1) - 1)
    ev = 0.5 * ((Fd[:, :, 1] ** 2).sum(1) - 1)
    cnt = np.zeros(g.n)
    sw, sv = np.zeros(g.n), np.zeros(g.n)
    for k in range(3):
        np.add.at(cnt, tv[:, k], 1)
        np.add.at(sw, tv[:, k], ew)
        np.add.at(sv, tv[:, k], ev)
    F[:, 6] = 100 * sw / np.maximum(cnt, 1)
    F[:, 7] = 100 * sv / np.maximum(cnt, 1)
    nbrs = [set() for _ in range(g.n)]
    for a, b, c_ in tv:
        for p, q in ((a, b), (b, c_), (c_, a)):
            nbrs[p].add(q)
            nbrs[q].add(p)
    for i in range(g.n):
        nb = sorted(nbrs[i])
        F[i, 8:11] = v[nb].mean(0) * dt / L
    F[:, 11] = radii / (relaxation * rq)
    return F


@pytest.mark.gpu
def test_device_features_match_numpy(gpu):
    from clothcraft_amd import VbdSolver
    sc = scenes.small_sheet_on_sphere(16, 0.2535)
    st = sc.state.copy()
    rng = np.random.default_rng(0)
    st.v[:] = rng.normal(scale=0.05, size=st.v.shape)
    st.x[:] += rng.normal(scale=2e-4, size=st.x.shape)
    sol = VbdSolver(sc.garment, sc.avatar, st)
    s = sc.settings
    f = gpu.warm.features(sol, st, s)
    dt, c = s["dt"], s["damping"]
    motion = np.linalg.norm((st.v * dt + np.array(s["gravity"]) * dt * dt) / (1 + c * dt), axis=1).max()
    rq = min(max(s["radius"] + motion, 2 * s["radius"]), 4 * s["radius"])
    det = gpu.detect(sol, st.x, rq, candidates=False)
    ref = _numpy_features(sc, st, s, det["radii"], s["relaxation"], rq)
    err = np.abs(f - ref).max(0)
    print("feature max abs err per column", np.array2string(err, precision=2))
    assert np.all(err <= 1e-5 * (1 + np.abs(ref).max(0)))


@pytest.mark.gpu
def test_warm_started_step_matches_oracle_from_the_same_guess(gpu):
    """With CV_FLAG_WARM_START the device step starts from the inertial
    prediction plus L * network(features); the oracle step started from that
    same guess lands at the same place."""
    from clothcraft_amd import VbdSolver
    sc = scenes.small_sheet_on_sphere(16, 0.2535)
    gpu.warm.set_params(M.init_params(13, out_scale=0.3))
    st = sc.state.copy()
    st.v[:] = np.array([0.0, -0.1, 0.0])
    s = dict(sc.settings, flags=8)
    sol = VbdSolver(sc.garment, sc.avatar, st)
    feats = gpu.warm.features(sol, st, s)
    pred = gpu.warm.infer(feats).astype(np.float64)
    dt, c = s["dt"], s["damping"]
    L = np.linalg.norm(s["gravity"]) * dt * dt / (1 + c * dt)
    guess = st.x + (st.v * dt + np.array(s["gravity"]) * dt * dt) / (1 + c * dt) + L * pred
    a, b = st.copy(), st.copy()
    ref = O.Solver(sc.garment, sc.avatar, st)
    ref.set_next_guess(guess)
    ref.step(a, dict(s, flags=0))
    gpu.step(VbdSolver(sc.garment, sc.avatar, st), b, s)
    err = np.abs(a.x - b.x).max() / np.linalg.norm(a.x.max(0) - a.x.min(0))
    plain = st.copy()
    gpu.step(VbdSolver(sc.garment, sc.avatar, st), plain, dict(s, flags=0))
    print(f"warm step vs oracle rel err {err:.3e}; differs from the plain step by "
          f"{np.abs(plain.x - b.x).max():.3e} m")
    assert err <= 1e-9
    assert np.abs(plain.x - b.x).max() > 1e-7  # the guess did change the step


def _objective(sc, st0, x, settings):
    """The step objective (Newton form, equal to the VBD objective up to a
    constant) with the step's OGC contacts frozen, from the oracle."""
    ref = O.Solver(sc.garment, sc.avatar, st0.copy())
    ref.step(st0.copy(), dict(settings, iterations=1))
    fz = O.frozen_c(ref.last_contacts(), settings["radius"], settings["stiffness"], sc.avatar.vertices)
    ns = O.newton_settings(dt=settings["dt"], gravity=settings["gravity"], damping=settings["damping"])
    return lambda xx: O.newton_objective(sc.garment, st0, xx, ns, None, fz)


@pytest.mark.gpu
def test_trained_warm_start_reaches_the_plain_energy_in_fewer_sweeps(gpu):
    """Trained on converged steps of two sewn garments, evaluated on a third
    (held-out seed), both starts without a prediction history (so the plain
    start is the inertial prediction): after K sweeps the learned start's
    objective gap to a 200-sweep solve is smaller than the plain start's, and
    the learned start reaches the plain start's K-sweep objective in fewer
    sweeps."""
    from clothcraft_amd import VbdSolver, garments, warmstart
    train = [garments.sewn_garment_on_body(spacing=0.012, seed=s) for s in (2, 5)]
    Xs, Ys = [], []
    for i, sc in enumerate(train):
        for pert in (0.0, 0.05):
            X, Y, _ = warmstart.generate(gpu, sc, 30, seed=10 + i, perturb=pert, start=5)
            Xs.append(X)
            Ys.append(Y)
    gpu.warm.set_params(M.init_params(0))
    losses = gpu.warm.train(np.vstack(Xs), np.vstack(Ys), epochs=40, batch=4096, lr=1e-3, seed=1)
    assert losses[-1] < 0.5 * losses[0]
    sc = garments.sewn_garment_on_body(spacing=0.012, seed=9)
    s = sc.settings
    _, _, states = warmstart.generate(gpu, sc, 3, converge_iterations=1, seed=3, start=20)
    K = 20
    gains, gaps = [], []
    for st0, _ in states:
        E = _objective(sc, st0, st0.x, s)
        ref = st0.copy()
        gpu.step(VbdSolver(sc.garment, sc.avatar, ref), ref, dict(s, iterations=200))
        e_ref = E(ref.x)

        def gap(k, flags):
            a = st0.copy()
            gpu.step(VbdSolver(sc.garment, sc.avatar, a), a, dict(s, iterations=k, flags=flags))
            return E(a.x) - e_ref

        plain, warm = gap(K, 0), gap(K, 8)
        k_eq = next((k for k in (1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 18, 20, 25, 30, 40) if gap(k, 8) <= plain), 60)
        gains.append(K / k_eq)
        gaps.append((plain, warm))
        print(f"K={K}: objective gap plain {plain:.4e} learned {warm:.4e}; learned start reaches the plain "
              f"{K}-sweep objective in {k_eq} sweeps")
    print(f"epoch losses {np.round(losses[::10], 5)}; sweeps-to-equal-objective gain per state "
          f"{np.round(gains, 2)} (mean {np.mean(gains):.2f})")
    assert np.mean([w for _, w in gaps]) < np.mean([p for p, _ in gaps])
    assert np.mean(gains) > 1.0


# Measured with the committed network on MI355X (deterministic device
# computation): iterations to the plain start's 40-iteration true residual.
CONFIG2_GAIN = 0.682


@pytest.mark.gpu
def test_committed_warm_start_on_the_config2_garment(gpu):
    """The committed network (clothcraft_amd/warm_params.npy, trained by
    clothcraft_amd/train_warm.py on config-2 and config-3 trajectories) on six
    held-out states of the full config-2 garment (seed 9, steps 20-60): the
    iterations each start needs to reach the true residual the plain start
    (the reference's adaptive guess with its prediction history) reaches
    after 40 iterations.  The learned start needs more: on this garment the
    residual plateaus after a few iterations for either start, and the
    correction left to learn (0.015 L, about 40 micrometres) is below the
    network's accuracy in the membrane's stiff directions.  The test records
    the gain and checks the value stated for it."""
    import os
    from clothcraft_amd import garments, warmstart
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "clothcraft_amd",
                        "warm_params.npy")
    gpu.warm.set_params(np.load(path))
    gpu.warm.set_smoothing(1)
    try:
        ev = garments.sewn_garment_on_body(seed=9)
        _, _, states = warmstart.generate(gpu, ev, 41, converge_iterations=1, seed=3, start=20, every=8)
        r = warmstart.evaluate(gpu, ev, states, K=40)
    finally:
        gpu.warm.set_smoothing(0)
    rows = [(x["plain_iterations"], x["warm_iterations"]) for x in r["rows"]]
    print(f"config-2 garment, {r['states']} states: iterations to the plain 40-iteration residual, plain "
          f"{r['plain_iterations_mean']:.1f} learned {r['warm_iterations_mean']:.1f} (gain {r['gain']:.3f}); "
          f"residual after 40 iterations plain {r['plain_residual_after_K_mean']:.4e} learned "
          f"{r['warm_residual_after_K_mean']:.4e}; per state {rows}")
    assert r["states"] == 6
    if CONFIG2_GAIN is not None:
        assert r["gain"] == pytest.approx(CONFIG2_GAIN, abs=0.02)
    assert r["gain"] < 1.0
