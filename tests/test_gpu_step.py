"""Device VBD step against the f64 oracle and the reference's closed forms."""
import numpy as np
import pytest

from clothcraft_amd import scenes

pytestmark = pytest.mark.gpu


def _bbox_diag(x):
    return float(np.linalg.norm(x.max(axis=0) - x.min(axis=0)))


def _one_step(gpu, sc, settings=None):
    import oracle as O
    from clothcraft_amd import VbdSolver
    settings = settings or sc.settings
    s_gpu, s_ref = sc.state.copy(), sc.state.copy()
    ref = O.Solver(sc.garment, sc.avatar, s_ref, threads=8)
    sol = VbdSolver(sc.garment, sc.avatar, s_gpu)
    m_ref = ref.step(s_ref, settings)
    m_gpu = gpu.step(sol, s_gpu, settings)
    return s_gpu, s_ref, m_gpu, m_ref


def test_free_fall_matches_damped_implicit_euler(gpu):
    # tests/gpu_motion.rs:140-190: mean displacement within 1e-5 of
    # g dt^2 / (1 + c dt), per vertex within 2e-5 (f32 device positions).
    from clothcraft_amd import VbdSolver
    for nx, ny in ((6, 4), (20, 12)):
        sc = scenes.free_panel(nx, ny)
        st = sc.state.copy()
        sol = VbdSolver(sc.garment, None, st)
        s = dict(sc.settings, dt=1 / 60, iterations=8)
        old = st.x.copy()
        gpu.step(sol, st, s)
        exp = np.array(s["gravity"]) * s["dt"] ** 2 / (1 + s["damping"] * s["dt"])
        d = st.x - old
        assert np.linalg.norm(d.mean(axis=0) - exp) < 1e-5
        assert np.abs(d - exp).max() < 2e-5


def test_one_step_parity_hanging_sheet(gpu):
    sc = scenes.config0()
    sg, sr, mg, mr = _one_step(gpu, sc)
    err = np.abs(sg.x - sr.x).max() / _bbox_diag(sr.x)
    print("config0 one-step rel err", err, "gpu", mg["residual_rms_n"], "ref", mr["residual_rms_n"],
          "colors", mg["colors"], mr["colors"])
    assert mg["colors"] == mr["colors"]
    assert err <= 1e-5


def test_one_step_parity_sheet_on_sphere_small(gpu):
    sc = scenes.small_sheet_on_sphere(16, 0.2535)
    sg, sr, mg, mr = _one_step(gpu, sc)
    err = np.abs(sg.x - sr.x).max() / _bbox_diag(sr.x)
    print("sphere16 one-step rel err", err, mg["contact_terms"], mr["contact_terms"],
          mg["detection_updates"], mr["detection_updates"])
    assert mg["contact_terms"] == mr["contact_terms"]
    assert err <= 1e-5


def test_invalid_grab_is_transactional(gpu):
    # tests/vbd.rs:98-127
    from clothcraft_amd import VbdError, VbdSolver
    sc = scenes.free_panel(2, 2)
    st = sc.state.copy()
    sol = VbdSolver(sc.garment, None, st)
    tri = [int(v) for v in sc.garment.tri_vertices[0]]
    center = st.x[tri].mean(axis=0)
    grab = dict(vertices=tri, weights=[1 / 3] * 3, target=tuple(center + [0, 0, 0.01]), stiffness=100.0)
    s = dict(sc.settings, gravity=(0.0, 0.0, 0.0), iterations=6)
    gpu.step(sol, st, s, grab)
    assert st.x[tri].mean(axis=0)[2] > center[2]
    old = st.copy()
    bad = dict(grab, target=(float("nan"), 0.0, 0.0))
    with pytest.raises(VbdError) as e:
        gpu.step(sol, st, s, bad)
    assert e.value.message == "Invalid mouse attachment"
    assert np.array_equal(st.x, old.x) and st.time == old.time
