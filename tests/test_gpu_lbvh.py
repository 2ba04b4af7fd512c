"""The LBVH broad phase at the garment's full size: the same candidate pairs
as the hashed grid (itself bit-exact against the oracle's pruned candidate
walk at the sizes the oracle runs), so device steps with either broad phase
are bitwise identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sorted_pairs(p):
    return p[np.lexsort((p[:, 1], p[:, 0]))] if len(p) else p


def test_lbvh_candidates_equal_the_grid_on_the_full_garment(gpu):
    from clothcraft_amd import OgcSettings, VbdSolver, garments
    sc = garments.sewn_garment_on_body()
    st = sc.state.copy()
    sol = VbdSolver(sc.garment, sc.avatar, st)
    gpu.run(sol, st, sc.settings, 10)
    r = sc.settings["radius"]
    got = {}
    try:
        for broad in ("grid", "lbvh"):
            gpu.set_broad_phase(broad)
            got[broad] = gpu.detect(sol, st.x, 2.5 * r, OgcSettings(radius=r), capacity=1 << 23)
    finally:
        gpu.set_broad_phase("grid")
    a, b = got["grid"], got["lbvh"]
    print(f"vf {len(a['vf'])} ee {len(a['ee'])} contacts {len(a['kind_ids'])}")
    assert len(a["vf"]) > 0 and len(a["kind_ids"]) > 0
    assert np.array_equal(_sorted_pairs(a["vf"]), _sorted_pairs(b["vf"]))
    assert np.array_equal(_sorted_pairs(a["ee"]), _sorted_pairs(b["ee"]))
    assert np.array_equal(a["kind_ids"], b["kind_ids"])
    assert np.array_equal(a["eps"].view(np.uint64), b["eps"].view(np.uint64))
    assert np.array_equal(a["radii"].view(np.uint64), b["radii"].view(np.uint64))
    assert a["minimum_distance"] == b["minimum_distance"]


def test_lbvh_steps_are_bitwise_the_grid_steps(gpu):
    from clothcraft_amd import VbdSolver, garments
    sc = garments.sewn_garment_on_body()
    st = sc.state.copy()
    sol = VbdSolver(sc.garment, sc.avatar, st)
    gpu.run(sol, st, sc.settings, 8)
    out, met = {}, {}
    try:
        for broad in ("grid", "lbvh"):
            gpu.set_broad_phase(broad)
            s = st.copy()
            m = gpu.run(VbdSolver(sc.garment, sc.avatar, s), s, sc.settings, 4)
            out[broad], met[broad] = s, m
    finally:
        gpu.set_broad_phase("grid")
    g, l = met["grid"], met["lbvh"]
    print(f"grid: {g['detection_updates']} detections, {g['gpu_compute_ms']:.1f} ms; "
          f"lbvh: {l['detection_updates']} detections, {l['gpu_compute_ms']:.1f} ms, "
          f"{l['broad_phase_fallbacks']} grid fallbacks")
    assert l["broad_phase_fallbacks"] == 0
    assert g["detection_updates"] == l["detection_updates"]
    assert g["contact_terms"] == l["contact_terms"] > 0
    assert np.array_equal(out["grid"].x.view(np.uint64), out["lbvh"].x.view(np.uint64))
    assert np.array_equal(out["grid"].v.view(np.uint64), out["lbvh"].v.view(np.uint64))


def test_lbvh_from_a_fresh_context_through_contact():
    """A context whose every detection uses the LBVH (no grid detection ever
    allocated the grid's arrays) steps a sheet onto the sphere through its
    first contacts exactly as a grid context does: the contact dedup sort has
    scratch of its own, sized to the raw contact count."""
    from clothcraft_amd import GpuVbd, VbdSolver, scenes
    sc = scenes.small_sheet_on_sphere(16, 0.2535)
    out, met = {}, {}
    for broad in ("lbvh", "grid"):
        gpu = GpuVbd(0)
        gpu.set_broad_phase(broad)
        s = sc.state.copy()
        met[broad] = gpu.run(VbdSolver(sc.garment, sc.avatar, s), s, sc.settings, 3)
        out[broad] = s
    print(f"lbvh: {met['lbvh']['contact_terms']} contact terms, {met['lbvh']['detection_updates']} detections")
    assert met["lbvh"]["contact_terms"] == met["grid"]["contact_terms"] > 0
    assert met["lbvh"]["broad_phase_fallbacks"] == 0
    assert np.array_equal(out["grid"].x.view(np.uint64), out["lbvh"].x.view(np.uint64))
