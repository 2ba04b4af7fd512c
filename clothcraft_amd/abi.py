"""ctypes mirror of include/clothvbd.h (plain data structs of the C ABI).

The same structs feed the product library (libclothvbd.so) and, in tests,
the f64 oracle, so both see identical input arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)

CV_OK = 0
STATUS_NAMES = {
    0: "CV_OK", 1: "CV_ERR_INVALID_ARGUMENT", 2: "CV_ERR_INVALID_SETTINGS",
    3: "CV_ERR_INVALID_OGC_SETTINGS", 4: "CV_ERR_INVALID_STATE", 5: "CV_ERR_INVALID_GRAB",
    6: "CV_ERR_DIMENSIONS", 7: "CV_ERR_PIN_MISMATCH", 8: "CV_ERR_TOPOLOGY",
    9: "CV_ERR_DEGENERATE_FACE", 10: "CV_ERR_INVALID_AVATAR", 11: "CV_ERR_INITIAL_INTERSECTION",
    12: "CV_ERR_INITIAL_TOUCHING", 13: "CV_ERR_ZERO_SEPARATION", 14: "CV_ERR_SEPARATION_F32",
    15: "CV_ERR_NONFINITE_RESULT", 16: "CV_ERR_TRUST_REGION", 17: "CV_ERR_FIXED_CHANGED",
    18: "CV_ERR_QUERY_RADIUS", 19: "CV_ERR_INVALID_POSITIONS", 20: "CV_ERR_DEVICE",
    21: "CV_ERR_CAPACITY", 22: "CV_ERR_INVALID_SEAM", 23: "CV_ERR_NOT_BUILT",
    24: "CV_ERR_INVALID_MATERIAL", 25: "CV_ERR_LINE_SEARCH", 26: "CV_ERR_GRADIENT_CHECK",
}

CV_FLAG_DIAGNOSTICS_OFF = 1 << 0
CV_FLAG_TRUE_RESIDUAL = 1 << 1
CV_FLAG_FAST_COLORING = 1 << 2
CV_FLAG_WARM_START = 1 << 3


class cv_material(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "warp", "weft", "shear", "compression_ratio", "compression_transition",
        "bend_warp", "bend_weft", "density", "thickness", "grain_degrees")]


class cv_garment(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_uint32), ("masses", f64p), ("pinned", u8p), ("pin_targets", f64p),
        ("n_triangles", C.c_uint32), ("tri_vertices", u32p), ("tri_inv_dm", f64p),
        ("tri_area", f64p), ("tri_material", u32p),
        ("n_hinges", C.c_uint32), ("hinge_vertices", u32p), ("hinge_weight", f64p),
        ("hinge_across", f64p), ("hinge_material", u32p),
        ("n_materials", C.c_uint32), ("materials", C.POINTER(cv_material)),
        ("n_stitches", C.c_uint32), ("stitch_offsets", u32p), ("stitch_vertices", u32p),
        ("stitch_weights", f64p), ("stitch_stiffness", f64p),
        ("n_probes", C.c_uint32), ("probe_offsets", u32p), ("probe_vertices", u32p),
        ("probe_weights", f64p),
    ]


class cv_surface(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("vertices", f64p),
                ("n_triangles", C.c_uint32), ("triangles", u32p)]


class cv_state(C.Structure):
    _fields_ = [("x", f64p), ("v", f64p), ("n", C.c_uint32), ("time", C.c_double)]


class cv_ogc_settings(C.Structure):
    _fields_ = [("radius", C.c_double), ("stiffness", C.c_double),
                ("relaxation", C.c_double), ("refresh_fraction", C.c_double)]


class cv_settings(C.Structure):
    _fields_ = [("dt", C.c_double), ("iterations", C.c_uint32), ("damping", C.c_double),
                ("gravity", C.c_double * 3), ("contact", cv_ogc_settings),
                ("residual_tolerance", C.c_double), ("flags", C.c_uint32),
                ("residual_interval", C.c_uint32)]


class cv_grab(C.Structure):
    _fields_ = [("vertices", C.c_uint32 * 3), ("weights", C.c_double * 3),
                ("target", C.c_double * 3), ("stiffness", C.c_double)]


class cv_metrics(C.Structure):
    _fields_ = [
        ("time_s", C.c_double), ("dt_s", C.c_double), ("step_ms", C.c_double),
        ("detection_ms", C.c_double), ("solve_ms", C.c_double), ("gpu_compute_ms", C.c_double),
        ("refit_ms", C.c_double), ("candidates_ms", C.c_double), ("narrow_ms", C.c_double),
        ("candidate_pairs", C.c_uint64), ("aabb_rejections", C.c_uint64),
        ("iterations", C.c_uint32),
        ("contacts", C.c_uint64), ("contact_terms", C.c_uint64), ("detection_updates", C.c_uint64),
        ("broad_phase_rebuilds", C.c_uint64), ("truncated_vertices", C.c_uint64),
        ("minimum_distance_m", C.c_double), ("max_seam_gap_m", C.c_double),
        ("extension_p95_percent", C.c_double), ("extension_max_percent", C.c_double),
        ("residual_rms_n", C.c_double),
        ("true_residual_rms_n", C.c_double), ("colors", C.c_uint32), ("iterations_run", C.c_uint32),
        ("sweep_ms", C.c_double), ("sweep_launches", C.c_uint64), ("sweep_bytes", C.c_double),
        ("f64", C.c_uint32), ("sweep_bytes_min", C.c_double), ("steps", C.c_uint32),
        ("sweep_timed_launches", C.c_uint64), ("sweep_timed_bytes_min", C.c_double),
        ("sweep_event_overhead_ms", C.c_double), ("broad_phase_fallbacks", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class cv_detection(C.Structure):
    _fields_ = [("n_vf_candidates", C.c_uint64), ("n_ee_candidates", C.c_uint64),
                ("vertex_face_candidates", C.c_uint64), ("edge_edge_candidates", C.c_uint64),
                ("n_contacts", C.c_uint64), ("active_contacts", C.c_uint64),
                ("minimum_distance", C.c_double)]


class cv_newton_settings(C.Structure):
    _fields_ = [("newton_iterations", C.c_uint32), ("pcg_iterations", C.c_uint32),
                ("absolute_tolerance", C.c_double), ("relative_tolerance", C.c_double)]


class cv_newton_metrics(C.Structure):
    _fields_ = [("newton_iterations", C.c_uint32), ("pcg_iterations", C.c_uint32),
                ("line_search_backtracks", C.c_uint32), ("converged", C.c_uint32),
                ("residual_rms_n", C.c_double), ("tolerance_n", C.c_double),
                ("objective_initial_j", C.c_double), ("objective_final_j", C.c_double),
                ("max_regularization", C.c_double)]


def ptr(a: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (or NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))
