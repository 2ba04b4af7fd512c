"""ctypes binding of libcvm_host.so: the host (x86) build of the product's
solver set-up and device math headers, used by the CPU test suite."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcvm_host.so")

TERM_DTYPE = np.dtype([("head", "<u4"), ("ids", "<u4", 4), ("f", "<f4", 11)])
assert TERM_DTYPE.itemsize == 64
TERM64_DTYPE = np.dtype({"names": ["head", "ids", "f"], "formats": ["<u4", ("<u4", 4), ("<f8", 11)],
                         "offsets": [0, 4, 24], "itemsize": 112})

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build_host()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.cvh_solver_create.argtypes = [C.POINTER(abi.cv_garment), C.POINTER(abi.cv_surface),
                                        C.POINTER(abi.cv_state), C.POINTER(vp), C.c_char_p, C.c_size_t]
        L.cvh_solver_destroy.argtypes = [vp]
        L.cvh_solver_terms.argtypes = [vp, abi.u32p, C.c_void_p, C.c_uint64]
        L.cvh_solver_terms.restype = C.c_uint64
        L.cvh_solver_edges.argtypes = [vp, abi.u32p, C.c_uint32]
        L.cvh_solver_edges.restype = C.c_uint32
        L.cvh_greedy_colors.argtypes = [vp, abi.u32p, abi.u32p, C.c_uint64, C.POINTER(abi.cv_grab),
                                        abi.u32p, C.c_int]
        L.cvh_greedy_colors.restype = C.c_uint32
        L.cvh_coloring_new.restype = C.c_void_p
        L.cvh_coloring_free.argtypes = [C.c_void_p]
        L.cvh_coloring_update.argtypes = [vp, C.c_void_p, abi.u32p, abi.u32p, C.c_uint64, C.POINTER(abi.cv_grab),
                                          C.c_int, abi.u32p]
        L.cvh_coloring_update.restype = C.c_uint32
        for name in ("cvh_term_deriv_f64", "cvh_term_deriv_f32"):
            getattr(L, name).argtypes = [C.c_void_p, abi.f64p, abi.f64p, C.c_double, C.c_double,
                                         abi.f64p]
        for name in ("cvh_term_energy_f64", "cvh_term_energy_f32"):
            getattr(L, name).argtypes = [C.c_void_p, abi.f64p, abi.f64p, C.c_double, C.c_double]
            getattr(L, name).restype = C.c_double
        L.cvh_contact_record.argtypes = [C.c_int, abi.u32p, C.c_int, C.c_double, C.c_uint32,
                                         C.c_void_p]
        L.cvh_newton_direction_f64.argtypes = [abi.f64p, abi.f64p, abi.f64p, C.c_double, abi.f64p]
        L.cvh_closest_triangle.argtypes = [abi.f64p, abi.f64p, abi.f64p, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        L.cvh_closest_segments.argtypes = [abi.f64p] * 4 + [abi.f64p, abi.f64p]
        _lib = L
    return _lib


def _p(a):
    return abi.ptr(a, C.c_double)


class HostSolver:
    """The product's cv_solver set-up, built for the host."""

    def __init__(self, garment, avatar, state):
        self._g = garment.to_c()
        self._a = avatar.to_c() if avatar is not None else None
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().cvh_solver_create(C.byref(self._g), C.byref(self._a) if self._a else None,
                                     C.byref(state.to_c()), C.byref(h), err, 512)
        if rc != 0:
            from .vbd import VbdError
            raise VbdError(rc, err.value.decode())
        self.h = h
        self.n = garment.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().cvh_solver_destroy(self.h)
            self.h = None

    def terms(self):
        off = np.zeros(self.n + 1, np.uint32)
        k = lib().cvh_solver_terms(self.h, abi.ptr(off, C.c_uint32), None, 0)
        recs = np.zeros(k, TERM_DTYPE)
        lib().cvh_solver_terms(self.h, abi.ptr(off, C.c_uint32), recs.ctypes.data, k)
        return off, recs

    def check_probe(self, x, targets, probe, settings, grab=None, origin=(0.0, 0.0, 0.0), coff=None,
                    crec=None):
        """The step's runtime gradient check on the host (check_probe):
        positions x [n_all,3] and targets [n,4] relative to origin, probe
        [n,4] (g, E).  Returns (status, message, max_error)."""
        L = lib()
        L.cvh_check_probe.argtypes = [C.c_void_p, abi.f64p, abi.f64p, abi.u32p, C.c_void_p, abi.f64p,
                                      C.POINTER(abi.cv_settings), C.POINTER(abi.cv_grab), abi.f64p,
                                      C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
        from .vbd import _grab as grab_c, _settings as settings_c
        xx = np.ascontiguousarray(x, np.float64)
        tt = np.ascontiguousarray(targets, np.float64)
        pp = np.ascontiguousarray(probe, np.float64)
        oo = np.ascontiguousarray(origin, np.float64)
        co = np.zeros(self.n + 1, np.uint32) if coff is None else np.ascontiguousarray(coff, np.uint32)
        cr = np.zeros(1, TERM64_DTYPE) if crec is None else crec
        mx = C.c_double(0.0)
        err = C.create_string_buffer(1024)
        g = grab_c(grab)
        rc = L.cvh_check_probe(self.h, _p(xx), _p(tt), abi.ptr(co, C.c_uint32), cr.ctypes.data, _p(oo),
                               C.byref(settings_c(settings)), C.byref(g) if g else None, _p(pp),
                               C.byref(mx), err, 1024)
        return rc, err.value.decode(), mx.value

    def terms64(self):
        L = lib()
        L.cvh_solver_terms64.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.cvh_solver_terms64.restype = C.c_uint64
        k = L.cvh_solver_terms64(self.h, None, 0)
        recs = np.zeros(k, TERM64_DTYPE)
        L.cvh_solver_terms64(self.h, recs.ctypes.data, k)
        return recs

    def edges(self):
        out = np.zeros((1 << 21, 2), np.uint32)
        k = lib().cvh_solver_edges(self.h, abi.ptr(out, C.c_uint32), 1 << 21)
        return out[:k].copy()

    def greedy_colors(self, kind_ids, counts, grab=None, smallest_last=False):
        ki = np.ascontiguousarray(kind_ids, np.uint32).reshape(-1, 5)
        cn = np.ascontiguousarray(counts, np.uint32)
        out = np.zeros(self.n, np.uint32)
        from .vbd import _grab
        g = _grab(grab)
        k = lib().cvh_greedy_colors(self.h, abi.ptr(ki, C.c_uint32), abi.ptr(cn, C.c_uint32), len(cn),
                                    C.byref(g) if g else None, abi.ptr(out, C.c_uint32), int(smallest_last))
        return out, int(k)


class IncrementalColoring:
    """update_colors (the step's colouring kept across detections) for tests."""

    def __init__(self, solver: "HostSolver"):
        self.solver = solver
        self.h = lib().cvh_coloring_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().cvh_coloring_free(self.h)
            self.h = None

    def update(self, kind_ids, counts, grab=None, smallest_last=False):
        ki = np.ascontiguousarray(kind_ids, np.uint32).reshape(-1, 5)
        cn = np.ascontiguousarray(counts, np.uint32)
        out = np.zeros(self.solver.n, np.uint32)
        from .vbd import _grab
        g = _grab(grab)
        k = lib().cvh_coloring_update(self.solver.h, self.h, abi.ptr(ki, C.c_uint32), abi.ptr(cn, C.c_uint32),
                                      len(cn), C.byref(g) if g else None, int(smallest_last),
                                      abi.ptr(out, C.c_uint32))
        return out, int(k)


def term_deriv(rec, xv, x, cr=0.003, ck=1e5, f32=False) -> np.ndarray:
    r = np.ascontiguousarray(np.asarray(rec, TERM_DTYPE).reshape(1))
    xv = np.ascontiguousarray(xv, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(10)
    fn = lib().cvh_term_deriv_f32 if f32 else lib().cvh_term_deriv_f64
    fn(r.ctypes.data, _p(xv), _p(x), cr, ck, _p(out))
    return out


def term_energy(rec, xv, x, cr=0.003, ck=1e5, f32=False) -> float:
    r = np.ascontiguousarray(np.asarray(rec, TERM_DTYPE).reshape(1))
    xv = np.ascontiguousarray(xv, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    fn = lib().cvh_term_energy_f32 if f32 else lib().cvh_term_energy_f64
    return fn(r.ctypes.data, _p(xv), _p(x), cr, ck)


def contact_record(kind, ids, count, eps, v):
    out = np.zeros(1, TERM_DTYPE)
    ids = np.ascontiguousarray(ids, np.uint32)
    lib().cvh_contact_record(kind, abi.ptr(ids, C.c_uint32), count, eps, v, out.ctypes.data)
    return out[0]


def sym(h6) -> np.ndarray:
    a = h6
    return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])


def sweep_color(order, pos, tgt, ball, term_off, terms, coff, crec, cr, ck, gsq, f32=True):
    """Host mirror of one device sweep launch (k_sweep) over `order`.
    pos/tgt/ball are float64 [n,4] arrays (positions relative to the origin,
    updated in place) in both precisions, as on the device; terms/crec are
    the records of the precision used.  In f32 each vertex solve works in f32
    relative to that vertex's own position.  Returns (clipped count,
    per-vertex attempt index)."""
    L = lib()
    fn = L.cvh_sweep_color_f32 if f32 else L.cvh_sweep_color_f64
    fp = C.POINTER(C.c_double)
    fn.argtypes = [C.c_uint32, abi.u32p, fp, fp, fp, abi.u32p, C.c_void_p, abi.u32p, C.c_void_p,
                   C.c_double, C.c_double, fp, C.POINTER(C.c_uint64), abi.u32p]
    clipped = C.c_uint64(0)
    att = np.zeros(len(pos), np.uint32)
    order = np.ascontiguousarray(order, np.uint32)
    fn(len(order), abi.ptr(order, C.c_uint32), abi.ptr(pos, C.c_double), abi.ptr(tgt, C.c_double),
       abi.ptr(ball, C.c_double), abi.ptr(term_off, C.c_uint32), terms.ctypes.data, abi.ptr(coff, C.c_uint32),
       crec.ctypes.data if len(crec) else None, cr, ck, abi.ptr(gsq, C.c_double), C.byref(clipped),
       abi.ptr(att, C.c_uint32))
    return int(clipped.value), att
