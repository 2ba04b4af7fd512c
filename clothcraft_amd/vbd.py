"""Python mirror of the reference stepping API over libclothvbd's C ABI.

`VbdSolver(garment, avatar, state)` mirrors `VbdSolver::new`
(src/vbd.rs:101-192); `GpuVbd(device).step(solver, state, settings, grab)`
mirrors `GpuVbd::step` (src/gpu_vbd.rs:125-472): same argument meaning, the
state is updated in place only on success, and every failure raises
`VbdError` carrying the reference's message with the state untouched.

The library must be built (python -m clothcraft_amd.build); there is no
fallback path: loading fails loudly when libclothvbd.so is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclothvbd.so")

_lib = None


class VbdError(RuntimeError):
    """A failed step or create: `status` is the cv_status code, `message` the
    reference error string (e.g. "Invalid mouse attachment")."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m clothcraft_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.cv_solver_create.argtypes = [C.POINTER(abi.cv_garment), C.POINTER(abi.cv_surface),
                                       C.POINTER(abi.cv_state), C.POINTER(vp)]
        L.cv_solver_destroy.argtypes = [vp]
        L.cv_solver_record_prediction.argtypes = [vp, abi.f64p, C.POINTER(abi.cv_state),
                                                  C.POINTER(abi.cv_settings)]
        L.cv_solver_colors.argtypes = [vp, abi.u32p, C.c_uint32, abi.u32p]
        L.cv_gpu_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.cv_gpu_destroy.argtypes = [vp]
        L.cv_gpu_set_check_gradients.argtypes = [vp, C.c_int]
        L.cv_gpu_set_lanes.argtypes = [vp, C.c_int]
        L.cv_gpu_set_batch_block.argtypes = [vp, C.c_int]
        L.cv_gpu_set_broad_phase.argtypes = [vp, C.c_int]
        L.cv_gpu_set_profiling.argtypes = [vp, C.c_int]
        L.cv_gpu_set_precision.argtypes = [vp, C.c_int]
        L.cv_gpu_step.argtypes = [vp, vp, C.POINTER(abi.cv_state), C.POINTER(abi.cv_settings),
                                  C.POINTER(abi.cv_grab), C.POINTER(abi.cv_metrics)]
        L.cv_gpu_run.argtypes = [vp, vp, C.POINTER(abi.cv_state), C.POINTER(abi.cv_settings),
                                 C.c_uint32, C.POINTER(abi.cv_metrics)]
        L.cv_gpu_probe.argtypes = [vp, vp, C.POINTER(abi.cv_state), C.POINTER(abi.cv_settings),
                                   C.POINTER(abi.cv_grab), abi.f64p]
        L.cv_gpu_detect.argtypes = [vp, vp, abi.f64p, C.c_double, C.POINTER(abi.cv_ogc_settings),
                                    C.POINTER(abi.cv_detection), abi.u32p, C.c_uint64, abi.u32p,
                                    C.c_uint64, abi.u32p, C.c_uint64, abi.f64p, abi.f64p]
        L.cv_gpu_newton.argtypes = [vp, vp, C.POINTER(abi.cv_state), C.POINTER(abi.cv_settings),
                                    C.POINTER(abi.cv_newton_settings), abi.f64p,
                                    C.POINTER(abi.cv_newton_metrics)]
        L.cv_gpu_colors.argtypes = [vp, vp, C.POINTER(abi.cv_grab), C.c_uint32, abi.u32p, C.c_uint32,
                                    abi.u32p]
        L.cv_gpu_step_batch.argtypes = [vp, C.POINTER(vp), C.POINTER(C.POINTER(abi.cv_state)),
                                        C.c_uint32, C.POINTER(abi.cv_settings),
                                        C.POINTER(abi.cv_metrics)]
        f32p = C.POINTER(C.c_float)
        L.cv_warm_set_params.argtypes = [vp, f32p]
        L.cv_warm_get_params.argtypes = [vp, f32p]
        L.cv_warm_infer.argtypes = [vp, f32p, C.c_uint64, f32p]
        L.cv_warm_gradient.argtypes = [vp, f32p, f32p, C.c_uint64, f32p, f32p]
        L.cv_warm_train.argtypes = [vp, f32p, f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                                    C.c_uint64, f32p]
        L.cv_warm_features.argtypes = [vp, vp, C.POINTER(abi.cv_state), C.POINTER(abi.cv_settings), f32p]
        L.cv_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.cv_last_error.restype = C.c_size_t
        L.cv_solver_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.cv_solver_last_error.restype = C.c_size_t
        L.cv_version.restype = C.c_char_p
        _lib = L
    return _lib


def _error(handle=None) -> str:
    buf = C.create_string_buffer(1024)
    lib().cv_last_error(handle, buf, 1024)
    return buf.value.decode()


@dataclass
class OgcSettings:
    """src/ogc/world.rs:33-54."""
    radius: float = 0.003
    stiffness: float = 100000.0
    relaxation: float = 0.45
    refresh_fraction: float = 0.01

    def to_c(self):
        return abi.cv_ogc_settings(self.radius, self.stiffness, self.relaxation,
                                   self.refresh_fraction)


@dataclass
class VbdSettings:
    """src/vbd.rs:18-36 (+ the library's extensions, zero = reference behaviour)."""
    dt: float = 1.0 / 240.0
    iterations: int = 16
    damping: float = 1.5
    gravity: tuple = (0.0, -9.81, 0.0)
    contact: OgcSettings = field(default_factory=OgcSettings)
    residual_tolerance: float = 0.0
    flags: int = 0
    residual_interval: int = 0

    @staticmethod
    def from_dict(d: dict) -> "VbdSettings":
        return VbdSettings(d["dt"], d["iterations"], d["damping"], tuple(d["gravity"]),
                           OgcSettings(d["radius"], d["stiffness"], d["relaxation"],
                                       d["refresh_fraction"]),
                           d.get("residual_tolerance", 0.0), d.get("flags", 0),
                           d.get("residual_interval", 0))

    def to_c(self):
        return abi.cv_settings(self.dt, self.iterations, self.damping,
                               (C.c_double * 3)(*self.gravity), self.contact.to_c(),
                               self.residual_tolerance, self.flags, self.residual_interval)


def _settings(s) -> abi.cv_settings:
    if isinstance(s, dict):
        s = VbdSettings.from_dict(s)
    return s.to_c()


@dataclass
class Grab:
    """src/vbd.rs:54-60."""
    vertices: tuple
    weights: tuple
    target: tuple
    stiffness: float

    def to_c(self):
        return abi.cv_grab((C.c_uint32 * 3)(*self.vertices), (C.c_double * 3)(*self.weights),
                           (C.c_double * 3)(*self.target), self.stiffness)


def _grab(g):
    if g is None:
        return None
    if isinstance(g, dict):
        g = Grab(**g)
    return g.to_c()


class VbdSolver:
    """VbdSolver::new (src/vbd.rs:101-192): validates the garment, pins and the
    start state (no intersection, positive separation) and builds the term
    incidence, conflict graph and OGC topology."""

    def __init__(self, garment, avatar, state):
        L = lib()
        self._g = garment.to_c()
        self._a = avatar.to_c() if avatar is not None else None
        h = C.c_void_p()
        rc = L.cv_solver_create(C.byref(self._g), C.byref(self._a) if self._a else None,
                                C.byref(state.to_c()), C.byref(h))
        if rc != 0:
            raise VbdError(rc, _error(None))
        self.h = h
        self.n = garment.n
        self.garment = garment
        self.avatar = avatar

    def __del__(self):
        if getattr(self, "h", None):
            lib().cv_solver_destroy(self.h)
            self.h = None

    def record_prediction(self, old_velocity, state, settings):
        """Prediction::record (src/vbd/prediction.rs:43-54)."""
        ov = np.ascontiguousarray(old_velocity, np.float64)
        rc = lib().cv_solver_record_prediction(self.h, abi.ptr(ov, C.c_double),
                                               C.byref(state.to_c()), C.byref(_settings(settings)))
        if rc != 0:
            buf = C.create_string_buffer(512)
            lib().cv_solver_last_error(self.h, buf, 512)
            raise VbdError(rc, buf.value.decode())

    def static_colors(self):
        out = np.zeros(self.n, np.uint32)
        k = C.c_uint32()
        lib().cv_solver_colors(self.h, abi.ptr(out, C.c_uint32), self.n, C.byref(k))
        return out, int(k.value)


class GpuVbd:
    """GpuVbd (src/gpu_vbd.rs:15-124) on one MI355X: one HIP stream, device
    buffers reused across steps, static term records cached per solver."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().cv_gpu_create(device, C.byref(h))
        if rc != 0:
            raise VbdError(rc, _error(None))
        self.h = h
        self._check = False

    def __del__(self):
        if getattr(self, "h", None):
            lib().cv_gpu_destroy(self.h)
            self.h = None

    @property
    def warm(self) -> "WarmStart":
        """The learned warm-start network of this device context."""
        return WarmStart(self)

    @property
    def check_gradients(self) -> bool:
        return self._check

    @check_gradients.setter
    def check_gradients(self, v: bool):
        self._check = bool(v)
        lib().cv_gpu_set_check_gradients(self.h, int(v))

    def set_lanes(self, lanes: int):
        rc = lib().cv_gpu_set_lanes(self.h, lanes)
        if rc != 0:
            raise VbdError(rc, "lanes must be 8, 16, 32 or 64")

    def set_batch_block(self, threads: int):
        """Threads per block of the batched step's shared sweep launches."""
        rc = lib().cv_gpu_set_batch_block(self.h, threads)
        if rc != 0:
            raise VbdError(rc, "batch block must be 64, 128 or 256 threads")

    BROAD_PHASES = {"grid": 0, "lbvh": 1}

    def set_broad_phase(self, kind: str):
        """OGC broad phase: "grid" (default hashed grid) or "lbvh" (linear BVH
        rebuilt per step, refitted at refreshes).  Both give the same
        candidate pairs."""
        if kind not in self.BROAD_PHASES:
            raise ValueError(f"broad phase must be one of {sorted(self.BROAD_PHASES)}")
        rc = lib().cv_gpu_set_broad_phase(self.h, self.BROAD_PHASES[kind])
        if rc != 0:
            raise VbdError(rc, "invalid broad phase")

    def set_precision(self, bits: int):
        """64 (default, reference parity) or 32 (faster f32 sweep)."""
        rc = lib().cv_gpu_set_precision(self.h, bits)
        if rc != 0:
            raise VbdError(rc, "precision must be 32 or 64")

    def set_profiling(self, on: bool):
        lib().cv_gpu_set_profiling(self.h, int(on))

    def _raise(self, rc):
        raise VbdError(rc, _error(self.h))

    def step(self, solver: VbdSolver, state, settings, grab=None) -> dict:
        st = state.to_c()
        m = abi.cv_metrics()
        g = _grab(grab)
        rc = lib().cv_gpu_step(self.h, solver.h, C.byref(st), C.byref(_settings(settings)),
                               C.byref(g) if g else None, C.byref(m))
        if rc != 0:
            self._raise(rc)
        state.time = st.time
        return m.as_dict()

    def run(self, solver: VbdSolver, state, settings, steps: int) -> dict:
        st = state.to_c()
        m = abi.cv_metrics()
        rc = lib().cv_gpu_run(self.h, solver.h, C.byref(st), C.byref(_settings(settings)), steps,
                              C.byref(m))
        if rc != 0:
            self._raise(rc)
        state.time = st.time
        return m.as_dict()

    def step_batch(self, solvers, states, settings) -> list:
        """One step of many independent solvers (cv_gpu_step_batch); states
        are updated in place; returns per-solver metrics."""
        n = len(solvers)
        hs = (C.c_void_p * n)(*[s.h for s in solvers])
        cs = [st.to_c() for st in states]
        ps = (C.POINTER(abi.cv_state) * n)(*[C.pointer(c) for c in cs])
        ms = (abi.cv_metrics * n)()
        rc = lib().cv_gpu_step_batch(self.h, hs, ps, n, C.byref(_settings(settings)), ms)
        for st, c in zip(states, cs):
            st.time = c.time
        if rc != 0:
            self._raise(rc)
        return [m.as_dict() for m in ms]

    def probe(self, solver: VbdSolver, state, settings, grab=None) -> np.ndarray:
        out = np.zeros((solver.n, 4))
        g = _grab(grab)
        rc = lib().cv_gpu_probe(self.h, solver.h, C.byref(state.to_c()),
                                C.byref(_settings(settings)), C.byref(g) if g else None,
                                abi.ptr(out, C.c_double))
        if rc != 0:
            self._raise(rc)
        return out

    def newton(self, solver: VbdSolver, start, settings, newton_iterations: int = 30,
               pcg_iterations: int = 250, absolute_tolerance: float = 1e-6,
               relative_tolerance: float = 1e-4):
        """Newton energy reference of one step on the device
        (ReferenceSolver::step, src/solver.rs:422-647; defaults of
        SolverSettings, src/solver.rs:25-36): minimises the step objective
        from `start` with the contacts of this solver's last device step or
        detection frozen.  Returns (x, metrics); `start` is not changed."""
        x = np.zeros((solver.n, 3))
        ns = abi.cv_newton_settings(newton_iterations, pcg_iterations, absolute_tolerance, relative_tolerance)
        m = abi.cv_newton_metrics()
        rc = lib().cv_gpu_newton(self.h, solver.h, C.byref(start.to_c()), C.byref(_settings(settings)),
                                 C.byref(ns), abi.ptr(x, C.c_double), C.byref(m))
        if rc != 0:
            self._raise(rc)
        return x, {name: getattr(m, name) for name, _ in m._fields_}

    def detect(self, solver: VbdSolver, x, rq, ogc: OgcSettings | None = None,
               candidates: bool = True, capacity: int = 1 << 22) -> dict:
        """Device OGC detection on f64 positions x (moving part): counts,
        canonical contacts, radii and (optionally) the candidate pairs."""
        ogc = ogc or OgcSettings()
        x = np.ascontiguousarray(x, np.float64)
        d = abi.cv_detection()
        vf = np.zeros((capacity, 2), np.uint32) if candidates else None
        ee = np.zeros((capacity, 2), np.uint32) if candidates else None
        cki = np.zeros((capacity, 5), np.uint32)
        ceps = np.zeros(capacity)
        radii = np.zeros(solver.n)
        rc = lib().cv_gpu_detect(self.h, solver.h, abi.ptr(x, C.c_double), rq,
                                 C.byref(ogc.to_c()), C.byref(d), abi.ptr(vf, C.c_uint32),
                                 capacity if candidates else 0, abi.ptr(ee, C.c_uint32),
                                 capacity if candidates else 0, abi.ptr(cki, C.c_uint32),
                                 capacity, abi.ptr(ceps, C.c_double), abi.ptr(radii, C.c_double))
        if rc != 0:
            self._raise(rc)
        out = {name: getattr(d, name) for name, _ in d._fields_}
        nc = int(d.n_contacts)
        out["kind_ids"] = cki[:nc].copy()
        out["eps"] = ceps[:nc].copy()
        out["radii"] = radii
        if candidates:
            out["vf"] = vf[:int(d.n_vf_candidates)].copy()
            out["ee"] = ee[:int(d.n_ee_candidates)].copy()
        return out

    def colors(self, solver: VbdSolver, grab=None, fast: bool = False):
        """Colouring of the last detect()'s contacts (and the grab) as a step
        colours them: the reference's greedy in vertex order, or the fast
        colouring on the device.  Returns (colour of each vertex, count)."""
        out = np.zeros(solver.n, np.uint32)
        k = C.c_uint32()
        g = _grab(grab)
        rc = lib().cv_gpu_colors(self.h, solver.h, C.byref(g) if g else None,
                                 abi.CV_FLAG_FAST_COLORING if fast else 0, abi.ptr(out, C.c_uint32),
                                 solver.n, C.byref(k))
        if rc != 0:
            self._raise(rc)
        return out, int(k.value)


WARM_PARAMS = 35075


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class WarmStart:
    """The learned warm start on one GpuVbd (no reference counterpart):
    parameters, device inference, the device backward pass and training."""

    def __init__(self, gpu: "GpuVbd"):
        self.gpu = gpu

    def _call(self, rc):
        if rc != 0:
            self.gpu._raise(rc)

    def set_params(self, params):
        p = _f32(params)
        assert p.size == WARM_PARAMS
        self._call(lib().cv_warm_set_params(self.gpu.h, abi.ptr(p, C.c_float)))

    def set_smoothing(self, passes: int):
        """One-ring averaging passes of the network's correction (cv_warm_set_smoothing)."""
        L = lib()
        L.cv_warm_set_smoothing.argtypes = [C.c_void_p, C.c_int]
        self._call(L.cv_warm_set_smoothing(self.gpu.h, int(passes)))

    def set_correction(self, scale: float = 1.0, normal_only: bool = False):
        """Factor of the network's correction and whether only its component
        along the vertex normal is kept (cv_warm_set_correction)."""
        L = lib()
        L.cv_warm_set_correction.argtypes = [C.c_void_p, C.c_double, C.c_int]
        self._call(L.cv_warm_set_correction(self.gpu.h, float(scale), int(normal_only)))

    def params(self) -> np.ndarray:
        out = np.zeros(WARM_PARAMS, np.float32)
        self._call(lib().cv_warm_get_params(self.gpu.h, abi.ptr(out, C.c_float)))
        return out

    def infer(self, features) -> np.ndarray:
        X = _f32(features).reshape(-1, 12)
        out = np.zeros((len(X), 3), np.float32)
        self._call(lib().cv_warm_infer(self.gpu.h, abi.ptr(X, C.c_float), len(X), abi.ptr(out, C.c_float)))
        return out

    def gradient(self, features, targets):
        X, T = _f32(features).reshape(-1, 12), _f32(targets).reshape(-1, 3)
        g = np.zeros(WARM_PARAMS, np.float32)
        loss = C.c_float()
        self._call(lib().cv_warm_gradient(self.gpu.h, abi.ptr(X, C.c_float), abi.ptr(T, C.c_float), len(X),
                                          abi.ptr(g, C.c_float), C.byref(loss)))
        return g, float(loss.value)

    def train(self, features, targets, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1):
        X, T = _f32(features).reshape(-1, 12), _f32(targets).reshape(-1, 3)
        losses = np.zeros(epochs, np.float32)
        self._call(lib().cv_warm_train(self.gpu.h, abi.ptr(X, C.c_float), abi.ptr(T, C.c_float), len(X), epochs,
                                       batch, lr, seed, abi.ptr(losses, C.c_float)))
        return losses

    def features(self, solver: "VbdSolver", state, settings) -> np.ndarray:
        out = np.zeros((solver.n, 12), np.float32)
        self._call(lib().cv_warm_features(self.gpu.h, solver.h, C.byref(state.to_c()),
                                          C.byref(_settings(settings)), abi.ptr(out, C.c_float)))
        return out


def version() -> str:
    return lib().cv_version().decode()
