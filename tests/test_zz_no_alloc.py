"""cv_warm_infer_noalloc, the allocation-free variant of cv_warm_infer: the
same predictions through the C ABI, on two calls in a row (the second with
fewer rows, so stale rows of the shared staging array are present), at the
row limit, and CV_ERR_CAPACITY above it."""
import ctypes as C

import numpy as np
import pytest

from clothcraft_amd import abi, warmstart
from clothcraft_amd.vbd import WarmStart, lib

pytestmark = pytest.mark.gpu

MAX_ROWS = 262144  # CV_WARM_NOALLOC_MAX_ROWS


def _noalloc(gpu, X):
    L = lib()
    L.cv_warm_infer_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_float)]
    L.cv_warm_infer_noalloc.restype = C.c_int
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 12)
    out = np.full((len(X), 3), np.nan, np.float32)
    rc = L.cv_warm_infer_noalloc(gpu.h, abi.ptr(X, C.c_float), len(X), abi.ptr(out, C.c_float))
    return rc, out


def test_noalloc_inference_equals_cv_warm_infer(gpu):
    w = WarmStart(gpu)
    w.set_params(warmstart.init_params(seed=11, out_scale=0.5))
    rng = np.random.default_rng(4)
    for n in (100003, 777, MAX_ROWS):  # a large call, then a smaller one, then the limit
        X = rng.normal(0.0, 1.0, (n, 12)).astype(np.float32)
        want = w.infer(X)
        rc, got = _noalloc(gpu, X)
        print(f"noalloc inference of {n} rows: status {rc}, max |diff| {np.abs(got - want).max():.3e}")
        assert rc == 0
        assert np.array_equal(got, want)
    big = np.zeros((MAX_ROWS + 1, 12), np.float32)
    rc, _ = _noalloc(gpu, big)
    assert abi.STATUS_NAMES[rc] == "CV_ERR_CAPACITY"
