"""The C-ABI library loads on a host without a GPU and exports every entry
point include/clothvbd.h declares (no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "clothvbd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    names = declared()
    for n in ("cv_solver_create", "cv_solver_destroy", "cv_gpu_create", "cv_gpu_step",
              "cv_gpu_set_check_gradients", "cv_gpu_step_batch", "cv_last_error", "cv_gpu_destroy"):
        assert n in names


def test_library_loads_and_exports_every_declared_symbol():
    from clothcraft_amd import vbd
    lib = vbd.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert vbd.version().startswith("clothvbd")


def test_library_is_built_for_gfx950():
    path = os.path.join(ROOT, "clothcraft_amd", "libclothvbd.so")
    blob = open(path, "rb").read()
    assert b"gfx950" in blob
