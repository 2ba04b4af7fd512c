"""Build libclothvbd.so (HIP, gfx950) and libcvm_host.so (host build of the
device math, for CPU tests) in-tree with hipcc / g++."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclothvbd.so")
HOST_LIB = os.path.join(HERE, "libcvm_host.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CV_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-I", CSRC]
# Sources and their extra flags.  Detection is f64 without FMA contraction so
# it rounds like the reference; the sweep may contract.
SOURCES = [
    ("vbd_device.hip", []),
    ("ogc_device.hip", ["-ffp-contract=off"]),
    ("mlp_device.hip", []),
    ("color_device.hip", []),
    ("newton_device.hip", []),
    ("solver.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("api.cpp", ["-x", "hip"]),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + out + err)
            raise RuntimeError(f"hipcc failed on {cmd[-3]}")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """Host (x86) build of vbd_math.hpp / ogc_math.hpp behind a small C ABI."""
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I", CSRC,
           "-Wno-unknown-pragmas", os.path.join(CSRC, "host_mirror.cpp"),
           os.path.join(CSRC, "solver.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
