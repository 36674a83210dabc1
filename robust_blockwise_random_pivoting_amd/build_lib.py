"""Build the in-tree HIP library ``librbrp.so`` for gfx950 with hipcc.

Each source is compiled to an object in ``_build/`` (in parallel), then linked
into ``robust_blockwise_random_pivoting_amd/librbrp.so``.  The kernel-level
test hooks (``include/rbrp_testing.h``) go into a separate
``librbrp_testing.so`` linked against it, so the product library exports only
the ``include/rbrp.h`` surface.  Run as
``python -m robust_blockwise_random_pivoting_amd.build_lib``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "librbrp.so")
TESTING_LIB = os.path.join(PKG, "librbrp_testing.so")
CXX_CHECK = os.path.join(PKG, "bin", "rbrp_hpp_check")
DRIVER = os.path.join(PKG, "bin", "rbrp_driver")

SOURCES = ["k_gemm.hip", "k_rbrp.hip", "k_sketch.hip", "k_qrcp.hip", "k_mixed.hip", "k_batched.hip",
           "k_panel.hip", "rbrp_host.cpp", "rbrp_sketch.cpp", "rbrp_cpqr.cpp", "rbrp_matrices.cpp", "rbrp_batched.cpp"]
TESTING_SOURCES = ["rbrp_testing.cpp", "k_ceiling.hip"]

ARCH = os.environ.get("RBRP_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def compile_flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        "-Werror=uninitialized",
        # no packed-FP32 instructions anywhere in the library
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
    ]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + compile_flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES) + len(TESTING_SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES + TESTING_SOURCES))
    tobjs = objs[len(SOURCES):]
    objs = objs[:len(SOURCES)]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # test hooks: their own library, resolving the kernel launchers in librbrp.so
    tmp = TESTING_LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + tobjs + \
        ["-L", PKG, "-lrbrp", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test-hook library link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TESTING_LIB)
    # C++ drop-in check of include/rbrp.hpp, linked against the library
    os.makedirs(os.path.dirname(CXX_CHECK), exist_ok=True)
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cxx", "rbrp_hpp_check.cpp"), "-L", PKG, "-lrbrp",
           "-Wl,-rpath,$ORIGIN/..", "-o", CXX_CHECK]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rbrp.hpp check failed to build:\n{r.stdout}\n{r.stderr}")
    # the experiment driver (gpu/driver.cpp's flags and output) over include/rbrp.hpp
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(PKG, "driver", "rbrp_driver.cpp"), "-L", PKG, "-lrbrp",
           "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rbrp_driver failed to build:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
