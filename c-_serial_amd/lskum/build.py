"""Builds liblskum.so (HIP kernels + C ABI + host set-up) for gfx950.

Run as ``python -m lskum.build`` from ``c-_serial_amd/`` or call ``build()``.
Host C++ (cloud generator, set-up) is compiled by g++ with OpenMP and
``-ffp-contract=off`` so that sub-stencil classification is bit-identical to
the reference's; HIP sources by hipcc, FP64 with FMA contraction.
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, "csrc")
REPO = os.path.dirname(ROOT)
INCLUDE = os.path.join(REPO, "include")
LIB_PATH = os.path.join(PKG_DIR, "liblskum.so")
FPI_PATH = os.path.join(PKG_DIR, "lskum_fpi")
BUILD_DIR = os.path.join(ROOT, "_build")

HIP_SOURCES = ["kernels.hip", "kernels_mixed.hip", "adjoint.hip", "trajectory.hip", "api.hip",
               "shape_api.hip", "diagnostics.hip"]
CXX_SOURCES = ["setup.cpp", "cloud.cpp", "restart.cpp", "output.cpp", "shape.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LSKUM_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # device code may fuse a*b+c into FMAs (rounding-level differences from
    # the reference, inside the parity bar) except where a pragma forbids
    # it; nothing else is relaxed
    "-ffp-contract=fast-honor-pragmas", "-fno-fast-math",
    # packed FP32 forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-fopenmp", "-Wall", "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs, "-lgomp",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB_PATH)
    # the reference's main() as a program over the C ABI (per-stage loop of
    # include/lskum_fpi.hpp), linked against the library beside it
    tmp = FPI_PATH + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, os.path.join(CSRC, "fpi_main.cpp"),
          "-o", tmp, "-L" + PKG_DIR, "-llskum", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FPI_PATH)
    if verbose:
        print("built", LIB_PATH, "and", FPI_PATH)
    return LIB_PATH


def build_oracle(with_reference=True):
    """Builds the CPU oracle (and the reference harness when its sources are
    present).  Test infrastructure: the product library never uses them."""
    oracle_dir = os.path.join(REPO, "oracle")
    _run(["make", "-B", "-C", oracle_dir])
    ref = os.environ.get("LSKUM_REFERENCE", "/root/reference")
    if with_reference and os.path.isfile(os.path.join(ref, "fpi_solver_mod.h")):
        _run(["make", "-B", "-C", oracle_dir, "ref", "REF=" + ref])
        if os.path.isfile(LIB_PATH):
            _run(["make", "-B", "-C", oracle_dir, "dropin", "REF=" + ref])


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
