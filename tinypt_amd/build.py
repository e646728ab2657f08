"""Build libtinypt_gpu.so (HIP for gfx950 + the C++ host scene build) in-tree.

Each translation unit is compiled separately so that the plain FP64 kernel and the host scene
build get -ffp-contract=off (their f64 arithmetic restates the reference's, which never fuses
into FMA), while the FP32 kernels contract within expressions.  Packed FP32 instructions are
disabled for all device code.

    python -m tinypt_amd.build [-v]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libtinypt_gpu.so")
OBJDIR = os.path.join(HERE, "lib", "obj")

ARCH = os.environ.get("TPT_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
GXX = shutil.which("g++") or "g++"

HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC]
# FP32 kernels: contraction only inside one source expression (-ffp-contract=on), so that the
# forward kernels and the adjoint's replay of the same code round identically (a path replayed
# for the gradient is then the path the image summed).  FP32 division and square root use the
# hardware reciprocal / square-root approximations instead of correctly rounded sequences
# (a few ulp; the FP64 kernels are unaffected).  FP32 subnormals flush to zero, so that those
# approximations are the bare instructions, without the exponent scaling that keeps subnormal
# operands exact; FP64 keeps its subnormals.
FP32_FLAGS = ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]
SOURCES = [
    ("api.hip", HIP_FLAGS),
    ("k_trace.hip", HIP_FLAGS + FP32_FLAGS),
    ("k_wavefront.hip", HIP_FLAGS + FP32_FLAGS),
    ("k_adjoint.hip", HIP_FLAGS + FP32_FLAGS),
    ("k_plain64.hip", HIP_FLAGS + ["-ffp-contract=off"]),
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC]
HOST_SOURCES = [
    ("scene_host.cpp", HOST_FLAGS),
    ("scene_xml.cpp", HOST_FLAGS),
    ("output.cpp", HOST_FLAGS),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs_list = []
    for src, flags in SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        jobs_list.append([HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
    for src, flags in HOST_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        jobs_list.append([GXX] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs_list))
    objs = [c[-1] for c in jobs_list]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
