"""Builds the CPU tests' host build of the FP32 traversal (libhost_trace.so, beside this file):
tpt_trace32.h compiled for the host over the product's host scene build (scene_host.cpp), so
tests/test_host_traversal.py can hold the traversal against the oracle without a GPU.  Test
infrastructure: the product library does not contain it.

    python tests/native/build_host_trace.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(os.path.dirname(HERE)), "tinypt_amd", "csrc")
# TPT_HOST_TRACE_LIB names a library built elsewhere (the sanitizer test's instrumented build),
# used as it is.
LIB_OVERRIDE = os.environ.get("TPT_HOST_TRACE_LIB")
LIB = LIB_OVERRIDE or os.path.join(HERE, "libhost_trace.so")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("TPT_OFFLOAD_ARCH", "gfx950")


def build(verbose: bool = False) -> str:
    """Builds the library; concurrent callers (pytest workers) take turns on a lock file and the
    later ones find it current."""
    import fcntl
    with open(os.path.join(HERE, ".build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        if os.path.exists(LIB) and not is_stale():
            return LIB
        tag = f"{os.getpid()}"
        tmp = f"{LIB}.tmp{tag}"
        scene_obj = os.path.join(HERE, f"scene_host.{tag}.o")
        hook_obj = os.path.join(HERE, f"host_trace.{tag}.o")
        cmds = [
            ["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC, "-c",
             os.path.join(CSRC, "scene_host.cpp"), "-o", scene_obj],
            [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-function", "-Xclang",
             "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC, "-c", os.path.join(HERE, "host_trace.hip"),
             "-o", hook_obj],
            [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", hook_obj, scene_obj, "-o", tmp],
        ]
        try:
            for c in cmds:
                if verbose:
                    print(" ".join(c), flush=True)
                r = subprocess.run(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
                if r.returncode != 0:
                    raise RuntimeError("host_trace build failed:\n" + " ".join(c) + "\n" + r.stdout)
            os.replace(tmp, LIB)
        finally:
            for f in (scene_obj, hook_obj, tmp):
                if os.path.exists(f):
                    os.remove(f)
    return LIB


def is_stale() -> bool:
    """True when the library is missing or older than a source it is built from."""
    if LIB_OVERRIDE:
        return False
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    srcs = [os.path.join(HERE, "host_trace.hip")] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)
                                                      if f.endswith((".h", ".cpp"))]
    return any(os.path.getmtime(s) > t for s in srcs)


if __name__ == "__main__":
    if os.path.exists(LIB):
        os.utime(os.path.join(HERE, "host_trace.hip"))  # an explicit build always compiles
    print(build(verbose="-v" in sys.argv))
