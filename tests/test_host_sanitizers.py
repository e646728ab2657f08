"""The host code behind the C ABI -- the Mitsuba-XML, OBJ and Radiance HDR readers, the host
scene build (BVHs, four-wide trees, light table) and the image and checkpoint writers -- run over
mutated inputs under AddressSanitizer and UndefinedBehaviorSanitizer (tests/native/host_fuzz.cpp,
built here with g++ from the product's own sources).  A memory error, leak or undefined operation
makes the run fail with the sanitizer's report.  CPU only."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "tinypt_amd", "csrc")


def test_host_code_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path / "host_fuzz")
    build = subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                            "-fno-omit-frame-pointer", "-I", CSRC, "-o", exe, os.path.join(HERE, "native", "host_fuzz.cpp"),
                            os.path.join(CSRC, "scene_host.cpp"), os.path.join(CSRC, "scene_xml.cpp"),
                            os.path.join(CSRC, "output.cpp")], capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-4000:]
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, os.path.join(HERE, "golden", "scenes"), str(scratch), "600"], capture_output=True,
                         text=True, errors="replace", timeout=600, env=env)
    out = run.stdout + run.stderr[-8000:]
    assert run.returncode == 0, out
    assert "writers: done" in run.stdout, out
    for word in ("AddressSanitizer", "LeakSanitizer", "runtime error"):
        assert word not in run.stderr, out


def test_traversal_under_sanitizers(tmp_path):
    """The FP32 traversal and the FP64 self-occlusion code -- the kernels' own source, compiled for
    the host in tests/native/host_trace.hip -- with AddressSanitizer and UndefinedBehaviorSanitizer
    on the host side, running the host traversal and self-occlusion tests: an out-of-bounds node,
    triangle or stack access there is one the GPU kernels would make."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    clangxx = "/opt/rocm/llvm/bin/clang++"
    runtimes = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not (os.path.exists(hipcc) and os.path.exists(clangxx) and runtimes):
        pytest.skip("no ROCm clang with its sanitizer runtime")
    native = os.path.join(HERE, "native")
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    scene_obj, hook_obj, lib = (str(tmp_path / n) for n in ("scene_host.o", "host_trace.o", "libhost_trace.so"))
    host_only = [a for f in san for a in ("-Xarch_host", f)]  # device code is not instrumented
    cmds = [
        [clangxx, "-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fno-omit-frame-pointer"]
        + san + ["-I", CSRC, "-c", os.path.join(CSRC, "scene_host.cpp"), "-o", scene_obj],
        [hipcc, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-Wno-unused-function", "-Xclang",
         "-target-feature", "-Xclang", "-packed-fp32-ops"] + host_only
        + ["-I", CSRC, "-c", os.path.join(native, "host_trace.hip"), "-o", hook_obj],
        [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + host_only + [hook_obj, scene_obj, "-o", lib],
    ]
    for c in cmds:
        r = subprocess.run(c, capture_output=True, text=True, errors="replace", timeout=900)
        assert r.returncode == 0, " ".join(c) + "\n" + r.stderr[-4000:]
    env = dict(os.environ, LD_PRELOAD=runtimes[-1], TPT_HOST_TRACE_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                          os.path.join(HERE, "test_host_traversal.py"), os.path.join(HERE, "test_host_selfocc.py")],
                         capture_output=True, text=True, errors="replace", timeout=1200, env=env, cwd=ROOT)
    out = run.stdout[-6000:] + run.stderr[-6000:]
    assert run.returncode == 0, out
    for word in ("AddressSanitizer", "runtime error"):
        assert word not in run.stdout + run.stderr, out
