"""The host code behind the C ABI -- the Mitsuba-XML, OBJ and Radiance HDR readers, the host
scene build (BVHs, four-wide trees, light table) and the image and checkpoint writers -- run over
mutated inputs under AddressSanitizer and UndefinedBehaviorSanitizer (tests/native/host_fuzz.cpp,
built here with g++ from the product's own sources).  A memory error, leak or undefined operation
makes the run fail with the sanitizer's report.  CPU only."""
import os
import shutil
import subprocess

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
