"""The C ABI from plain C: tests/native/c_abi_render.c (built here with gcc against
include/tinypt_gpu.h and lib/libtinypt_gpu.so) loads a sample scene file, uploads it and renders
it with both precisions, as a non-Python caller of the library would.  Its sums equal, bit for
bit, those of the same calls made through the Python package."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import loader

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SCENES = os.path.join(HERE, "golden", "scenes")


@pytest.fixture(scope="module")
def c_client(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    libdir = os.path.join(ROOT, "tinypt_amd", "lib")
    exe = str(tmp_path_factory.mktemp("cabi") / "c_abi_render")
    subprocess.run([gcc, "-O2", "-Wall", "-Wextra", "-Werror", "-o", exe, os.path.join(HERE, "native", "c_abi_render.c"),
                    f"-L{libdir}", "-ltinypt_gpu", f"-Wl,-rpath,{libdir}"], check=True)
    return exe


@pytest.mark.parametrize("name,w,h,spp", [("cornell.xml", 48, 48, 64), ("mesh.xml", 64, 36, 32)])
def test_c_client_renders_as_the_python_package(gpu_ctx, c_client, tmp_path, name, w, h, spp):
    path = os.path.join(SCENES, name)
    prefix = str(tmp_path / "out")
    r = subprocess.run([c_client, path, str(w), str(h), str(spp), prefix], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert os.path.getsize(prefix + "_fp32.ppm") > 0
    cfg = T.RenderConfig(width=w, height=h, spp=spp, max_bounces=8, rr_start=3, seed=7)
    scene = loader.load_scene(path, cfg)
    cfg.width, cfg.height, cfg.spp = w, h, spp
    ds = T.DeviceScene(scene, gpu_ctx)
    for prec, tag in ((T.FP32_WAVEFRONT, "fp32"), (T.FP64_PLAIN, "fp64")):
        cfg.precision = prec
        out = T.render(ds, cfg, want_sq=False)
        raw = np.fromfile(f"{prefix}_{tag}.bin", dtype=np.float64)
        acc, acc_w = raw[:w * h * 3].reshape(h, w, 3), raw[w * h * 3:].reshape(h, w)
        assert np.array_equal(acc_w, out.acc_w)
        assert np.array_equal(acc, out.acc), (tag, np.abs(acc - out.acc).max())
        assert acc.sum() > 0
