"""The C ABI: the built library loads, exports every entry point the header declares, and the
ctypes structs match the C structs field for field (checked against a gcc-compiled probe of
include/tinypt_gpu.h).  No device is used."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tinypt_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = T.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in the header but not exported"
        assert n in _abi.ENTRY_POINTS, f"{n} has no ctypes signature"
    assert set(_abi.ENTRY_POINTS) == set(names)
    assert lib.tpt_abi_version() == _abi.ABI_VERSION == 3


def test_struct_layouts_match_header(tmp_path):
    structs = {
        "tpt_camera_desc": (_abi.CameraDesc, ["eye", "vfov_deg", "aperture"]),
        "tpt_scene_desc": (_abi.SceneDesc, ["camera", "n_materials", "mat_param", "n_spheres", "mesh_tri_begin",
                                            "tri_mat", "inst_mat_override", "env_kind", "env_rgb"]),
        "tpt_render_config": (_abi.RenderConfigC, ["spp", "precision", "adaptive_threshold", "seed", "samples_per_pass",
                                                   "traversal_refill"]),
        "tpt_stats": (_abi.Stats, ["paths", "elapsed_ms", "trace_rays", "light_self_occluded"]),
        "tpt_scene_info": (_abi.SceneInfo, ["n_tris", "max_tlas_depth", "device_bytes"]),
        "tpt_pass_plan": (_abi.PassPlan, ["paths_alloc", "samples_per_pass", "pixels_per_pass", "path_sets"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for s, (_, fields) in structs.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f in fields:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l.strip())
    for s, (cls, fields) in structs.items():
        assert int(got[f"{s} size"]) == C.sizeof(cls), s
        for f in fields:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def test_null_and_invalid_arguments_are_rejected_without_a_device():
    lib = T.load_library()
    cfg = T.RenderConfig(width=4, height=4, spp=1).to_c()
    acc = np.zeros(48)
    rc = lib.tpt_render(None, None, C.byref(cfg), acc.ctypes.data_as(C.POINTER(C.c_double)),
                        acc.ctypes.data_as(C.POINTER(C.c_double)), None, None)
    assert rc == _abi.TPT_ERR_INVALID_ARGUMENT
    assert b"null" in lib.tpt_last_error()
    h = C.c_void_p()
    assert lib.tpt_scene_upload(None, None, C.byref(h)) == _abi.TPT_ERR_INVALID_ARGUMENT
    assert lib.tpt_ctx_create(0, None) == _abi.TPT_ERR_INVALID_ARGUMENT


def test_bad_scene_descriptor_is_reported():
    sc = scenes.default_scene(8, 8)
    sc.world.spheres[0].mat_id = 99  # out of range
    with pytest.raises(T.TptError, match="material out of range"):
        T.HostScene(sc)


def test_host_only_scene_cannot_be_rendered():
    sc = scenes.cornell_box(8, 8)
    hs = T.HostScene(sc)
    assert hs.info()["n_tris"] == 36
    lib = T.load_library()
    cfg = T.RenderConfig(width=8, height=8, spp=1).to_c()
    acc = np.zeros(8 * 8 * 3)
    w = np.zeros(64)
    rc = lib.tpt_render(None, hs.h, C.byref(cfg), acc.ctypes.data_as(C.POINTER(C.c_double)),
                        w.ctypes.data_as(C.POINTER(C.c_double)), None, None)
    assert rc == _abi.TPT_ERR_INVALID_ARGUMENT


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is present")
def test_context_creation_fails_cleanly_without_a_device():
    with pytest.raises(T.TptError):
        T.Context(0)


def test_c_client_compiles_and_links(tmp_path):
    """The plain C client of the header (tests/native/c_abi_render.c, which the GPU tests run)
    builds as C with warnings as errors and links against the library: the header is valid C and
    every call it makes resolves.  It is not run here (it needs a device)."""
    import shutil
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    libdir = os.path.join(ROOT, "tinypt_amd", "lib")
    exe = str(tmp_path / "c_abi_render")
    subprocess.run([gcc, "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "c_abi_render.c"), f"-L{libdir}", "-ltinypt_gpu",
                    f"-Wl,-rpath,{libdir}", "-Wl,--no-undefined"], check=True)
    assert os.path.getsize(exe) > 0
