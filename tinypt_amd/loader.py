"""Scene files: load_scene (mitsuba.rs:137-223) and load_obj_triangles (obj_loader.rs:101-109).

Both run in the native library (``tpt_scene_load_xml``, ``tpt_obj_load``); this module turns the
loaded descriptor back into the Python ``Scene`` so it can be inspected, edited and rendered:

    cfg = tinypt_amd.RenderConfig()
    scene = tinypt_amd.loader.load_scene("cornell.xml", cfg)   # cfg gets film/sampler/integrator
    out = tinypt_amd.render(scene, cfg)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi, load_library, TptError
from .scene import Camera, EnvMap, Instance, Material, Mesh, RenderConfig, Scene, Sphere, Transform, World


def _check(lib, rc):
    if rc != 0:
        raise TptError(rc, lib.tpt_last_error().decode(errors="replace"))


def load_obj_triangles(path: str) -> np.ndarray:
    """Triangles (n, 3, 3) of an OBJ file (v / f records, polygons fan-triangulated)."""
    lib = load_library()
    n = C.c_int64()
    _check(lib, lib.tpt_obj_load(os.fsencode(path), C.byref(n), None))
    out = np.zeros((max(n.value, 1), 3, 3))
    _check(lib, lib.tpt_obj_load(os.fsencode(path), C.byref(n), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out[:n.value]


def load_scene(path: str, config: RenderConfig, warnings: list | None = None) -> Scene:
    """Load a Mitsuba-XML scene.  ``config`` is updated in place from <film> (width, height),
    <sampler> (spp) and <integrator> (max_bounces, rr_start); other fields keep their values.
    Loader warnings (also printed to stderr) are appended to ``warnings`` when given."""
    lib = load_library()
    cfg = config.to_c()
    h = C.c_void_p()
    _check(lib, lib.tpt_scene_load_xml(os.fsencode(path), C.byref(cfg), C.byref(h)))
    try:
        dp = C.POINTER(_abi.SceneDesc)()
        _check(lib, lib.tpt_loaded_scene_desc(h, C.byref(dp)))
        d = dp.contents
        if warnings is not None:
            w = lib.tpt_loaded_scene_warnings(h).decode(errors="replace")
            warnings.extend(x for x in w.split("\n") if x)
        scene = _scene_from_desc(d)
    finally:
        lib.tpt_loaded_scene_destroy(h)
    config.width, config.height, config.spp = cfg.width, cfg.height, cfg.spp
    config.max_bounces, config.rr_start = cfg.max_bounces, cfg.rr_start
    return scene


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _scene_from_desc(d) -> Scene:
    cam = Camera(tuple(d.camera.eye), tuple(d.camera.target), tuple(d.camera.up), d.camera.vfov_deg, d.camera.aspect,
                 d.camera.focus_dist, d.camera.aperture)
    kinds = _arr(d.mat_kind, d.n_materials, np.int32)
    params = _arr(d.mat_param, 4 * d.n_materials, np.float64).reshape(-1, 4)
    mats = [Material(int(k), tuple(float(v) for v in p)) for k, p in zip(kinds, params)]
    world = World()
    c = _arr(d.sphere_center, 3 * d.n_spheres, np.float64).reshape(-1, 3)
    r = _arr(d.sphere_radius, d.n_spheres, np.float64)
    sm = _arr(d.sphere_mat, d.n_spheres, np.int32)
    for i in range(d.n_spheres):
        world.spheres.append(Sphere(tuple(c[i]), float(r[i]), int(sm[i])))
    begin = _arr(d.mesh_tri_begin, d.n_meshes + 1, np.int64)
    ntri = int(begin[-1]) if d.n_meshes else 0
    tv = _arr(d.tri_vertices, 9 * ntri, np.float64).reshape(-1, 3, 3)
    tm = _arr(d.tri_mat, ntri, np.int32)
    for m in range(d.n_meshes):
        world.meshes.append(Mesh(tv[begin[m]:begin[m + 1]], tm[begin[m]:begin[m + 1]]))
    im = _arr(d.inst_mesh, d.n_instances, np.int32)
    ix = _arr(d.inst_xform, 12 * d.n_instances, np.float64).reshape(-1, 12)
    io = _arr(d.inst_mat_override, d.n_instances, np.int32)
    for i in range(d.n_instances):
        a = ix[i, :9].reshape(3, 3)
        world.instances.append(Instance(int(im[i]), Transform.from_affine(a, tuple(ix[i, 9:])),
                                        None if io[i] < 0 else int(io[i])))
    env = None
    if d.env_kind == _abi.ENV_IMAGE:
        rgb = _arr(d.env_rgb, 3 * d.env_width * d.env_height, np.float64)
        env = EnvMap.from_pixels(d.env_width, d.env_height, rgb.reshape(d.env_height, d.env_width, 3))
    return Scene(cam, world, mats, env)
