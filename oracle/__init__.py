"""ctypes wrapper of the CPU FP64 oracle (oracle/tpt_oracle.cpp).

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker / CPU baseline.  The product package (tinypt_amd) never imports it.

Every function here forwards to the C restatement of the reference renderer; see the header of
tpt_oracle.cpp for the files it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

from tinypt_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("aabb", "tri", "sph", "inst", "rays_ext", "rays_shadow", "paths", "ext_bounces",
                                                  "aabb_ext", "tri_ext", "sph_ext", "inst_ext", "light_samples",
                                                  "light_self")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class RenderOpts(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_bounces", C.c_int32),
        ("rr_start", C.c_int32), ("tile", C.c_int32), ("adaptive", C.c_int32), ("adaptive_min_spp", C.c_int32),
        ("adaptive_threshold", C.c_double), ("seed", C.c_uint64), ("morton", C.c_int32), ("threads", C.c_int32),
        ("per_sample_seed", C.c_int32), ("crop_x0", C.c_int32), ("crop_y0", C.c_int32), ("crop_x1", C.c_int32),
        ("crop_y1", C.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "liboracle.so"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_rng_u32": (None, [C.c_uint64, C.c_int64, C.POINTER(C.c_uint32)]),
            "orc_rng_f64": (None, [C.c_uint64, C.c_int64, _dp]),
            "orc_seed_for": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_morton2": (C.c_uint64, [C.c_uint64] * 4),
            "orc_srgb_to_linear": (C.c_double, [C.c_double]),
            "orc_linear_to_srgb": (C.c_double, [C.c_double]),
            "orc_sphere_hit": (C.c_int, [_dp, C.c_double, _dp, C.c_double, C.c_double, _dp]),
            "orc_triangle_hit": (C.c_int, [_dp, _dp, C.c_double, C.c_double, _dp]),
            "orc_aabb_hit": (C.c_int, [_dp, _dp, C.c_double, C.c_double, _dp]),
            "orc_camera_rays": (None, [C.POINTER(_abi.CameraDesc), C.c_uint64, C.c_int64, _dp, _dp]),
            "orc_transform_apply": (None, [_dp, C.c_int32, C.c_int64, _dp, _dp]),
            "orc_task_tiles": (C.c_int64, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
            "orc_checkpoint_save": (C.c_int32, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, _dp, _dp]),
            "orc_checkpoint_load": (C.c_int32, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint32, _u64p, _dp, _dp]),
            "orc_prim_distance": (None, [C.c_void_p, C.c_int64, _dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), _dp, _dp]),
            "orc_light_shadow": (None, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), _dp, _dp, _dp,
                                        C.POINTER(C.c_int32)]),
            "orc_material_sample": (None, [C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_int64, _dp]),
            "orc_material_eval": (None, [C.c_int, _dp, _dp, _dp, _dp, _dp]),
            "orc_ggx_vndf": (None, [_dp, C.c_double, C.c_double, C.c_double, _dp]),
            "orc_scene_create": (C.c_void_p, [C.POINTER(_abi.SceneDesc)]),
            "orc_scene_destroy": (None, [C.c_void_p]),
            "orc_scene_n_lights": (C.c_int64, [C.c_void_p]),
            "orc_scene_light_total": (C.c_double, [C.c_void_p]),
            "orc_scene_lights": (None, [C.c_void_p, _dp]),
            "orc_scene_sample_light": (None, [C.c_void_p, _dp, C.c_double, C.c_uint64, C.c_int64, _dp]),
            "orc_scene_bvh_nodes": (C.c_int64, [C.c_void_p, C.c_int32]),
            "orc_scene_bvh": (None, [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]),
            "orc_scene_trace": (None, [C.c_void_p, C.c_int64, _dp, _dp, C.POINTER(Counters), C.c_int32]),
            "orc_env_total": (C.c_double, [C.c_void_p]),
            "orc_env_lookup": (None, [C.c_void_p, C.c_int64, _dp, _dp, _dp]),
            "orc_env_sample": (None, [C.c_void_p, C.c_uint64, C.c_int64, _dp]),
            "orc_scene_radiance": (None, [C.c_void_p, C.c_int64, _dp, _u64p, C.c_int32, C.c_int32, _dp]),
            "orc_render": (C.c_double, [C.c_void_p, C.POINTER(RenderOpts), _dp, _dp, _dp, C.POINTER(Counters)]),
            "orc_resolve_pixels": (None, [C.c_int64, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------- rng.rs
def rng_u32(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_rng_u32(seed & 0xFFFFFFFFFFFFFFFF, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def rng_f64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().orc_rng_f64(seed & 0xFFFFFFFFFFFFFFFF, n, _d(out))
    return out


def seed_for(px: int, py: int, s: int) -> int:
    return int(lib().orc_seed_for(px, py, s))


def morton2(tx, ty, tiles_x, tiles_y) -> int:
    return int(lib().orc_morton2(tx, ty, tiles_x, tiles_y))


# ---------------------------------------------------------------- geometry.rs
def sphere_hit(c, r, ray, tmin, tmax):
    out = np.zeros(7)
    ok = lib().orc_sphere_hit(_d(_f64(c)), float(r), _d(_f64(ray)), tmin, tmax, _d(out))
    return out if ok else None


def triangle_hit(v, ray, tmin, tmax):
    out = np.zeros(7)
    ok = lib().orc_triangle_hit(_d(_f64(v).reshape(-1)), _d(_f64(ray)), tmin, tmax, _d(out))
    return out if ok else None


def aabb_hit(box, ray, tmin, tmax):
    t0 = np.zeros(1)
    ok = lib().orc_aabb_hit(_d(_f64(box)), _d(_f64(ray)), tmin, tmax, _d(t0))
    return float(t0[0]) if ok else None


def camera_rays(cam_desc: _abi.CameraDesc, seed: int, sxy) -> np.ndarray:
    sxy = _f64(sxy).reshape(-1, 2)
    out = np.zeros((len(sxy), 7))
    lib().orc_camera_rays(C.byref(cam_desc), seed, len(sxy), _d(sxy), _d(out))
    return out


# ---------------------------------------------------------------- transform.rs
TRANSFORM_OPS = {"apply_point": 0, "apply_point_inv": 1, "apply_vec_inv": 2, "apply_normal": 3}


def transform_apply(xf12, op: str, vecs) -> np.ndarray:
    """The oracle's Transform (transform.rs:24-93) built from xf12 = A (row-major) + t, applied
    to each row of vecs (n x 3) by the named method."""
    v = _f64(np.asarray(vecs, dtype=np.float64).reshape(-1, 3))
    out = np.zeros_like(v)
    lib().orc_transform_apply(_d(_f64(xf12)), TRANSFORM_OPS[op], len(v), _d(v), _d(out))
    return out


# ---------------------------------------------------------------- material.rs
def material_sample(kind, param, ray, hit, seed, n) -> np.ndarray:
    """n draws of Material::sample from one RNG stream; rows: valid, d(3), weight(3), pdf,
    is_delta, origin(3)."""
    out = np.zeros((n, 12))
    lib().orc_material_sample(kind, _d(_f64(param)), _d(_f64(ray)), _d(_f64(hit)), seed, n, _d(out))
    return out


def material_eval(kind, param, wo, wi, n):
    out = np.zeros(4)
    lib().orc_material_eval(kind, _d(_f64(param)), _d(_f64(wo)), _d(_f64(wi)), _d(_f64(n)), _d(out))
    return out[:3].copy(), float(out[3])


def ggx_vndf(wo, alpha, u1, u2):
    out = np.zeros(3)
    lib().orc_ggx_vndf(_d(_f64(wo)), alpha, u1, u2, _d(out))
    return out


# ---------------------------------------------------------------- scenes
class OracleScene:
    """A reference-faithful FP64 scene built from a tinypt_amd.scene.Scene."""

    def __init__(self, scene):
        desc, keep = scene.to_desc()
        self._keep = (desc, keep)
        self.h = lib().orc_scene_create(C.byref(desc))
        self.scene = scene

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def lights(self) -> np.ndarray:
        n = lib().orc_scene_n_lights(self.h)
        out = np.zeros((n, 10))
        if n:
            lib().orc_scene_lights(self.h, _d(out))
        return out

    def light_total(self) -> float:
        return float(lib().orc_scene_light_total(self.h))

    def sample_light(self, p, time, seed, n) -> np.ndarray:
        out = np.zeros((n, 11))
        lib().orc_scene_sample_light(self.h, _d(_f64(p)), time, seed, n, _d(out))
        return out

    def bvh(self, mesh: int) -> dict:
        n = lib().orc_scene_bvh_nodes(self.h, mesh)
        ntri = len(self.scene.world.meshes[mesh].tris)
        bbox = np.zeros((n, 6))
        left = np.zeros(n, np.int32)
        right = np.zeros(n, np.int32)
        start = np.zeros(n, np.uint32)
        count = np.zeros(n, np.uint32)
        idx = np.zeros(max(ntri, 1), np.int64)
        lib().orc_scene_bvh(self.h, mesh, _d(bbox), left.ctypes.data_as(C.POINTER(C.c_int32)),
                            right.ctypes.data_as(C.POINTER(C.c_int32)), start.ctypes.data_as(C.POINTER(C.c_uint32)),
                            count.ctypes.data_as(C.POINTER(C.c_uint32)), idx.ctypes.data_as(C.POINTER(C.c_int64)))
        return dict(bbox=bbox, left=left, right=right, start=start, count=count, indices=idx[:ntri])

    def trace(self, rays, threads: int = 0):
        """rays (n, 8): o, d, tmin, tmax.  Returns dict(t, kind, inst, prim, mat, n, p, counters)."""
        rays = _f64(rays).reshape(-1, 8)
        out = np.zeros((len(rays), 11))
        cnt = Counters()
        lib().orc_scene_trace(self.h, len(rays), _d(rays), _d(out), C.byref(cnt), threads)
        return dict(t=out[:, 0], kind=out[:, 1].astype(np.int32), inst=out[:, 2].astype(np.int32),
                    prim=out[:, 3].astype(np.int32), mat=out[:, 4].astype(np.int32), n=out[:, 5:8], p=out[:, 8:11],
                    counters=cnt.as_dict())

    def prim_distance(self, rays, kind, inst, prim):
        """FP64 distance along each ray to one given primitive (kind 1: triangle `prim` of
        instance `inst`; kind 2: sphere `inst`), +inf where it is not hit in [tmin, tmax], and
        the smallest barycentric coordinate of the crossing (1 for spheres)."""
        rays = _f64(rays).reshape(-1, 8)
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))
        kind, inst, prim = i32(kind), i32(inst), i32(prim)
        out, margin = np.zeros(len(rays)), np.zeros(len(rays))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        lib().orc_prim_distance(self.h, len(rays), _d(rays), ip(kind), ip(inst), ip(prim), _d(out), _d(margin))
        return out, margin

    def env_total(self) -> float:
        return float(lib().orc_env_total(self.h))

    def env_lookup(self, dirs):
        dirs = _f64(dirs).reshape(-1, 3)
        rgb = np.zeros((len(dirs), 3))
        pdf = np.zeros(len(dirs))
        lib().orc_env_lookup(self.h, len(dirs), _d(dirs), _d(rgb), _d(pdf))
        return rgb, pdf

    def env_sample(self, seed, n):
        out = np.zeros((n, 7))
        lib().orc_env_sample(self.h, seed, n, _d(out))
        return out[:, :3], out[:, 3:6], out[:, 6]

    def radiance(self, rays, seeds, max_bounces=8, rr_start=3):
        rays = _f64(rays).reshape(-1, 7)
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        out = np.zeros((len(rays), 3))
        lib().orc_scene_radiance(self.h, len(rays), _d(rays), seeds.ctypes.data_as(_u64p), max_bounces, rr_start, _d(out))
        return out

    def render(self, width, height, spp, max_bounces=8, rr_start=3, seed=0, adaptive=False, adaptive_min_spp=8,
               adaptive_threshold=0.02, threads=0, per_sample_seed=False, crop=None, tile=16, morton=True):
        """render() (render.rs:42-278).  Returns dict(acc (h,w,3), acc_w (h,w), acc_sq (h,w,3),
        seconds, counters)."""
        o = RenderOpts()
        o.width, o.height, o.spp, o.max_bounces, o.rr_start = width, height, spp, max_bounces, rr_start
        o.tile = tile
        o.adaptive = 1 if adaptive else 0
        o.adaptive_min_spp = adaptive_min_spp
        o.adaptive_threshold = adaptive_threshold
        o.seed = seed & 0xFFFFFFFFFFFFFFFF
        o.morton = 1 if morton else 0
        o.threads = threads
        o.per_sample_seed = 1 if per_sample_seed else 0
        if crop is not None:
            o.crop_x0, o.crop_y0, o.crop_x1, o.crop_y1 = crop
        acc = np.zeros((height, width, 3))
        acc_w = np.zeros((height, width))
        acc_sq = np.zeros((height, width, 3))
        cnt = Counters()
        secs = lib().orc_render(self.h, C.byref(o), _d(acc), _d(acc_w), _d(acc_sq), C.byref(cnt))
        return dict(acc=acc, acc_w=acc_w, acc_sq=acc_sq, seconds=float(secs), counters=cnt.as_dict())


def task_tiles(w, h, tile=16, morton=True) -> np.ndarray:
    """The tile tasks of a w x h image in task-id order (render.rs:76-106): (n, 4) x0, y0, x1, y1."""
    n = int(lib().orc_task_tiles(w, h, tile, 1 if morton else 0, None))
    out = np.zeros((n, 4), dtype=np.int64)
    lib().orc_task_tiles(w, h, tile, 1 if morton else 0, out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def checkpoint_save(path, scene_hash, next_id, acc, acc_w):
    """save_checkpoint (checkpoint.rs:48-80) of acc (h, w, 3) and acc_w (h, w)."""
    h, w = acc_w.shape
    rc = lib().orc_checkpoint_save(str(path).encode(), scene_hash & 0xFFFFFFFFFFFFFFFF, w, h, next_id,
                                   _d(_f64(acc)), _d(_f64(acc_w)))
    if rc != 0:
        raise OSError(f"checkpoint write failed: {path}")


def checkpoint_load(path, scene_hash, w, h):
    """load_checkpoint (checkpoint.rs:84-138): (next_id, acc, acc_w), or None."""
    acc, acc_w = np.zeros((h, w, 3)), np.zeros((h, w))
    nid = C.c_uint64()
    rc = lib().orc_checkpoint_load(str(path).encode(), scene_hash & 0xFFFFFFFFFFFFFFFF, w, h, C.byref(nid), _d(acc),
                                   _d(acc_w))
    if rc < 0:
        raise OSError(f"checkpoint read failed: {path}")
    return (int(nid.value), acc, acc_w) if rc == 1 else None


def resolve_pixels(acc, acc_w) -> np.ndarray:
    """output.rs:96-102: acc / max(acc_w, 1)."""
    acc = _f64(acc)
    acc_w = _f64(acc_w)
    out = np.zeros_like(acc)
    lib().orc_resolve_pixels(acc_w.size, _d(acc), _d(acc_w), _d(out))
    return out
