"""Host-side scene description: a Python mirror of the reference crate's scene types.

Names and argument meanings follow the reference so that code written against it reads the
same here:

* ``Material`` variants (material.rs:30-43) -> ``Lambert``, ``Metal``, ``Dielectric``, ``Ggx``,
  ``Subsurface``, ``DiffuseLight``.
* ``Camera.look_at`` / ``Camera.look_at_dof`` (ray.rs:36-86).
* ``Transform`` (transform.rs:11-112), with the same f64 operation order so composed Mitsuba
  transforms come out bit-identical.
* ``Sphere``, ``Mesh``, ``Instance``, ``World`` (geometry.rs, world.rs:20-62).
* ``EnvMap.from_pixels`` / ``EnvMap.constant`` / ``EnvMap.scaled`` (env.rs:37-89).
* ``Scene`` (scene.rs:14-23) and ``RenderConfig`` (config.rs:5-91).

``Scene.to_desc()`` flattens everything into the C ABI's ``tpt_scene_desc``.  BVHs and light
CDFs are built by the library from that descriptor (bvh.rs:51-242, world.rs:127-167).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi

Vec3 = Tuple[float, float, float]


# ---------------------------------------------------------------- colour (math.rs:20-51,176-191)
def srgb_to_linear(c: float) -> float:
    """math.rs:176-182."""
    if c <= 0.04045:
        return c / 12.92
    return ((c + 0.055) / 1.055) ** 2.4


def linear_to_srgb(c: float) -> float:
    """math.rs:185-191."""
    if c <= 0.0031308:
        return 12.92 * c
    return 1.055 * c ** (1.0 / 2.4) - 0.055


def color_from_srgb(r: float, g: float, b: float) -> Vec3:
    """Color::from_srgb (math.rs:28-30)."""
    return (srgb_to_linear(r), srgb_to_linear(g), srgb_to_linear(b))


def luminance(c: Sequence[float]) -> float:
    """Color::luminance (math.rs:40): Rec.709 weights."""
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


# ---------------------------------------------------------------- materials (material.rs:30-43)
@dataclass(frozen=True)
class Material:
    kind: int
    param: Tuple[float, float, float, float]

    @property
    def is_delta(self) -> bool:
        """Material::is_delta (material.rs:70-72)."""
        return self.kind in (_abi.MAT_METAL, _abi.MAT_DIELECTRIC)

    def emitted(self) -> Optional[Vec3]:
        """Material::emitted (material.rs:62-67)."""
        if self.kind == _abi.MAT_DIFFUSE_LIGHT:
            return tuple(self.param[:3])
        return None


def Lambert(albedo: Vec3) -> Material:
    return Material(_abi.MAT_LAMBERT, (float(albedo[0]), float(albedo[1]), float(albedo[2]), 0.0))


def Metal(albedo: Vec3) -> Material:
    return Material(_abi.MAT_METAL, (float(albedo[0]), float(albedo[1]), float(albedo[2]), 0.0))


def Dielectric(ior: float, absorption: Vec3 = (0.0, 0.0, 0.0)) -> Material:
    return Material(_abi.MAT_DIELECTRIC, (float(absorption[0]), float(absorption[1]), float(absorption[2]), float(ior)))


def Ggx(albedo: Vec3, alpha: float) -> Material:
    return Material(_abi.MAT_GGX, (float(albedo[0]), float(albedo[1]), float(albedo[2]), float(alpha)))


def Subsurface(albedo: Vec3, scatter_dist: float) -> Material:
    return Material(_abi.MAT_SUBSURFACE, (float(albedo[0]), float(albedo[1]), float(albedo[2]), float(scatter_dist)))


def DiffuseLight(emit: Vec3) -> Material:
    return Material(_abi.MAT_DIFFUSE_LIGHT, (float(emit[0]), float(emit[1]), float(emit[2]), 0.0))


# ---------------------------------------------------------------- camera (ray.rs:22-86)
@dataclass
class Camera:
    eye: Vec3
    target: Vec3
    up: Vec3
    vfov_deg: float
    aspect: float
    focus_dist: float
    aperture: float = 0.0

    @staticmethod
    def look_at(eye: Vec3, target: Vec3, up: Vec3, fov_deg: float, aspect: float) -> "Camera":
        """Camera::look_at (ray.rs:36-56): pinhole, focus at |eye - target|."""
        d = [eye[i] - target[i] for i in range(3)]
        return Camera(tuple(eye), tuple(target), tuple(up), fov_deg, aspect, math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), 0.0)

    @staticmethod
    def look_at_dof(eye: Vec3, target: Vec3, up: Vec3, fov_deg: float, aspect: float, focus_dist: float,
                    aperture: float) -> "Camera":
        """Camera::look_at_dof (ray.rs:59-86)."""
        return Camera(tuple(eye), tuple(target), tuple(up), fov_deg, aspect, focus_dist, aperture)


# ---------------------------------------------------------------- transform (transform.rs, math.rs:100-170)
def _mat_mul(a, b):
    return [[a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] for j in range(3)] for i in range(3)]


def _mat_vec(m, v):
    return (m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
            m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2])


def _mat_invert(m):
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    inv_det = 1.0 / det
    return [[(m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv_det, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv_det,
             (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv_det],
            [(m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv_det, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv_det,
             (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv_det],
            [(m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv_det, (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv_det,
             (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv_det]]


def _norm(v):
    l = max(math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), 1e-30)
    return (v[0] / l, v[1] / l, v[2] / l)


def _rotation(axis: Vec3, deg: float):
    """Rodrigues rotation (transform.rs:101-112)."""
    x, y, z = _norm(axis)
    rad = deg * (math.pi / 180.0)  # f64::to_radians
    s, c = math.sin(rad), math.cos(rad)
    one_c = 1.0 - c
    return [[c + x * x * one_c, x * y * one_c - z * s, x * z * one_c + y * s],
            [y * x * one_c + z * s, c + y * y * one_c, y * z * one_c - x * s],
            [z * x * one_c - y * s, z * y * one_c + x * s, c + z * z * one_c]]


@dataclass
class Transform:
    """Object->world affine p' = A p + t (transform.rs:11-20)."""
    a: list
    t: Vec3

    @staticmethod
    def from_affine(a, t: Vec3) -> "Transform":
        return Transform([list(map(float, r)) for r in a], tuple(map(float, t)))

    @staticmethod
    def identity() -> "Transform":
        return Transform.from_affine([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], (0.0, 0.0, 0.0))

    @staticmethod
    def new(t: Vec3, rot_y_deg: float, s: float) -> "Transform":
        """transform.rs:36-38: p' = R_y(s p) + t."""
        return Transform.from_affine(_mat_mul(_rotation((0.0, 1.0, 0.0), rot_y_deg), [[s, 0, 0], [0, s, 0], [0, 0, s]]), t)

    @staticmethod
    def translate(t: Vec3) -> "Transform":
        return Transform.from_affine([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], t)

    @staticmethod
    def rotate(axis: Vec3, deg: float) -> "Transform":
        return Transform.from_affine(_rotation(axis, deg), (0.0, 0.0, 0.0))

    @staticmethod
    def scale(s: Vec3) -> "Transform":
        return Transform.from_affine([[s[0], 0.0, 0.0], [0.0, s[1], 0.0], [0.0, 0.0, s[2]]], (0.0, 0.0, 0.0))

    @staticmethod
    def from_matrix4(m) -> "Transform":
        return Transform.from_affine([[m[r][c] for c in range(3)] for r in range(3)], (m[0][3], m[1][3], m[2][3]))

    def compose(self, inner: "Transform") -> "Transform":
        """self . inner (transform.rs:70-72)."""
        at = _mat_vec(self.a, inner.t)
        return Transform.from_affine(_mat_mul(self.a, inner.a), (at[0] + self.t[0], at[1] + self.t[1], at[2] + self.t[2]))

    def apply_point(self, p: Vec3) -> Vec3:
        v = _mat_vec(self.a, p)
        return (v[0] + self.t[0], v[1] + self.t[1], v[2] + self.t[2])

    def apply_point_inv(self, p: Vec3) -> Vec3:
        return _mat_vec(_mat_invert(self.a), (p[0] - self.t[0], p[1] - self.t[1], p[2] - self.t[2]))

    def apply_normal(self, n: Vec3) -> Vec3:
        ai = _mat_invert(self.a)
        nt = [[ai[j][i] for j in range(3)] for i in range(3)]
        return _norm(_mat_vec(nt, n))

    def flat(self) -> List[float]:
        return [self.a[0][0], self.a[0][1], self.a[0][2], self.a[1][0], self.a[1][1], self.a[1][2],
                self.a[2][0], self.a[2][1], self.a[2][2], self.t[0], self.t[1], self.t[2]]


# ---------------------------------------------------------------- geometry / world
@dataclass
class Sphere:
    c: Vec3
    r: float
    mat_id: int


class Mesh:
    """Triangle mesh (world.rs:20-33).  ``tris``: float64 array (n, 3, 3) of v0, v1, v2.
    The library builds the mesh BVH (Mesh::new -> Bvh::build)."""

    def __init__(self, tris, mat_ids):
        self.tris = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3))
        m = np.asarray(mat_ids, dtype=np.int32)
        if m.ndim == 0:
            m = np.full(len(self.tris), int(m), dtype=np.int32)
        self.mat_ids = np.ascontiguousarray(m)
        if len(self.mat_ids) != len(self.tris):
            raise ValueError("one material id per triangle")


@dataclass
class Instance:
    mesh_id: int
    xform: Transform
    mat_override: Optional[int] = None


@dataclass
class World:
    spheres: List[Sphere] = field(default_factory=list)
    meshes: List[Mesh] = field(default_factory=list)
    instances: List[Instance] = field(default_factory=list)


class EnvMap:
    """Equirectangular environment (env.rs:20-33); ``data`` float64 (h, w, 3) linear RGB."""

    def __init__(self, width: int, height: int, data):
        self.width = int(width)
        self.height = int(height)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(self.height, self.width, 3))

    @staticmethod
    def from_pixels(w: int, h: int, data) -> "EnvMap":
        return EnvMap(w, h, data)

    @staticmethod
    def constant(color: Vec3) -> "EnvMap":
        """env.rs:47-49."""
        return EnvMap(1, 1, np.array(color, dtype=np.float64).reshape(1, 1, 3))

    def scaled(self, factor: float) -> "EnvMap":
        """env.rs:52-58."""
        if factor == 1.0:
            return self
        return EnvMap(self.width, self.height, self.data * factor)


@dataclass
class Scene:
    """scene.rs:14-23.  ``env=None`` selects the procedural sky (integrator.rs:25-37)."""
    cam: Camera
    world: World
    mats: List[Material]
    env: Optional[EnvMap] = None

    def n_tris(self) -> int:
        return int(sum(len(m.tris) for m in self.world.meshes))

    def to_desc(self):
        """Flatten into ``tpt_scene_desc``.  Returns (desc, keepalive); keep ``keepalive``
        referenced while the descriptor is in use."""
        keep = {}

        def arr(name, a, dtype):
            a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            ctype = {np.float64: C.c_double, np.int32: C.c_int32, np.int64: C.c_int64}[dtype]
            return a.ctypes.data_as(C.POINTER(ctype))

        d = _abi.SceneDesc()
        cam = self.cam
        d.camera.eye[:] = [float(x) for x in cam.eye]
        d.camera.target[:] = [float(x) for x in cam.target]
        d.camera.up[:] = [float(x) for x in cam.up]
        d.camera.vfov_deg = float(cam.vfov_deg)
        d.camera.aspect = float(cam.aspect)
        d.camera.focus_dist = float(cam.focus_dist)
        d.camera.aperture = float(cam.aperture)

        d.n_materials = len(self.mats)
        d.mat_kind = arr("mk", [m.kind for m in self.mats], np.int32)
        d.mat_param = arr("mp", [list(m.param) for m in self.mats], np.float64)

        w = self.world
        d.n_spheres = len(w.spheres)
        d.sphere_center = arr("sc", [list(s.c) for s in w.spheres], np.float64)
        d.sphere_radius = arr("sr", [s.r for s in w.spheres], np.float64)
        d.sphere_mat = arr("sm", [s.mat_id for s in w.spheres], np.int32)

        d.n_meshes = len(w.meshes)
        counts = [len(m.tris) for m in w.meshes]
        begin = np.zeros(len(counts) + 1, dtype=np.int64)
        begin[1:] = np.cumsum(counts)
        d.mesh_tri_begin = arr("mb", begin, np.int64)
        if w.meshes:
            d.tri_vertices = arr("tv", np.concatenate([m.tris.reshape(-1, 9) for m in w.meshes]), np.float64)
            d.tri_mat = arr("tm", np.concatenate([m.mat_ids for m in w.meshes]), np.int32)
        else:
            d.tri_vertices = arr("tv", [], np.float64)
            d.tri_mat = arr("tm", [], np.int32)

        d.n_instances = len(w.instances)
        d.inst_mesh = arr("im", [i.mesh_id for i in w.instances], np.int32)
        d.inst_xform = arr("ix", [i.xform.flat() for i in w.instances], np.float64)
        d.inst_mat_override = arr("io", [(-1 if i.mat_override is None else i.mat_override) for i in w.instances], np.int32)

        if self.env is None:
            d.env_kind = _abi.ENV_SKY
            d.env_width = 0
            d.env_height = 0
            d.env_rgb = arr("er", [], np.float64)
        else:
            d.env_kind = _abi.ENV_IMAGE
            d.env_width = self.env.width
            d.env_height = self.env.height
            d.env_rgb = arr("er", self.env.data.reshape(-1), np.float64)
        return d, keep


# ---------------------------------------------------------------- config (config.rs:5-91)
@dataclass
class RenderConfig:
    width: int = 1920
    height: int = 1080
    spp: int = 512
    max_bounces: int = 8          # constants.rs:59
    rr_start: int = 3             # constants.rs:61
    tile: int = 16
    adaptive_enabled: bool = False
    adaptive_min_spp: int = 8
    adaptive_threshold: float = 0.02
    morton_enabled: bool = True
    seed: int = 0
    precision: int = _abi.FP32_WAVEFRONT
    samples_per_pass: int = 0
    traversal_refill: int = 0     # 0: by scene shape, -1: off, 1..63: refill threshold (lanes)
    # host-side settings (config.rs:17-45, 79-86): checkpoints and image output
    checkpoint_enabled: bool = False
    checkpoint_every_tasks: int = 128
    scene_hash: int = 0x4859503000000001
    output_path: str = "out.ppm"
    tonemap: str = "aces"
    exposure: float = 0.0

    def to_c(self) -> _abi.RenderConfigC:
        c = _abi.RenderConfigC()
        c.width = int(self.width)
        c.height = int(self.height)
        c.spp = int(self.spp)
        c.max_bounces = int(self.max_bounces)
        c.rr_start = int(self.rr_start)
        c.adaptive_enabled = 1 if self.adaptive_enabled else 0
        c.adaptive_min_spp = int(self.adaptive_min_spp)
        c.precision = int(self.precision)
        c.adaptive_threshold = float(self.adaptive_threshold)
        c.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        c.samples_per_pass = int(self.samples_per_pass)
        c.traversal_refill = int(self.traversal_refill)
        return c
