"""Seeded scene generators for the renderer's benchmark and test workloads.

* ``default_scene``   -- sample/default.xml as loaded by mitsuba.rs:137-222 (six spheres, black
                         constant background), or scene.rs:26-95 with the procedural sky.
* ``cornell_box``     -- the rectangles + cubes of sample/cornell.xml:31-82 with configurable
                         wall albedos and light radiance.
* ``displaced_icosphere`` / ``synthetic_env`` / ``mesh_scene`` -- a subdivided icosphere with
                         seeded value-noise displacement, a ground rectangle, per-object
                         materials and a log-normal sky with a sun lobe.
* ``instanced_scene`` -- mesh variants instanced at seeded rotations, scales and positions.
* ``inverse_scene``   -- Cornell box plus GGX spheres for gradient checks.

Canonical Mitsuba shapes follow mitsuba.rs:414-469; transforms compose as parse_transform
(mitsuba.rs:483-491): trafo = op1 . op2 . ... . opN.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .scene import (Camera, DiffuseLight, Dielectric, EnvMap, Ggx, Instance, Lambert, Material, Mesh, Metal, Scene,
                    Sphere, Subsurface, Transform, World, color_from_srgb)


# ---------------------------------------------------------------- canonical shapes (mitsuba.rs:414-469)
def unit_rectangle_tris() -> np.ndarray:
    a, b, c, d = (-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (1.0, 1.0, 0.0), (-1.0, 1.0, 0.0)
    return np.array([[a, b, c], [a, c, d]], dtype=np.float64)


def unit_cube_tris() -> np.ndarray:
    v = [(-1.0, -1.0, -1.0), (1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (-1.0, 1.0, -1.0),
         (-1.0, -1.0, 1.0), (1.0, -1.0, 1.0), (1.0, 1.0, 1.0), (-1.0, 1.0, 1.0)]
    quads = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]
    tris = []
    for q in quads:
        tris.append([v[q[0]], v[q[1]], v[q[2]]])
        tris.append([v[q[0]], v[q[2]], v[q[3]]])
    return np.array(tris, dtype=np.float64)


def unit_disk_tris(n: int = 64) -> np.ndarray:
    tris = []
    for i in range(n):
        a0 = 2.0 * math.pi * i / n
        a1 = 2.0 * math.pi * (i + 1) / n
        tris.append([(0.0, 0.0, 0.0), (math.cos(a0), math.sin(a0), 0.0), (math.cos(a1), math.sin(a1), 0.0)])
    return np.array(tris, dtype=np.float64)


def compose_ops(ops: Sequence[Transform]) -> Transform:
    """parse_transform (mitsuba.rs:483-491)."""
    acc = Transform.identity()
    for op in ops:
        acc = acc.compose(op)
    return acc


class SceneBuilder:
    """Accumulates shapes in the order mitsuba.rs:352-411 would (one material per shape)."""

    def __init__(self):
        self.world = World()
        self.mats: List[Material] = []

    def material(self, m: Material) -> int:
        self.mats.append(m)
        return len(self.mats) - 1

    def sphere(self, c, r, mat: Material) -> int:
        mid = self.material(mat)
        self.world.spheres.append(Sphere(tuple(map(float, c)), float(r), mid))
        return mid

    def mesh_instance(self, tris: np.ndarray, xform: Transform, mat: Material, tri_mat=None) -> int:
        mid = self.material(mat)
        mats = mid if tri_mat is None else tri_mat
        self.world.meshes.append(Mesh(tris, mats))
        self.world.instances.append(Instance(len(self.world.meshes) - 1, xform, None))
        return mid

    def build(self, cam: Camera, env: Optional[EnvMap]) -> Scene:
        return Scene(cam, self.world, self.mats, env)


# ---------------------------------------------------------------- default scene
def default_scene(width: int = 256, height: int = 256, background: str = "black") -> Scene:
    """sample/default.xml loaded by load_scene (black 1x1 constant env, mitsuba.rs:218), or with
    background="sky" the built-in build_scene (scene.rs:26-95, env None -> sky())."""
    aspect = width / height
    eye, target = (0.0, 1.2, 4.0), (0.0, 0.5, 0.0)
    d = [eye[i] - target[i] for i in range(3)]
    focus = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    cam = Camera.look_at_dof(eye, target, (0.0, 1.0, 0.0), 40.0, aspect, focus, 0.0)
    b = SceneBuilder()
    b.sphere((0.0, -1000.0, 0.0), 1000.0, Lambert(color_from_srgb(0.5, 0.5, 0.5)))
    b.sphere((-1.8, 0.5, 0.0), 0.5, Lambert(color_from_srgb(0.8, 0.3, 0.3)))
    b.sphere((-0.6, 0.5, 0.0), 0.5, Metal(color_from_srgb(0.8, 0.8, 0.8)))
    b.sphere((0.6, 0.5, 0.0), 0.5, Ggx(color_from_srgb(0.95, 0.78, 0.35), 0.25))
    b.sphere((1.8, 0.5, 0.0), 0.5, Dielectric(1.5 / 1.0, (0.02, 0.05, 0.02)))
    b.sphere((0.0, 3.0, -1.0), 0.8, DiffuseLight((8.0, 7.0, 5.0)))
    env = EnvMap.constant((0.0, 0.0, 0.0)) if background == "black" else None
    return b.build(cam, env)


# ---------------------------------------------------------------- Cornell box
CORNELL_WHITE = (0.73, 0.73, 0.73)
CORNELL_RED = (0.65, 0.05, 0.05)
CORNELL_GREEN = (0.12, 0.45, 0.15)
CORNELL_LIGHT = (17.0, 12.0, 4.0)


def cornell_geometry(b: SceneBuilder, white=CORNELL_WHITE, red=CORNELL_RED, green=CORNELL_GREEN,
                     light=CORNELL_LIGHT, light_scale: float = 0.35):
    """Shapes of sample/cornell.xml:31-82 in file order."""
    T, R, S = Transform.translate, Transform.rotate, Transform.scale
    rect, cube = unit_rectangle_tris(), unit_cube_tris()
    b.mesh_instance(rect, compose_ops([R((1, 0, 0), -90.0)]), Lambert(white))                               # floor y=0
    b.mesh_instance(rect, compose_ops([T((0, 2, 0)), R((1, 0, 0), 90.0)]), Lambert(white))                 # ceiling y=2
    b.mesh_instance(rect, compose_ops([T((0, 1, -1))]), Lambert(white))                                    # back z=-1
    b.mesh_instance(rect, compose_ops([T((-1, 1, 0)), R((0, 1, 0), 90.0)]), Lambert(red))                  # left x=-1
    b.mesh_instance(rect, compose_ops([T((1, 1, 0)), R((0, 1, 0), -90.0)]), Lambert(green))                # right x=1
    b.mesh_instance(cube, compose_ops([T((-0.35, 0.6, -0.2)), R((0, 1, 0), 18.0), S((0.3, 0.6, 0.3))]),
                    Lambert(white))                                                                        # tall box
    b.mesh_instance(cube, compose_ops([T((0.35, 0.3, 0.35)), R((0, 1, 0), -18.0), S((0.3, 0.3, 0.3))]),
                    Lambert(white))                                                                        # short box
    b.mesh_instance(rect, compose_ops([T((0, 1.98, 0)), R((1, 0, 0), 90.0), S((light_scale, light_scale, 1.0))]),
                    DiffuseLight(light))                                                                   # light


def cornell_camera(width: int, height: int) -> Camera:
    eye, target = (0.0, 1.0, 3.9), (0.0, 1.0, 0.0)
    return Camera.look_at_dof(eye, target, (0.0, 1.0, 0.0), 40.0, width / height, 3.9, 0.0)


def cornell_box(width: int = 512, height: int = 512, white=CORNELL_WHITE, red=CORNELL_RED, green=CORNELL_GREEN,
                light=CORNELL_LIGHT) -> Scene:
    """Cornell box built from rectangles and two cubes; black background (no env emitter)."""
    b = SceneBuilder()
    cornell_geometry(b, white, red, green, light)
    return b.build(cornell_camera(width, height), EnvMap.constant((0.0, 0.0, 0.0)))


# ---------------------------------------------------------------- displaced icosphere
def icosphere(subdivisions: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Unit icosphere: (verts (V,3), faces (F,3), face_root (F,) base-face id)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                      [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                      [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                      [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    root = np.arange(20, dtype=np.int64)
    for _ in range(subdivisions):
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e.sort(axis=1)
        # unique edges in row order, through one integer key per edge (lo * V + hi sorts as the
        # rows do)
        nv = np.int64(len(verts))
        keys, inv = np.unique(e[:, 0] * nv + e[:, 1], return_inverse=True)
        inv = inv.reshape(-1)
        uniq = np.stack([keys // nv, keys % nv], 1)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mid])
        nf = len(faces)
        a = inv[:nf] + base
        bb = inv[nf:2 * nf] + base
        c = inv[2 * nf:] + base
        v0, v1, v2 = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([np.stack([v0, a, c], 1), np.stack([v1, bb, a], 1), np.stack([v2, c, bb], 1),
                                np.stack([a, bb, c], 1)])
        root = np.concatenate([root, root, root, root])
    return verts, faces, root


def _hash_u64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele et al. 2014) on uint64 arrays."""
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """Trilinear value noise in [-1, 1] on a hashed integer lattice, smoothstep-weighted."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    w = f * f * (3.0 - 2.0 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix = (i0[:, 0] + dx).astype(np.uint64)
                iy = (i0[:, 1] + dy).astype(np.uint64)
                iz = (i0[:, 2] + dz).astype(np.uint64)
                with np.errstate(over="ignore"):
                    key = (ix * np.uint64(73856093)) ^ (iy * np.uint64(19349663)) ^ (iz * np.uint64(83492791)) ^ \
                          (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15))
                h = _hash_u64(key)
                val = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * val
    return out


def displaced_icosphere(subdivisions: int = 8, amplitude: float = 0.05, seed: int = 3,
                        freq: float = 4.0) -> Tuple[np.ndarray, np.ndarray]:
    """Triangles (F,3,3) of an icosphere whose vertices sit at radius 1 + amplitude*vnoise(p),
    and the base-face id (0..19) of every triangle."""
    verts, faces, root = icosphere(subdivisions)
    r = 1.0 + amplitude * value_noise(verts, seed, freq)
    verts = verts * r[:, None]
    return verts[faces], root


def synthetic_env(width: int = 2048, height: int = 1024, seed: int = 4, sun_peak: float = 1.0e4,
                  sun_width_deg: float = 0.5, sun_dir=(0.4, 0.6, -0.7)) -> EnvMap:
    """Log-normal sky gradient exp(mu(elevation) + 0.25 N(0,1)) plus a Gaussian sun lobe."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = math.pi * v[:, None] * np.ones((1, width))
    phi = 2.0 * math.pi * u[None, :] * np.ones((height, 1))
    d = np.stack([np.cos(phi) * np.sin(theta), np.cos(theta), np.sin(phi) * np.sin(theta)], -1)
    elev = math.pi / 2 - theta
    mu = -0.5 + 1.2 * (elev / (math.pi / 2))
    sky = np.exp(mu + 0.25 * rng.standard_normal((height, width)))
    tint = np.stack([0.55 + 0.1 * (elev > 0), 0.7 + 0.05 * (elev > 0), np.ones_like(elev)], -1)
    img = sky[..., None] * tint
    s = np.asarray(sun_dir, dtype=np.float64)
    s = s / np.linalg.norm(s)
    gamma = np.degrees(np.arccos(np.clip((d * s).sum(-1), -1.0, 1.0)))
    sun = sun_peak * np.exp(-(gamma / sun_width_deg) ** 2)
    img = img + sun[..., None] * np.array([1.0, 0.92, 0.8])
    return EnvMap.from_pixels(width, height, img)


def random_material(rng: np.random.Generator, kinds=("lambert", "ggx", "dielectric")) -> Material:
    """Lambert albedo U[.1,.9], GGX alpha U[.05,.5], dielectric ior 1.5 with absorption U[0,2],
    subsurface scatter_dist U[.01,.2]."""
    k = kinds[int(rng.integers(len(kinds)))]
    if k == "lambert":
        return Lambert(tuple(rng.uniform(0.1, 0.9, 3)))
    if k == "ggx":
        return Ggx(tuple(rng.uniform(0.1, 0.95, 3)), float(rng.uniform(0.05, 0.5)))
    if k == "dielectric":
        return Dielectric(1.5, tuple(rng.uniform(0.0, 2.0, 3)))
    return Subsurface(tuple(rng.uniform(0.1, 0.9, 3)), float(rng.uniform(0.01, 0.2)))


def mesh_scene(width: int = 1920, height: int = 1080, subdivisions: int = 8, seed: int = 3,
               env: Optional[EnvMap] = None, env_size=(2048, 1024)) -> Scene:
    """Displaced icosphere (20 base-face objects, each with a seeded material) on a ground
    rectangle, lit by the synthetic HDR environment."""
    rng = np.random.default_rng(seed)
    tris, root = displaced_icosphere(subdivisions, 0.05, seed)
    b = SceneBuilder()
    face_mats = [b.material(random_material(rng)) for _ in range(20)]
    tri_mat = np.asarray(face_mats, dtype=np.int32)[root]
    b.world.meshes.append(Mesh(tris, tri_mat))
    b.world.instances.append(Instance(0, Transform.translate((0.0, 1.05, 0.0)), None))
    ground = compose_ops([Transform.rotate((1, 0, 0), -90.0), Transform.scale((8.0, 8.0, 1.0))])
    b.mesh_instance(unit_rectangle_tris(), ground, Lambert(tuple(rng.uniform(0.1, 0.9, 3))))
    cam = Camera.look_at_dof((0.0, 1.6, 4.2), (0.0, 0.95, 0.0), (0.0, 1.0, 0.0), 40.0, width / height, 4.25, 0.0)
    if env is None:
        env = synthetic_env(env_size[0], env_size[1], seed=4)
    return b.build(cam, env)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation matrix from a uniform unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_scene(width: int = 1920, height: int = 1080, n_variants: int = 4, n_instances: int = 4096,
                    subdivisions: int = 8, box: float = 25.0, seed: int = 6, env: Optional[EnvMap] = None,
                    env_size=(2048, 1024)) -> Scene:
    """Displaced-icosphere variants instanced at seeded random rotations, scales U[.2,1] and
    positions U[-box, box]^3, with per-instance Lambert/GGX/dielectric/subsurface materials."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    base_mat = b.material(Lambert((0.5, 0.5, 0.5)))
    for k in range(n_variants):
        tris, _ = displaced_icosphere(subdivisions, 0.05, seed * 100 + k)
        b.world.meshes.append(Mesh(tris, base_mat))
    for i in range(n_instances):
        rot = random_rotation(rng)
        s = float(rng.uniform(0.2, 1.0))
        pos = rng.uniform(-box, box, 3)
        mid = b.material(random_material(rng, ("lambert", "ggx", "dielectric", "subsurface")))
        b.world.instances.append(Instance(i % n_variants, Transform.from_affine(rot * s, tuple(pos)), mid))
    cam = Camera.look_at_dof((0.0, 0.0, 2.2 * box), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 55.0, width / height, 2.2 * box, 0.0)
    if env is None:
        env = synthetic_env(env_size[0], env_size[1], seed=4)
    return b.build(cam, env)


# ---------------------------------------------------------------- inverse rendering
def inverse_scene(params: np.ndarray, width: int = 256, height: int = 256, n_spheres: int = 16, radius: float = 0.12,
                  with_cubes: bool = True) -> Scene:
    """Cornell box plus ``n_spheres`` GGX spheres on a grid; params (n_spheres, 4) = F0 rgb,
    alpha.  The material of sphere i is ``inverse_material_ids(n_spheres, with_cubes)[i]``."""
    params = np.asarray(params, dtype=np.float64).reshape(n_spheres, 4)
    b = SceneBuilder()
    cornell_geometry(b)
    if not with_cubes:  # drop the two boxes (shapes 5 and 6), keep walls and light
        del b.world.meshes[5:7], b.world.instances[5:7], b.mats[5:7]
        for k, inst in enumerate(b.world.instances):
            inst.mesh_id = k
            b.world.meshes[k].mat_ids[:] = k
    k = int(math.ceil(math.sqrt(n_spheres)))
    for i in range(n_spheres):
        gx, gy = i % k, i // k
        x = -0.75 + 1.5 * (gx + 0.5) / k
        if with_cubes:
            # a vertical grid in front of the boxes (z = 0.8, above the short box's top at
            # y = 0.6): every sphere in view, none touching a box
            y = 0.75 + 0.9 * gy / max(k - 1, 1)
            z = 0.8
        else:  # on the floor
            y = radius + 0.01
            z = -0.75 + 1.5 * (gy + 0.5) / k
        b.sphere((x, y, z), radius, Ggx(tuple(params[i, :3]), float(params[i, 3])))
    return b.build(cornell_camera(width, height), EnvMap.constant((0.0, 0.0, 0.0)))


def inverse_params(seed: int = 5, n_spheres: int = 16) -> np.ndarray:
    """Initial parameters U[.2,.8] (n_spheres, 4)."""
    return np.random.default_rng(seed).uniform(0.2, 0.8, (n_spheres, 4))


def inverse_material_ids(n_spheres: int = 16, with_cubes: bool = True) -> List[int]:
    """Material index of each sphere in ``inverse_scene`` (the Cornell shapes come first)."""
    first = 8 if with_cubes else 6
    return [first + i for i in range(n_spheres)]
