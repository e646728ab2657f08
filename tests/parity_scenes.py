"""Small scenes for the GPU parity tests and the statistics used to compare two renders.

Every scene here is small enough for the FP64 oracle to render in about a second, and together
they exercise every material, both light kinds, the environment map with its NEE, instancing
with rotation and non-uniform scale, and both background kinds.
"""
import os

import numpy as np

from tinypt_amd import scenes
from tinypt_amd.scene import (Camera, Dielectric, DiffuseLight, EnvMap, Ggx, Lambert, Metal, Scene, Sphere, Subsurface,
                              Transform, World)


def oracle_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def floorlight() -> Scene:
    """A floor rectangle under a rotated, scaled rectangle light."""
    b = scenes.SceneBuilder()
    b.mesh_instance(scenes.unit_rectangle_tris(), scenes.compose_ops([Transform.rotate((1, 0, 0), -90.0)]),
                    Lambert((0.5, 0.5, 0.5)))
    b.mesh_instance(scenes.unit_rectangle_tris(),
                    scenes.compose_ops([Transform.translate((0, 1, 0)), Transform.rotate((1, 0, 0), 90.0),
                                        Transform.scale((0.3, 0.3, 1))]), DiffuseLight((5, 5, 5)))
    return b.build(Camera.look_at((0, 2, 3), (0, 0, 0), (0, 1, 0), 40, 1), EnvMap.constant((0, 0, 0)))


def lambert_sphere_env(albedo=0.6, e=1.0) -> Scene:
    """A convex Lambert sphere in a constant environment: sphere pixels converge to albedo * E."""
    w = World()
    w.spheres.append(Sphere((0.0, 0.0, 0.0), 1.0, 0))
    return Scene(Camera.look_at((0, 0, 4), (0, 0, 0), (0, 1, 0), 20, 1), w, [Lambert((albedo,) * 3)],
                 EnvMap.constant((e, e, e)))


def material_zoo() -> Scene:
    """Every material kind (sphere and mesh instances) under a sphere light and a small image
    environment, with a mirror-and-glass pair that makes long delta chains."""
    b = scenes.SceneBuilder()
    T, R, S = Transform.translate, Transform.rotate, Transform.scale
    b.mesh_instance(scenes.unit_rectangle_tris(), scenes.compose_ops([R((1, 0, 0), -90.0), S((4.0, 4.0, 1.0))]),
                    Lambert((0.6, 0.6, 0.6)))
    b.sphere((-1.2, 0.4, 0.0), 0.4, Ggx((0.9, 0.6, 0.3), 0.2))
    b.sphere((-0.4, 0.4, 0.3), 0.4, Dielectric(1.5, (0.1, 0.3, 0.6)))
    b.sphere((0.4, 0.4, 0.0), 0.4, Metal((0.8, 0.8, 0.8)))
    b.sphere((1.2, 0.4, 0.3), 0.4, Subsurface((0.7, 0.5, 0.4), 0.1))
    b.mesh_instance(scenes.unit_cube_tris(), scenes.compose_ops([T((0.0, 0.25, -1.0)), R((0, 1, 0), 30.0),
                                                                  S((0.8, 0.25, 0.3))]), Ggx((0.3, 0.8, 0.4), 0.45))
    b.sphere((0.5, 2.5, 1.0), 0.3, DiffuseLight((10.0, 9.0, 8.0)))
    env = scenes.synthetic_env(64, 32, seed=4, sun_peak=40.0, sun_width_deg=6.0)
    return b.build(Camera.look_at((0, 1.3, 3.5), (0, 0.4, 0), (0, 1, 0), 45, 1), env)


def mesh_small() -> Scene:
    """The displaced-icosphere mesh scene at 2 subdivisions (320 triangles, 20 materials) under a
    64x32 synthetic environment with its sun lobe."""
    return scenes.mesh_scene(32, 32, subdivisions=2, env_size=(64, 32))


def instanced_small() -> Scene:
    """64 instances of 2 icosphere variants (rotation, uniform scale, per-instance materials
    including subsurface) under a 64x32 environment."""
    return scenes.instanced_scene(32, 32, n_variants=2, n_instances=64, subdivisions=2, box=3.0, env_size=(64, 32))


# name -> (builder, width, height, spp, max_bounces, rr_start, seed)
PARITY_CASES = {
    "cornell": (lambda: scenes.cornell_box(32, 32), 32, 32, 256, 12, 4, 2),
    "default_black": (lambda: scenes.default_scene(32, 32, "black"), 32, 32, 256, 8, 3, 1),
    "default_sky": (lambda: scenes.default_scene(24, 16, "sky"), 24, 16, 256, 8, 3, 11),
    "floorlight": (floorlight, 32, 32, 512, 4, 8, 7),
    "material_zoo": (material_zoo, 32, 32, 256, 8, 3, 99),
    "mesh_small": (mesh_small, 32, 32, 128, 16, 3, 3),
    "instanced_small": (instanced_small, 32, 32, 128, 8, 3, 5),
}


def compare(acc_a, sq_a, n_a, acc_b, sq_b, n_b):
    """Per-channel z scores of two independent per-pixel means, and the z score of the image
    means.  acc/sq are per-pixel sums of samples and of squares, n the samples per pixel."""
    n_a = np.maximum(np.asarray(n_a, dtype=np.float64), 1.0)[..., None]
    n_b = np.maximum(np.asarray(n_b, dtype=np.float64), 1.0)[..., None]
    ma, mb = acc_a / n_a, acc_b / n_b
    va = np.maximum(sq_a / n_a - ma * ma, 0.0) / n_a
    vb = np.maximum(sq_b / n_b - mb * mb, 0.0) / n_b
    den = np.sqrt(va + vb)
    diff = ma - mb
    z = np.where(den > 0, diff / np.where(den > 0, den, 1.0), np.where(np.abs(diff) > 1e-5 * (np.abs(mb) + 1e-3), np.inf, 0.0))
    npx = ma.shape[0] * ma.shape[1]
    zm = (ma.sum((0, 1)) - mb.sum((0, 1))) / np.sqrt((va + vb).sum((0, 1)) + 1e-300)
    return z, zm, ma.mean(), mb.mean(), npx
