"""The FP32 traversal (tpt_trace32.h) built for the host (tests/native/libhost_trace.so) against
the oracle's FP64 World::hit, on the CPU.

Covers both top-level forms: small scenes whose instances are flattened into one world-space
tree (Cornell box, material scene), and scenes traversed instance by instance (the displaced
mesh is above the flattening limit when instanced 64 times).  The same (kind, instance,
triangle) must come back on at least 99.99% of the rays a render traces, and every disagreement
on random stress rays must be one FP32 cannot decide (tests/trace_checks.py: equal-distance
ties, edge crossings, the reference's tmin rule for instances, rays the oracle itself answers
differently when nudged by 1e-5).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from tinypt_amd import scenes

from parity_scenes import PARITY_CASES
from trace_checks import classify, path_rays, report, stress_rays

LIB = os.path.join(os.path.dirname(__file__), "native", "libhost_trace.so")


def _lib():
    """The host build of the traversal, (re)built by tests/native/build_host_trace.py when a source is newer."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "native"))
    import build_host_trace
    if build_host_trace.is_stale():
        build_host_trace.build()
    return C.CDLL(LIB)


def host_trace(sc, rays):
    desc, keep = sc.to_desc()
    r = np.ascontiguousarray(rays, dtype=np.float32)
    n = len(r)
    t = np.zeros(n, np.float32)
    kind, inst, prim = (np.zeros(n, np.int32) for _ in range(3))
    flat, bound, deepest = C.c_int32(), C.c_int32(), C.c_int32()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    rc = _lib().tpt_test_host_trace(C.byref(desc), C.c_int64(n), r.ctypes.data_as(C.POINTER(C.c_float)),
                                    t.ctypes.data_as(C.POINTER(C.c_float)), ip(kind), ip(inst), ip(prim), C.byref(flat),
                                    C.byref(bound), C.byref(deepest))
    assert rc == 0
    # the kernels size their stacks by this bound (LDS part + overflow area)
    assert deepest.value <= bound.value, (deepest.value, bound.value)
    print(f"stack entries used {deepest.value} of bound {bound.value}")
    return dict(t=t, kind=kind, inst=inst, prim=prim), bool(flat.value)


def _check(sc, rays, expect_flat, path=False):
    """Path rays: raw agreement with the oracle >= 99.99%.  Stress rays: every disagreement
    explained (a tie, an edge, the reference's tmin rule or a nudge; tests/trace_checks.py)."""
    g, flat = host_trace(sc, rays)
    assert flat == expect_flat
    rep = classify(sc, rays, g)
    report("path rays" if path else "stress rays", rep)
    if path:
        assert rep["agree"] >= 0.9999, rep
    assert len(rep["other"]) <= 1e-4 * len(rays), rep["other"][:10]
    assert rep["t_rel_p999"] < 1e-5


def test_flattened_cornell():
    rng = np.random.default_rng(1)
    sc = scenes.cornell_box(16, 16)
    _check(sc, stress_rays(rng, 40_000, (-0.95, 0.05, -0.95), (0.95, 1.95, 0.95)), True)
    _check(sc, path_rays(sc, rng, 20_000), True, path=True)


def test_flattened_material_scene_with_spheres_and_tmax():
    rng = np.random.default_rng(2)
    sc = PARITY_CASES["material_zoo"][0]()
    targets = np.array([[-1.2, 0.4, 0.0], [-0.4, 0.4, 0.3], [0.4, 0.4, 0.0], [1.2, 0.4, 0.3], [0.0, 0.25, -1.0]])
    _check(sc, stress_rays(rng, 40_000, (-3, 0.05, -3), (3, 3, 3), targets, tmax=3.0), True)
    _check(sc, path_rays(sc, rng, 20_000), True, path=True)


def test_two_level_instanced():
    rng = np.random.default_rng(3)
    sc = scenes.instanced_scene(16, 16, n_variants=2, n_instances=64, subdivisions=6, box=3.0, env_size=(8, 4))
    targets = np.array([inst.xform.t for inst in sc.world.instances])
    _check(sc, stress_rays(rng, 20_000, (-4.5, -4.5, -4.5), (4.5, 4.5, 4.5), targets), False)
    _check(sc, path_rays(sc, rng, 20_000), False, path=True)


def test_flattening_keeps_mirrored_instance_normals():
    """An instance with det(A) < 0 flips e1 x e2 in world space; the hit must still report the
    reference's instance and triangle (normal orientation is checked on the GPU renders)."""
    rng = np.random.default_rng(4)
    b = scenes.SceneBuilder()
    from tinypt_amd.scene import Camera, EnvMap, Lambert, Transform
    b.mesh_instance(scenes.unit_cube_tris(), scenes.compose_ops([Transform.scale((-1.0, 0.5, 1.0))]), Lambert((0.5,) * 3))
    b.mesh_instance(scenes.unit_cube_tris(), scenes.compose_ops([Transform.translate((2.0, 0, 0))]), Lambert((0.5,) * 3))
    sc = b.build(Camera.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 1), EnvMap.constant((1.0, 1.0, 1.0)))
    _check(sc, stress_rays(rng, 20_000, (-3, -3, -3), (3, 3, 3), np.array([[0.0, 0, 0], [2.0, 0, 0]])), True)


def test_instance_world_boxes_are_tight_and_conservative():
    """An instance's top-level box is built from the corners of a cut of its mesh tree through the
    instance transform: it contains every transformed vertex (the traversal may skip the instance
    otherwise) and, for rotated round meshes, stays close to the exact box, where the root box's
    eight corners grow it by up to sqrt(3) per axis."""
    rng = np.random.default_rng(6)
    tris, _ = scenes.displaced_icosphere(4, 0.05, 7)
    tv = np.ascontiguousarray(tris.reshape(-1, 9), dtype=np.float64)
    lib = _lib()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    worst_cut, mean_root, mean_cut = 0.0, [], []
    for _ in range(40):
        a = np.ascontiguousarray(scenes.random_rotation(rng) * rng.uniform(0.2, 1.0))
        t = rng.uniform(-25, 25, 3)
        bc, br = np.zeros(6), np.zeros(6)
        assert lib.tpt_test_instance_box(C.c_int64(len(tv)), dp(tv), dp(a), dp(t), dp(bc), dp(br)) == 0
        w = tris.reshape(-1, 3) @ a.T + t
        lo, hi = w.min(0), w.max(0)
        assert np.all(bc[:3] <= lo) and np.all(bc[3:] >= hi)  # conservative
        ext = hi - lo
        worst_cut = max(worst_cut, float(((bc[3:] - bc[:3]) / ext).max()))
        mean_cut.append(float(((bc[3:] - bc[:3]) / ext).mean()))
        mean_root.append(float(((br[3:] - br[:3]) / ext).mean()))
    print(f"cut box / exact extent: worst {worst_cut:.3f}, mean {np.mean(mean_cut):.3f}; root-corner box / exact: mean {np.mean(mean_root):.3f}")
    assert worst_cut < 1.15 and np.mean(mean_cut) < 1.1
    assert np.mean(mean_root) > 1.3
