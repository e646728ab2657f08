"""The FP32 traversal (tpt_trace32.h) built for the host (tests/native/libhost_trace.so) against
the oracle's FP64 World::hit, on the CPU.

Covers both top-level forms: small scenes whose instances are flattened into one world-space
tree (Cornell box, material scene), and scenes traversed instance by instance (the displaced
mesh is above the flattening limit when instanced 64 times).  The same (kind, instance,
triangle) must come back on all rays but those whose answer FP32 cannot decide: equal-distance
ties and rays within 1e-5 of an edge (the oracle itself flips on a nudged copy).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from tinypt_amd import scenes

from parity_scenes import PARITY_CASES

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


def _rays(rng, n, lo, hi, targets=None, tmax=1e30):
    o = rng.uniform(lo, hi, (n, 3))
    if targets is None:
        d = rng.standard_normal((n, 3))
    else:
        d = targets[rng.integers(len(targets), size=n)] + rng.normal(0, 0.3, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8))
    r[:, :3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, 1e-4, tmax
    return r.astype(np.float32).astype(np.float64)


def _check(sc, rays, expect_flat):
    g, flat = host_trace(sc, rays)
    assert flat == expect_flat
    orc = oracle.OracleScene(sc)
    r = orc.trace(rays, threads=8)
    same = (g["kind"] == r["kind"]) & (g["inst"] == r["inst"]) & (g["prim"] == r["prim"])
    unexplained = 0
    for i in np.nonzero(~same)[0]:
        if g["kind"][i] and r["kind"][i] and abs(float(g["t"][i]) - r["t"][i]) <= 1e-4 * max(1.0, r["t"][i]):
            continue
        ok = False
        for ax in range(3):
            for s in (-1e-5, 1e-5):
                q = rays[i:i + 1].copy()
                q[0, ax] += s * max(1.0, abs(q[0, ax]))
                rq = orc.trace(q, threads=1)
                ok |= rq["kind"][0] == g["kind"][i] and rq["inst"][0] == g["inst"][i] and rq["prim"][0] == g["prim"][i]
        unexplained += not ok
    hit = same & (r["kind"] != 0)
    rel = np.abs(g["t"][hit] - r["t"][hit]) / np.maximum(r["t"][hit], 1.0)
    print(f"agree {same.mean():.6f}, unexplained {unexplained}, t rel err p99.9 {np.quantile(rel, 0.999):.2e}")
    assert unexplained <= 1e-4 * len(rays)
    assert np.quantile(rel, 0.999) < 1e-4


def test_flattened_cornell():
    rng = np.random.default_rng(1)
    _check(scenes.cornell_box(16, 16), _rays(rng, 40_000, (-0.95, 0.05, -0.95), (0.95, 1.95, 0.95)), True)


def test_flattened_material_scene_with_spheres_and_tmax():
    rng = np.random.default_rng(2)
    sc = PARITY_CASES["material_zoo"][0]()
    targets = np.array([[-1.2, 0.4, 0.0], [-0.4, 0.4, 0.3], [0.4, 0.4, 0.0], [1.2, 0.4, 0.3], [0.0, 0.25, -1.0]])
    _check(sc, _rays(rng, 40_000, (-3, 0.05, -3), (3, 3, 3), targets, tmax=3.0), True)


def test_two_level_instanced():
    rng = np.random.default_rng(3)
    sc = scenes.instanced_scene(16, 16, n_variants=2, n_instances=64, subdivisions=6, box=3.0, env_size=(8, 4))
    targets = np.array([inst.xform.t for inst in sc.world.instances])
    _check(sc, _rays(rng, 20_000, (-4.5, -4.5, -4.5), (4.5, 4.5, 4.5), targets), False)


def test_flattening_keeps_mirrored_instance_normals():
    """An instance with det(A) < 0 flips e1 x e2 in world space; the hit must still report the
    reference's instance and triangle (normal orientation is checked on the GPU renders)."""
    rng = np.random.default_rng(4)
    b = scenes.SceneBuilder()
    from tinypt_amd.scene import Camera, EnvMap, Lambert, Transform
    b.mesh_instance(scenes.unit_cube_tris(), scenes.compose_ops([Transform.scale((-1.0, 0.5, 1.0))]), Lambert((0.5,) * 3))
    b.mesh_instance(scenes.unit_cube_tris(), scenes.compose_ops([Transform.translate((2.0, 0, 0))]), Lambert((0.5,) * 3))
    sc = b.build(Camera.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 1), EnvMap.constant((1.0, 1.0, 1.0)))
    _check(sc, _rays(rng, 20_000, (-3, -3, -3), (3, 3, 3), np.array([[0.0, 0, 0], [2.0, 0, 0]])), True)


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
