"""GPU closest-hit queries (tpt_trace_closest) against the FP64 oracle's World::hit.

The FP32 traversal must name the same (kind, instance, primitive) as the reference's FP64
World::hit (world.rs:81-121) on at least 99.99% of the rays a render traces (camera rays and
two generations of bounce rays: tests/trace_checks.py), raw, with nothing set aside (the
1.31M-triangle mesh and the 4096-instance scene: test_gpu_heavy_scenes.py).  On
random stress rays every disagreement must be one FP32 cannot decide (equal-distance ties,
edge crossings, the reference's tmin rule for instances, rays the oracle itself answers
differently when nudged by 1e-5).  Distances of agreeing hits: p99.9 of |dt| / max(t, 1)
below 1e-5.
"""
import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import scenes

from parity_scenes import PARITY_CASES, oracle_threads
from trace_checks import classify, path_rays, report, stress_rays

pytestmark = pytest.mark.gpu


def _compare(name, sc, rays, path):
    ds = T.DeviceScene(sc, T.default_context())
    g = ds.trace_closest(rays)
    rep = classify(sc, rays, g, threads=oracle_threads())
    report(name, rep)
    if path:
        assert rep["agree"] >= 0.9999, (name, rep["agree"])
    assert len(rep["other"]) <= 1e-4 * len(rays), rep["other"][:10]
    assert rep["t_rel_p999"] < 1e-5, rep["t_rel_p999"]
    return rep


def test_trace_cornell(gpu_ctx):
    rng = np.random.default_rng(1)
    sc = scenes.cornell_box(64, 64)
    _compare("cornell stress", sc, stress_rays(rng, 200_000, (-0.95, 0.05, -0.95), (0.95, 1.95, 0.95)), False)
    _compare("cornell path", sc, path_rays(sc, rng, 100_000), True)


def test_trace_spheres_and_miss(gpu_ctx):
    rng = np.random.default_rng(7)
    sc = scenes.default_scene(64, 64, "sky")
    targets = np.array([[-1.8, 0.5, 0.0], [-0.6, 0.5, 0.0], [0.6, 0.5, 0.0], [1.8, 0.5, 0.0], [0.0, 3.0, -1.0]])
    _compare("spheres stress", sc, stress_rays(rng, 100_000, (-4, 0.1, -4), (4, 4, 4), targets), False)
    _compare("spheres path", sc, path_rays(sc, rng, 100_000), True)


def test_trace_instanced(gpu_ctx):
    rng = np.random.default_rng(11)
    sc = PARITY_CASES["instanced_small"][0]()
    targets = np.array([inst.xform.t for inst in sc.world.instances], dtype=np.float64)
    _compare("instanced stress", sc, stress_rays(rng, 200_000, (-4.5, -4.5, -4.5), (4.5, 4.5, 4.5), targets), False)
    _compare("instanced path", sc, path_rays(sc, rng, 100_000), True)


def test_trace_mesh_with_tmax(gpu_ctx):
    """Finite tmax (shadow-ray style intervals) and a mesh with 5,120 triangles."""
    rng = np.random.default_rng(3)
    sc = scenes.mesh_scene(32, 32, subdivisions=4, env_size=(16, 8))
    _compare("mesh_tmax stress", sc, stress_rays(rng, 100_000, (-2.5, 0.05, -2.5), (2.5, 2.5, 2.5),
                                                 np.array([[0.0, 1.05, 0.0]]), tmax=2.0), False)
