"""GPU closest-hit queries (tpt_trace_closest) against the FP64 oracle's World::hit.

The FP32 traversal must name the same (kind, instance, primitive) as the reference's FP64
World::hit (world.rs:81-121) on at least 99.99% of rays once rays on which the answer is not
determined at FP32 precision are set aside: a disagreement is set aside when the two hits lie at
the same distance (shared edges, coplanar faces) or when the oracle itself returns the GPU's
answer for the ray nudged by 1e-5 (grazing and silhouette rays).
"""
import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import scenes

from parity_scenes import PARITY_CASES, oracle_threads

pytestmark = pytest.mark.gpu


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _rays(rng, n, lo, hi, targets=None, tmin=1e-4, tmax=1e30):
    o = rng.uniform(lo, hi, (n, 3))
    if targets is None:
        d = _unit(rng.standard_normal((n, 3)))
    else:
        t = targets[rng.integers(len(targets), size=n)] + rng.normal(0, 0.3, (n, 3))
        d = _unit(t - o)
    r = np.zeros((n, 8))
    r[:, :3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, tmin, tmax
    # the GPU sees FP32 rays: give the oracle the same rounded values
    return r.astype(np.float32).astype(np.float64)


def _compare(name, sc, rays):
    ds = T.DeviceScene(sc, T.default_context())
    g = ds.trace_closest(rays)
    orc = oracle.OracleScene(sc)
    r = orc.trace(rays, threads=oracle_threads())
    same = (g["kind"] == r["kind"]) & (g["inst"] == r["inst"]) & (g["prim"] == r["prim"])
    bad = np.nonzero(~same)[0]
    unexplained = []
    for i in bad:
        both_hit = g["kind"][i] != 0 and r["kind"][i] != 0
        if both_hit and abs(float(g["t"][i]) - r["t"][i]) <= 1e-4 * max(1.0, r["t"][i]):
            continue  # same distance: a tie the two precisions break differently
        explained = False
        for ax in range(3):
            for s in (-1e-5, 1e-5):
                q = rays[i:i + 1].copy()
                q[0, ax] += s * max(1.0, abs(q[0, ax]))
                rq = orc.trace(q, threads=1)
                if rq["kind"][0] == g["kind"][i] and rq["inst"][0] == g["inst"][i] and rq["prim"][0] == g["prim"][i]:
                    explained = True
        if not explained:
            unexplained.append(int(i))
    hit = same & (r["kind"] != 0)
    rel = np.abs(g["t"][hit].astype(np.float64) - r["t"][hit]) / np.maximum(r["t"][hit], 1.0)
    q999 = float(np.quantile(rel, 0.999)) if hit.any() else 0.0
    print(f"{name}: {len(rays)} rays, hit {float((r['kind'] != 0).mean()):.3f}, agree {float(same.mean()):.6f}, "
          f"unexplained {len(unexplained)}, t rel err p99.9 {q999:.2e}", flush=True)
    assert len(unexplained) <= 1e-4 * len(rays), unexplained[:10]
    assert q999 < 1e-4


def test_trace_cornell(gpu_ctx):
    rng = np.random.default_rng(1)
    sc = scenes.cornell_box(64, 64)
    rays = _rays(rng, 200_000, (-0.95, 0.05, -0.95), (0.95, 1.95, 0.95))
    _compare("cornell", sc, rays)


def test_trace_spheres_and_miss(gpu_ctx):
    rng = np.random.default_rng(7)
    sc = scenes.default_scene(64, 64, "sky")
    targets = np.array([[-1.8, 0.5, 0.0], [-0.6, 0.5, 0.0], [0.6, 0.5, 0.0], [1.8, 0.5, 0.0], [0.0, 3.0, -1.0]])
    rays = _rays(rng, 100_000, (-4, 0.1, -4), (4, 4, 4), targets)
    _compare("spheres", sc, rays)


def test_trace_instanced(gpu_ctx):
    rng = np.random.default_rng(11)
    sc = PARITY_CASES["instanced_small"][0]()
    targets = np.array([inst.xform.t for inst in sc.world.instances], dtype=np.float64)
    rays = _rays(rng, 200_000, (-4.5, -4.5, -4.5), (4.5, 4.5, 4.5), targets)
    _compare("instanced", sc, rays)


def test_trace_mesh_with_tmax(gpu_ctx):
    """Finite tmax (shadow-ray style intervals) and a mesh with 5,120 triangles."""
    rng = np.random.default_rng(3)
    sc = scenes.mesh_scene(32, 32, subdivisions=4, env_size=(16, 8))
    rays = _rays(rng, 100_000, (-2.5, 0.05, -2.5), (2.5, 2.5, 2.5), np.array([[0.0, 1.05, 0.0]]), tmax=2.0)
    _compare("mesh_tmax", sc, rays)
