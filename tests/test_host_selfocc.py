"""The FP32 kernels' light self-occlusion decision against the FP64 oracle, on the CPU.

The reference's shadow ray to a sampled point of an area light ends 1e-4 short of the light,
so whether the light blocks its own sample is decided by the last bits of the f64 arithmetic
(integrator.rs:178-213, world.rs:81-121): about 30% of samples on a rectangle light.  The FP32
kernels restate that decision in FP64 (tpt_selfocc64.h) with inputs of the reference's
provenance: the shading point that the FP64 World::hit computes along the path's FP32 ray, and
53-bit surface uniforms.  These tests hold the restatement (traced for the host by
tests/native/libhost_trace.so) against the oracle:

* sample for sample, against the oracle fed the same FP32 ray and uniforms;
* in rate, against the oracle's own FP64 rays -- the rate the reference's images carry.  Fed
  the FP32 hit point o + t d instead, the rate was 0.7 points high on the Cornell box and 2.8
  points high under the rotated light of `floorlight` (the direct light 1-4% too dark)."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from tinypt_amd import _abi, scenes

from parity_scenes import floorlight, material_zoo

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "native"))


def _lib():
    import build_host_trace
    if build_host_trace.is_stale():
        build_host_trace.build()
    return C.CDLL(build_host_trace.LIB)


def product_decisions(sc, rays32, light, uv):
    desc, keep = sc.to_desc()
    out = np.zeros(len(light), np.int32)
    rc = _lib().tpt_test_selfocc(C.byref(desc), C.c_int64(len(light)), rays32.ctypes.data_as(C.POINTER(C.c_float)),
                                 light.ctypes.data_as(C.POINTER(C.c_int32)), uv.ctypes.data_as(C.POINTER(C.c_double)),
                                 out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert rc == 0, rc
    return out


def oracle_decisions(o, light, p, uv):
    """0 unoccluded, 1 / 3 blocked by the sampled light's own primitive / instance, 2 by other
    geometry (orc_light_shadow)."""
    out = np.zeros(len(light), np.int32)
    oracle.lib().orc_light_shadow(o.h, len(light), light.ctypes.data_as(C.POINTER(C.c_int32)), oracle._d(p),
                                  oracle._d(uv), oracle._d(np.zeros(len(light))), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def camera_hits(sc, o, n, seed):
    rng = np.random.default_rng(seed)
    c = _abi.CameraDesc()
    c.eye[:], c.target[:], c.up[:] = sc.cam.eye, sc.cam.target, sc.cam.up
    c.vfov_deg, c.aspect, c.focus_dist, c.aperture = sc.cam.vfov_deg, sc.cam.aspect, sc.cam.focus_dist, sc.cam.aperture
    r = oracle.camera_rays(c, seed, rng.uniform(-1, 1, (n, 2)))
    rays = np.concatenate([r[:, :6], np.full((n, 1), 1e-4), np.full((n, 1), 1e30)], 1)
    rays32 = np.ascontiguousarray(rays.astype(np.float32))
    return rng, rays, rays32


CASES = [("cornell", lambda: scenes.cornell_box(64, 64), 1_000_000),
         ("floorlight", floorlight, 400_000),
         ("material_zoo", material_zoo, 200_000)]  # a sphere light


@pytest.mark.parametrize("name,make,n", CASES, ids=[c[0] for c in CASES])
def test_self_occlusion_matches_the_oracle(name, make, n):
    sc = make()
    o = oracle.OracleScene(sc)
    rng, rays, rays32 = camera_hits(sc, o, n, 17)
    n_lights = len(o.lights())
    light = rng.integers(0, n_lights, n).astype(np.int32)
    uv = np.ascontiguousarray(rng.random((n, 2)))
    prod = product_decisions(sc, rays32, light, uv)
    # the oracle on the same FP32 rays (promoted) and uniforms: the same decisions
    h32 = o.trace(rays32.astype(np.float64))
    both = (prod >= 0) & (h32["kind"] > 0)
    same = oracle_decisions(o, light[both], np.ascontiguousarray(h32["p"][both]), np.ascontiguousarray(uv[both]))
    clear = same != 2  # no other geometry in the way
    agree = (prod[both][clear] == np.isin(same[clear], (1, 3))).mean()
    assert agree > 0.999, agree
    # the rate on the oracle's own FP64 rays: what the reference's images carry
    h = o.trace(rays)
    ok = (h["kind"] > 0) & (prod >= 0)
    ref = oracle_decisions(o, light[ok], np.ascontiguousarray(h["p"][ok]), np.ascontiguousarray(uv[ok]))
    keep = ref != 2
    r_ref = np.isin(ref[keep], (1, 3)).mean()
    r_prod = prod[ok][keep].mean()
    m = keep.sum()
    se = np.sqrt(2 * r_ref * (1 - r_ref) / m)
    print(f"{name}: self-occluded {r_prod:.5f} (FP32 kernels' decision) vs {r_ref:.5f} (oracle), {m} samples, "
          f"z {(r_prod - r_ref) / se:+.2f}")
    assert 0.05 < r_ref < 0.95
    assert abs(r_prod - r_ref) < 4 * se, (r_prod, r_ref, se)


def test_shared_reciprocal_quotient_equals_division():
    """The decision normalises directions with one correctly rounded reciprocal and a residual
    correction per component (Markstein's theorem); its quotients must be those of the
    division, bit for bit, signed zeros included."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    v = rng.normal(size=(n, 3)) * 10.0 ** rng.uniform(-6, 6, size=(n, 1))
    v32 = v.astype(np.float32).astype(np.float64)  # FP32-derived directions, as the kernels' rays
    v = np.concatenate([v, v32])
    l = np.sqrt((v * v).sum(axis=1))
    other = 10.0 ** rng.uniform(-3, 3, size=len(v))  # divisors unrelated to the vector
    v = np.concatenate([v, v])
    l = np.concatenate([l, other])
    zeros = np.zeros((8, 3))
    zeros[1::2] = -0.0
    zeros[2:4, 0] = 0.5
    v = np.concatenate([v, zeros])
    l = np.concatenate([l, np.full(8, 1.7)])
    out = np.empty_like(v)
    rc = _lib().tpt_test_div3(C.c_int64(len(l)), np.ascontiguousarray(v).ctypes.data_as(C.POINTER(C.c_double)),
                              np.ascontiguousarray(l).ctypes.data_as(C.POINTER(C.c_double)),
                              out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    want = v / l[:, None]
    assert np.array_equal(out.view(np.uint64), want.view(np.uint64))
