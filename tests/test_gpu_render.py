"""GPU parity of the render path against the FP64 oracle.

FP32 wavefront: per-pixel means agree with the oracle's within 3 sigma (the fraction of channel
z scores above 3 stays near its null rate) and the image means agree within 4 sigma.  The plain
FP64 kernel follows the reference's per-pixel random stream, so most pixels are bit-identical to
the oracle; the rest diverge where a libm last-bit difference flips a branch, and are checked
statistically like the FP32 form.
"""
import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import scenes
from tinypt_amd.scene import EnvMap, Scene, World, Camera

from parity_scenes import PARITY_CASES, compare, lambert_sphere_env, oracle_threads

pytestmark = pytest.mark.gpu

Z_FRAC_MAX = 0.012  # share of |z| > 3 channels allowed (null rate 0.0027 for Gaussian means)
Z_MEAN_MAX = 4.0


def _oracle_render(sc, w, h, spp, mb, rr, seed, **kw):
    return oracle.OracleScene(sc).render(w, h, spp, max_bounces=mb, rr_start=rr, seed=seed, threads=oracle_threads(),
                                         **kw)


def _check(name, out, ref, n_out, n_ref):
    z, zm, m_out, m_ref, npx = compare(out.acc, out.acc_sq, n_out, ref["acc"], ref["acc_sq"], n_ref)
    frac = float((np.abs(z) > 3).mean())
    print(f"{name}: mean {m_out:.6f} vs oracle {m_ref:.6f}; |z|>3 fraction {frac:.4f}; image-mean z {zm}", flush=True)
    assert frac <= Z_FRAC_MAX, (name, frac)
    assert np.all(np.abs(zm) < Z_MEAN_MAX), (name, zm)


@pytest.mark.parametrize("case", sorted(PARITY_CASES))
def test_fp32_wavefront_matches_oracle(gpu_ctx, case):
    build, w, h, spp, mb, rr, seed = PARITY_CASES[case]
    sc = build()
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=w, height=h, spp=spp, max_bounces=mb, rr_start=rr,
                                                               seed=seed))
    ref = _oracle_render(sc, w, h, spp, mb, rr, seed)
    assert (out.acc_w == spp).all()
    assert out.stats["paths"] == w * h * spp
    _check("fp32 " + case, out, ref, out.acc_w, ref["acc_w"])


@pytest.mark.parametrize("case", ["cornell", "default_black", "floorlight", "material_zoo", "mesh_small"])
def test_fp64_plain_matches_oracle(gpu_ctx, case):
    build, w, h, spp, mb, rr, seed = PARITY_CASES[case]
    sc = build()
    spp = min(spp, 128)
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=w, height=h, spp=spp, max_bounces=mb, rr_start=rr,
                                                               seed=seed, precision=T.FP64_PLAIN))
    ref = _oracle_render(sc, w, h, spp, mb, rr, seed)
    assert (out.acc_w == spp).all()
    same = float((out.acc == ref["acc"]).all(-1).mean())
    close = float((np.abs(out.acc - ref["acc"]) <= 1e-9 * np.maximum(np.abs(ref["acc"]), 1e-3)).all(-1).mean())
    print(f"fp64 plain {case}: bit-identical pixels {same:.4f}, within 1e-9 relative {close:.4f}", flush=True)
    assert out.stats["paths"] == ref["counters"]["paths"]
    _check("fp64 " + case, out, ref, out.acc_w, ref["acc_w"])
    # diagnostic: libm last-bit differences (OCML vs glibc) move some paths (a flipped RR
    # decision, a different edge hit); a wrong random stream would leave almost no pixel equal
    assert close > 0.1, close


def test_fp64_plain_bit_identical_on_floorlight(gpu_ctx):
    """No transcendental on this scene's paths except sqrt, so every pixel is bit-identical."""
    build, w, h, spp, mb, rr, seed = PARITY_CASES["floorlight"]
    sc = build()
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=w, height=h, spp=64, max_bounces=2, rr_start=8,
                                                               seed=seed, precision=T.FP64_PLAIN))
    ref = _oracle_render(sc, w, h, 64, 2, 8, seed)
    same = float((out.acc == ref["acc"]).all(-1).mean())
    print(f"floorlight plain bit-identical pixels {same:.4f}", flush=True)
    assert same > 0.95, same
    assert np.array_equal(out.acc_w, ref["acc_w"])


def test_constant_env_no_geometry_is_exact(gpu_ctx):
    """Closed form (integrator.rs:71-83): a bounce-0 miss returns E, no MIS, for every sample."""
    e = (0.25, 0.5, 1.5)
    sc = Scene(Camera.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), 60, 1), World(), [T.Lambert((0.5, 0.5, 0.5))],
               EnvMap.constant(e))
    for prec in (T.FP32_WAVEFRONT, T.FP64_PLAIN):
        out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=7, height=5, spp=16, seed=3, precision=prec))
        img = T.resolve_pixels(out.acc, out.acc_w)
        np.testing.assert_allclose(img, np.broadcast_to(np.array(e), img.shape), rtol=1e-6)


def test_firefly_clamp_on_bright_env(gpu_ctx):
    """lum(E) = 100 > 50: the camera-ray miss is clamped to luminance 50 (integrator.rs:80)."""
    sc = Scene(Camera.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), 60, 1), World(), [T.Lambert((0.5, 0.5, 0.5))],
               EnvMap.constant((100.0, 100.0, 100.0)))
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=4, height=4, spp=4, seed=1))
    img = T.resolve_pixels(out.acc, out.acc_w)
    np.testing.assert_allclose(img, 50.0, rtol=1e-6)


def test_lambert_sphere_in_constant_env(gpu_ctx):
    """Closed form: sphere pixels converge to albedo * E (convex sphere, NEE-env + MIS)."""
    sc = lambert_sphere_env(0.6, 1.0)
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=16, height=16, spp=2048, max_bounces=2,
                                                               rr_start=8, seed=5))
    img = T.resolve_pixels(out.acc, out.acc_w)
    centre = img[6:10, 6:10]
    np.testing.assert_allclose(centre.mean(), 0.6, rtol=0.01)


def test_black_world_renders_zero(gpu_ctx):
    """No emitter and a black environment: every sample is exactly zero."""
    b = scenes.SceneBuilder()
    b.sphere((0, 0, -3), 1.0, T.Lambert((0.8, 0.8, 0.8)))
    sc = b.build(Camera.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), 60, 1), EnvMap.constant((0.0, 0.0, 0.0)))
    for prec in (T.FP32_WAVEFRONT, T.FP64_PLAIN):
        out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=9, height=3, spp=8, seed=2, precision=prec))
        assert not out.acc.any()
        assert (out.acc_w == 8).all()


@pytest.mark.parametrize("case", ["cornell", "mesh_small"])
@pytest.mark.parametrize("mb,rr", [(0, 3), (1, 0), (2, 1)])
def test_short_paths_match_oracle(gpu_ctx, case, mb, rr):
    """Paths of at most two bounces, with Russian roulette from bounce rr (integrator.rs:53-95).
    max_bounces 0 traces nothing and gathers nothing (every pixel exactly zero, acc_w = spp).
    With one bounce the first bounce's kernels, which generate the camera rays themselves,
    carry the whole path: emission seen from the camera, NEE at the first hit (environment NEE
    in the mesh scene), no scattered ray."""
    build, w, h, *_ = PARITY_CASES[case]
    sc = build()
    spp = 256
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=w, height=h, spp=spp, max_bounces=mb, rr_start=rr,
                                                               seed=13))
    assert (out.acc_w == spp).all()
    if mb == 0:
        assert not out.acc.any() and out.stats["rays_extension"] == 0
        return
    ref = _oracle_render(sc, w, h, spp, mb, rr, 13)
    _check(f"{case} max_bounces {mb} rr {rr}", out, ref, out.acc_w, ref["acc_w"])


@pytest.mark.parametrize("wh", [(1, 1), (37, 23), (3, 64)])
def test_ragged_sizes_fp32(gpu_ctx, wh):
    w, h = wh
    sc = scenes.cornell_box(w, h)
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=w, height=h, spp=64, max_bounces=12, rr_start=4,
                                                               seed=2))
    ref = _oracle_render(sc, w, h, 64, 12, 4, 2)
    assert (out.acc_w == 64).all()
    z, zm, *_ = compare(out.acc, out.acc_sq, out.acc_w, ref["acc"], ref["acc_sq"], ref["acc_w"])
    assert np.all(np.abs(zm) < Z_MEAN_MAX), zm


def test_samples_per_pass_does_not_change_the_image(gpu_ctx):
    """Each (pixel, sample) path has its own seed, so the grouping of samples into passes only
    changes the FP64 summation grouping."""
    sc = scenes.cornell_box(24, 24)
    ds = T.DeviceScene(sc, gpu_ctx)
    outs = [T.render(ds, T.RenderConfig(width=24, height=24, spp=96, max_bounces=12, rr_start=4, seed=9,
                                        samples_per_pass=k)) for k in (0, 1, 7, 96)]
    for o in outs[1:]:
        np.testing.assert_allclose(o.acc, outs[0].acc, rtol=1e-12, atol=1e-12)
        assert o.stats["rays_extension"] == outs[0].stats["rays_extension"]


@pytest.mark.parametrize("name", ["cornell", "material_zoo", "mesh_small", "instanced_small"])
def test_traversal_refill_does_not_change_the_image(gpu_ctx, name):
    """The persistent refilling traversal kernels find the same hits as the one-ray-per-lane
    kernels: every threshold, and the scene-shape default, give bit-identical sums."""
    build, w, h, spp, mb, rr, seed = PARITY_CASES[name]
    ds = T.DeviceScene(build(), gpu_ctx)
    kw = dict(width=w, height=h, spp=spp // 4, max_bounces=mb, rr_start=rr, seed=seed)
    outs = [T.render(ds, T.RenderConfig(traversal_refill=r, **kw)) for r in (-1, 0, 1, 16, 48, 63, 200)]
    for o in outs[1:]:
        assert np.array_equal(o.acc, outs[0].acc) and np.array_equal(o.acc_sq, outs[0].acc_sq)
        for k in ("rays_extension", "rays_shadow", "light_samples", "light_self_occluded"):
            assert o.stats[k] == outs[0].stats[k], k


def test_seed_changes_the_image(gpu_ctx):
    sc = scenes.cornell_box(16, 16)
    ds = T.DeviceScene(sc, gpu_ctx)
    a = T.render(ds, T.RenderConfig(width=16, height=16, spp=8, seed=1))
    b = T.render(ds, T.RenderConfig(width=16, height=16, spp=8, seed=2))
    c = T.render(ds, T.RenderConfig(width=16, height=16, spp=8, seed=1))
    assert not np.array_equal(a.acc, b.acc)
    assert np.array_equal(a.acc, c.acc)


@pytest.mark.parametrize("prec", [T.FP32_WAVEFRONT, T.FP64_PLAIN])
def test_adaptive_matches_oracle(gpu_ctx, prec):
    """render.rs:143-179: Welford on luminance, stop once rel-std < threshold after min_spp."""
    sc = scenes.cornell_box(32, 32)
    kw = dict(width=32, height=32, spp=256, max_bounces=12, rr_start=4, seed=2, adaptive_enabled=True,
              adaptive_min_spp=16, adaptive_threshold=0.08)
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(precision=prec, **kw))
    ref = _oracle_render(sc, 32, 32, 256, 12, 4, 2, adaptive=True, adaptive_min_spp=16, adaptive_threshold=0.08)
    assert out.acc_w.min() >= 16 and out.acc_w.max() <= 256
    n_out, n_ref = out.acc_w.mean(), ref["acc_w"].mean()
    print(f"adaptive prec {prec}: mean spp {n_out:.2f} vs oracle {n_ref:.2f}; "
          f"stopped early {(out.acc_w < 256).mean():.3f} vs {(ref['acc_w'] < 256).mean():.3f}", flush=True)
    assert abs(n_out - n_ref) / n_ref < 0.05
    assert out.stats["paths"] == int(out.acc_w.sum())
    img_o = T.resolve_pixels(out.acc, out.acc_w)
    img_r = T.resolve_pixels(ref["acc"], ref["acc_w"])
    assert abs(img_o.mean() - img_r.mean()) / img_r.mean() < 0.02
    if prec == T.FP64_PLAIN:
        same = float((out.acc_w == ref["acc_w"]).mean())
        print(f"adaptive plain: identical per-pixel sample counts {same:.4f}", flush=True)
        assert same > 0.3


def test_full_size_cornell_fp32_vs_fp64_plain(gpu_ctx):
    """The benchmark's scene at its full 512x512 size: the FP32 and FP64 GPU forms agree on the
    image mean (4 sigma), every pixel took its samples, and the counters add up."""
    sc = scenes.cornell_box(512, 512)
    ds = T.DeviceScene(sc, gpu_ctx)
    kw = dict(width=512, height=512, spp=16, max_bounces=12, rr_start=4, seed=2)
    a = T.render(ds, T.RenderConfig(**kw))
    b = T.render(ds, T.RenderConfig(precision=T.FP64_PLAIN, **kw))
    z, zm, ma, mb, _ = compare(a.acc, a.acc_sq, a.acc_w, b.acc, b.acc_sq, b.acc_w)
    print(f"full-size cornell: fp32 mean {ma:.6f} fp64 plain {mb:.6f} image z {zm}", flush=True)
    assert np.all(np.abs(zm) < Z_MEAN_MAX)
    assert (a.acc_w == 16).all() and (b.acc_w == 16).all()
    assert a.stats["paths"] == b.stats["paths"] == 512 * 512 * 16
    assert float((np.abs(z) > 3).mean()) < Z_FRAC_MAX
