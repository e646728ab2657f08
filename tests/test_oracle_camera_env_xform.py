"""Oracle camera (ray.rs:36-120), sky (integrator.rs:25-28), bilinear environment lookup
(env.rs:92-121) and transforms (transform.rs:24-164), pinned by closed forms and by the
reference's own transform tests.

The expected values are derived here from geometry (view angles, orthonormal bases, focus
points, linear functions that bilinear interpolation must reproduce), not from the oracle's
code, so a dropped term, a swapped axis or a wrong sign in the oracle fails them."""
import math

import numpy as np
import pytest

import oracle
from tinypt_amd import _abi
from tinypt_amd.scene import Camera, EnvMap, Lambert, Scene, Transform, World


def cam_desc(eye, target, up, vfov, aspect, focus=None, aperture=0.0):
    c = _abi.CameraDesc()
    c.eye[:] = eye
    c.target[:] = target
    c.up[:] = up
    c.vfov_deg = vfov
    c.aspect = aspect
    c.focus_dist = focus if focus is not None else float(np.linalg.norm(np.subtract(eye, target)))
    c.aperture = aperture
    return c


# ---------------------------------------------------------------- camera (ray.rs:36-120)
CAMS = [((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1.0),
        ((1.0, 2.0, -4.0), (0.5, -0.3, 2.0), (0.0, 1.0, 0.0), 65.0, 16 / 9),
        ((-3.0, 1.0, 2.0), (2.0, 0.5, -1.0), (0.2, 1.0, 0.1), 25.0, 0.5)]


@pytest.mark.parametrize("eye,target,up,vfov,aspect", CAMS)
def test_pinhole_centre_and_edge_rays_follow_the_view_angles(eye, target, up, vfov, aspect):
    eye, target, up = map(np.array, (eye, target, up))
    fwd = (target - eye) / np.linalg.norm(target - eye)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)  # u = norm(up x w), w = -fwd
    upv = np.cross(right, fwd)      # v = w x u
    th = math.tan(math.radians(vfov) / 2)
    sxy = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-0.3, 0.7], [0.55, -0.9]], float)
    r = oracle.camera_rays(cam_desc(eye, target, up, vfov, aspect), 7, sxy)
    np.testing.assert_allclose(r[:, :3], np.tile(eye, (len(sxy), 1)), rtol=0, atol=1e-15)  # pinhole: origin = eye
    d = r[:, 3:6]
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose(d[0], fwd, atol=1e-15)  # the centre ray looks at the target
    # the tangent of the angle off the axis, per screen axis: sx * aspect * tan(vfov/2) to the
    # right, sy * tan(vfov/2) up (sy = +1 is the top of the image)
    along = d @ fwd
    np.testing.assert_allclose((d @ right) / along, sxy[:, 0] * aspect * th, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose((d @ upv) / along, sxy[:, 1] * th, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("eye,target,up,vfov,aspect", CAMS)
def test_thin_lens_rays_meet_at_the_focus_point(eye, target, up, vfov, aspect):
    focus, aperture = 2.5, 0.6
    sxy = np.random.default_rng(3).uniform(-1, 1, (400, 2))
    pin = oracle.camera_rays(cam_desc(eye, target, up, vfov, aspect, focus, 0.0), 9, sxy)
    lens = oracle.camera_rays(cam_desc(eye, target, up, vfov, aspect, focus, aperture), 9, sxy)
    fp = pin[:, :3] + pin[:, 3:6] * focus  # the pinhole ray's point at the focus distance
    o, d = lens[:, :3], lens[:, 3:6]
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    # every lens ray passes through its pixel's focus point
    t = np.einsum("ij,ij->i", fp - o, d)
    np.testing.assert_allclose(o + d * t[:, None], fp, atol=1e-12)
    # origins lie on the lens disk: radius aperture / 2, perpendicular to the view axis
    off = o - np.asarray(eye, float)
    fwd = np.asarray(target, float) - np.asarray(eye, float)
    fwd /= np.linalg.norm(fwd)
    np.testing.assert_allclose(off @ fwd, 0.0, atol=1e-14)
    rr = np.linalg.norm(off, axis=1)
    assert rr.max() < aperture / 2
    # uniform on the disk: E[r^2] = R^2 / 2 (R = 0.3), std of the mean ~ R^2 / sqrt(12 n)
    assert abs((rr ** 2).mean() - 0.045) < 4 * 0.09 / math.sqrt(12 * len(rr))


def test_ray_time_is_one_draw_per_pinhole_ray():
    """time = shutter_open + (close - open) U = U with the default [0, 1) shutter; a pinhole
    camera draws nothing else (ray.rs:91-108)."""
    sxy = np.zeros((50, 2))
    r = oracle.camera_rays(cam_desc((0, 0, 3), (0, 0, 0), (0, 1, 0), 40, 1), 12345, sxy)
    np.testing.assert_array_equal(r[:, 6], oracle.rng_f64(12345, 50))


# ---------------------------------------------------------------- sky (integrator.rs:25-28)
def empty_scene(env=None, w=1.0):
    cam = Camera.look_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 90.0, w)
    return Scene(cam, World(), [Lambert((0.5, 0.5, 0.5))], env)


def test_sky_closed_form_values():
    """Straight up is the zenith colour (0.6, 0.75, 1), straight down the nadir colour
    (0.9, 0.95, 1), the horizon their mean, and the sky is linear in d.y in between."""
    o = oracle.OracleScene(empty_scene())
    ys = np.linspace(-1, 1, 41)
    dirs = np.stack([np.sqrt(1 - ys ** 2), ys, np.zeros_like(ys)], 1)
    rays = np.concatenate([np.zeros((len(ys), 3)), dirs, np.zeros((len(ys), 1))], 1)
    L = o.radiance(rays, np.arange(len(ys)) + 1, max_bounces=4, rr_start=3)
    np.testing.assert_allclose(L[-1], (0.6, 0.75, 1.0), rtol=0, atol=1e-15)
    np.testing.assert_allclose(L[0], (0.9, 0.95, 1.0), rtol=0, atol=1e-15)
    np.testing.assert_allclose(L[20], (0.75, 0.85, 1.0), rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.diff(L, 2, axis=0), 0.0, atol=1e-14)


def test_sky_image_rows_follow_the_view_directions():
    """A 4x4 render of the sky with a 90-degree camera: the expected pixel is the sky averaged
    over the pixel's jitter square (integrated on a fine grid here); the top row is the bluest
    (sy = 1 - (y + jy) * 2 / h, render.rs:156-157)."""
    w = h = 4
    spp = 256
    r = oracle.OracleScene(empty_scene()).render(w, h, spp, max_bounces=4, rr_start=3, seed=5)
    img = r["acc"] / spp
    th = math.tan(math.radians(45.0))
    g = (np.arange(64) + 0.5) / 64
    for y in range(h):
        for x in range(w):
            sx = (x + g[:, None]) * 2 / w - 1
            sy = 1 - (y + g[None, :]) * 2 / h
            d = np.stack(np.broadcast_arrays(sx * th, sy * th, -np.ones_like(sx * sy)), -1)
            dy = d[..., 1] / np.linalg.norm(d, axis=-1)
            t = 0.5 * (dy + 1)
            expect = np.array([(1 - t) * 0.9 + t * 0.6, (1 - t) * 0.95 + t * 0.75, np.ones_like(t)]).reshape(3, -1).mean(1)
            var = np.maximum(r["acc_sq"][y, x] / spp - img[y, x] ** 2, 0) / spp
            assert np.all(np.abs(img[y, x] - expect) <= 4 * np.sqrt(var) + 1e-12), (x, y, img[y, x], expect)
    assert img[0, :, 0].mean() < img[-1, :, 0].mean()  # red falls towards the zenith


# ---------------------------------------------------------------- environment lookup (env.rs:92-121)
def dirs_from(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], 1)


def test_bilinear_lookup_reproduces_a_vertical_ramp():
    """texel (x, y) = (y, 2y, 3y): bilinear interpolation reproduces the ramp exactly at
    y = (theta / pi) (H - 1), for any phi."""
    W, H = 16, 9
    data = np.zeros((H, W, 3))
    data[..., 0] = np.arange(H)[:, None]
    data[..., 1] = 2 * np.arange(H)[:, None]
    data[..., 2] = 3 * np.arange(H)[:, None]
    o = oracle.OracleScene(empty_scene(EnvMap(W, H, data)))
    rng = np.random.default_rng(4)
    theta = np.concatenate([[0.0, math.pi, math.pi / 2], rng.uniform(0, math.pi, 300)])
    phi = rng.uniform(0, 2 * math.pi, len(theta))
    rgb, _ = o.env_lookup(dirs_from(theta, phi))
    y = theta / math.pi * (H - 1)
    np.testing.assert_allclose(rgb, np.stack([y, 2 * y, 3 * y], 1), rtol=1e-12, atol=1e-12)


def test_bilinear_lookup_wraps_in_x():
    """texel (x, y) = x: x = (phi / 2 pi) W; between the last column and the first the lookup
    blends column W-1 into column 0 (x1 = (x0 + 1) mod W)."""
    W, H = 8, 5
    data = np.zeros((H, W, 3))
    data[..., :] = np.arange(W)[None, :, None]
    o = oracle.OracleScene(empty_scene(EnvMap(W, H, data)))
    phi = np.concatenate([np.linspace(0, 2 * math.pi, 97, endpoint=False), [0.01, 2 * math.pi - 1e-3]])
    theta = np.full(len(phi), 1.1)
    rgb, _ = o.env_lookup(dirs_from(theta, phi))
    x = np.mod(np.arctan2(np.sin(phi), np.cos(phi)), 2 * math.pi) / (2 * math.pi) * W
    expect = np.where(x <= W - 1, x, (W - 1) * (1 - (x - (W - 1))))
    np.testing.assert_allclose(rgb[:, 0], expect, rtol=1e-11, atol=1e-11)
    np.testing.assert_array_equal(rgb[:, 0], rgb[:, 1])
    # phi = 0 is the +x direction, phi = pi/2 the +z direction
    rgb2, _ = o.env_lookup(np.array([[1.0, 0, 0], [0, 0, 1.0]]))
    np.testing.assert_allclose(rgb2[:, 0], [0.0, W / 4], atol=1e-12)


# ---------------------------------------------------------------- transforms (transform.rs:123-164)
def close(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) < 1e-9


def xf12(x: Transform):
    return x.flat()


def test_legacy_new_matches_trs():  # transform.rs:123-127
    xf = Transform.new((1.0, 2.0, 3.0), 90.0, 2.0)
    assert close(oracle.transform_apply(xf12(xf), "apply_point", [1.0, 0.0, 0.0])[0], (1.0, 2.0, 1.0))


def test_rotate_about_x_axis():  # transform.rs:131-135
    xf = Transform.rotate((1.0, 0.0, 0.0), 90.0)
    assert close(oracle.transform_apply(xf12(xf), "apply_point", [0.0, 1.0, 0.0])[0], (0.0, 0.0, 1.0))


def test_nonuniform_scale_normal_uses_inverse_transpose():  # transform.rs:139-150
    xf = Transform.scale((2.0, 1.0, 1.0))
    assert close(oracle.transform_apply(xf12(xf), "apply_normal", [1.0, 0.0, 0.0])[0], (1.0, 0.0, 0.0))
    n2 = oracle.transform_apply(xf12(xf), "apply_normal", np.array([1.0, 1.0, 0.0]) / math.sqrt(2))[0]
    assert close(n2, np.array([0.5, 1.0, 0.0]) / np.linalg.norm([0.5, 1.0, 0.0]))


def test_inverse_roundtrips():  # transform.rs:154-160
    xf = Transform.rotate((0.3, 1.0, 0.5), 37.0).compose(Transform.scale((2.0, 0.5, 1.5)))
    p = np.array([1.0, -2.0, 3.0])
    q = oracle.transform_apply(xf12(xf), "apply_point", p)
    assert close(oracle.transform_apply(xf12(xf), "apply_point_inv", q)[0], p)


def test_compose_applies_inner_first():  # transform.rs:164-172
    xf = Transform.identity().compose(Transform.translate((0.0, 1.0, 0.0))).compose(Transform.scale((2.0, 2.0, 2.0)))
    assert close(oracle.transform_apply(xf12(xf), "apply_point", [1.0, 0.0, 0.0])[0], (2.0, 1.0, 0.0))


def test_vec_inv_ignores_translation_and_normal_is_unit():
    rng = np.random.default_rng(8)
    a = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    t = rng.normal(size=3)
    xf = list(a.reshape(-1)) + list(t)
    v = rng.normal(size=(50, 3))
    np.testing.assert_allclose(oracle.transform_apply(xf, "apply_vec_inv", v), np.linalg.solve(a, v.T).T, atol=1e-12)
    n = oracle.transform_apply(xf, "apply_normal", v)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)
    # a normal stays perpendicular to transformed tangents: (A^-T n) . (A s) = n . s
    s = np.cross(v, rng.normal(size=(50, 3)))
    s -= (np.einsum("ij,ij->i", s, v) / np.einsum("ij,ij->i", v, v))[:, None] * v
    np.testing.assert_allclose(np.einsum("ij,ij->i", n, (a @ s.T).T), 0.0, atol=1e-11)
