"""Oracle BSDFs (material.rs) -- the reference's own tests (material.rs:362-475) restated, plus
closed-form checks the reference does not make (GGX reciprocity/normalisation, dielectric
branches, subsurface)."""
import numpy as np
import pytest

import oracle
from tinypt_amd import _abi

FLOOR_RAY = [0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 0.0]          # material.rs:354-358
FLOOR_HIT = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]       # t, p, n (+Y)
N = np.array([0.0, 1.0, 0.0])
WO = np.array([0.0, 1.0, 0.0])


def test_lambert_sample_pdf_matches_eval():  # material.rs:362-373
    s = oracle.material_sample(_abi.MAT_LAMBERT, [0.6, 0.4, 0.2, 0], FLOOR_RAY, FLOOR_HIT, 1, 1000)
    assert (s[:, 0] == 1).all()
    for row in s:
        _, pdf = oracle.material_eval(_abi.MAT_LAMBERT, [0.6, 0.4, 0.2, 0], WO, row[1:4], N)
        assert abs(row[7] - pdf) < 1e-9


def test_ggx_sample_pdf_matches_eval():  # material.rs:377-389
    s = oracle.material_sample(_abi.MAT_GGX, [0.9, 0.9, 0.9, 0.3], FLOOR_RAY, FLOOR_HIT, 7, 1000)
    for row in s[s[:, 0] == 1]:
        _, pdf = oracle.material_eval(_abi.MAT_GGX, [0.9, 0.9, 0.9, 0.3], WO, row[1:4], N)
        assert abs(row[7] - pdf) < 1e-9


def test_ggx_weight_matches_f_cos_over_pdf():  # material.rs:394-411
    p = [0.9, 0.8, 0.7, 0.35]
    s = oracle.material_sample(_abi.MAT_GGX, p, FLOOR_RAY, FLOOR_HIT, 11, 2000)
    for row in s[s[:, 0] == 1]:
        f, pdf = oracle.material_eval(_abi.MAT_GGX, p, WO, row[1:4], N)
        cos_o = max(N @ row[1:4], 0.0)
        np.testing.assert_allclose(row[4:7], f * (cos_o / max(pdf, 1e-6)), rtol=0, atol=1e-9)


def test_lambert_weight_is_albedo():  # material.rs:415-424
    s = oracle.material_sample(_abi.MAT_LAMBERT, [0.5, 0.7, 0.3, 0], FLOOR_RAY, FLOOR_HIT, 3, 1)
    np.testing.assert_allclose(s[0, 4:7], [0.5, 0.7, 0.3], atol=1e-12)


def test_ggx_eval_is_reciprocal():  # material.rs:428-438
    p = [0.8, 0.5, 0.2, 0.25]
    wo = np.array([0.4, 1.0, 0.1]) / np.linalg.norm([0.4, 1.0, 0.1])
    wi = np.array([-0.3, 1.0, 0.2]) / np.linalg.norm([-0.3, 1.0, 0.2])
    f1, _ = oracle.material_eval(_abi.MAT_GGX, p, wo, wi, N)
    f2, _ = oracle.material_eval(_abi.MAT_GGX, p, wi, wo, N)
    np.testing.assert_allclose(f1, f2, atol=1e-12)


def _uniform_hemisphere(seed, n):
    u = oracle.rng_f64(seed, 2 * n).reshape(n, 2)
    z = u[:, 0]
    r = np.sqrt(np.maximum(1 - z * z, 0))
    phi = 2 * np.pi * u[:, 1]
    return np.stack([r * np.cos(phi), z, r * np.sin(phi)], 1)


def test_lambert_pdf_integrates_to_one():  # material.rs:443-462
    wi = _uniform_hemisphere(99, 200000)
    pdf = np.array([oracle.material_eval(_abi.MAT_LAMBERT, [1, 1, 1, 0], WO, w, N)[1] for w in wi[:20000]])
    assert abs(2 * np.pi * pdf.mean() - 1.0) < 0.02


def py_vndf_as_written(wo, alpha, u1, u2):
    """sample_ggx_vndf exactly as material.rs:327-347 writes it.  Note the t2 blend
    sqrt(1 - s) * t2 + s * sqrt(1 - t1^2) differs from Heitz 2018's (1 - s) * sqrt(1 - t1^2) + s * t2;
    the reference's estimator is kept as is."""
    nrm = lambda v: v / max(np.linalg.norm(v), 1e-30)
    v = nrm(np.array([alpha * wo[0], alpha * wo[1], wo[2]]))
    if v[2] < 0.9999:
        t1 = nrm(np.array([-v[1], v[0], 0.0]))
        t2 = np.cross(v, t1)
    else:
        t1, t2 = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    t1p, t2p0 = r * np.cos(phi), r * np.sin(phi)
    s = 0.5 * (1 + v[2])
    t2p = np.sqrt(1 - s) * t2p0 + s * np.sqrt(max(1 - t1p * t1p, 0))
    nh = nrm(t1 * t1p + t2 * t2p + v * np.sqrt(max(1 - t1p * t1p - t2p * t2p, 0)))
    return nrm(np.array([alpha * nh[0], alpha * nh[1], nh[2]]))


def test_ggx_vndf_matches_independent_restatement():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        t = rng.uniform(0, 1.55)
        ph = rng.uniform(0, 2 * np.pi)
        wo = np.array([np.sin(t) * np.cos(ph), np.sin(t) * np.sin(ph), np.cos(t)])
        a, u1, u2 = rng.uniform(0.001, 1), rng.random(), rng.random()
        np.testing.assert_allclose(oracle.ggx_vndf(wo, a, u1, u2), py_vndf_as_written(wo, a, u1, u2), atol=1e-13)


def test_ggx_normal_incidence_rejects_every_sample():
    """With wo == n the blend puts every sampled half vector in the tangent plane
    (s = 1 -> t2p = sqrt(1 - t1p^2)), the reflected direction points below the surface and
    Material::sample returns None (material.rs:159-161): a reference quirk the GPU keeps."""
    s = oracle.material_sample(_abi.MAT_GGX, [1, 1, 1, 0.3], FLOOR_RAY, FLOOR_HIT, 5, 2000)
    assert (s[:, 0] == 0).all()
    wo = np.array([np.sin(1.0), np.cos(1.0), 0.0])
    s = oracle.material_sample(_abi.MAT_GGX, [1, 1, 1, 0.3], [0, 1, 0, *(-wo), 0], FLOOR_HIT, 5, 2000)
    assert 0.2 < s[:, 0].mean() < 0.9


def test_ggx_vndf_sample_is_unit_and_upper():
    rng = np.random.default_rng(1)
    for _ in range(200):
        t = rng.uniform(0, 1.5)
        wo = np.array([np.sin(t), 0.0, np.cos(t)])
        m = oracle.ggx_vndf(wo, rng.uniform(0.01, 1), rng.random(), rng.random())
        assert abs(np.linalg.norm(m) - 1) < 1e-12 and m[2] >= 0


def test_classification():  # material.rs:466-475 via sample(): delta flag and emission
    s = oracle.material_sample(_abi.MAT_METAL, [1, 1, 1, 0], FLOOR_RAY, FLOOR_HIT, 1, 1)
    assert s[0, 8] == 1 and s[0, 7] == 0
    np.testing.assert_allclose(s[0, 1:4], [0, 1, 0])  # mirror
    s = oracle.material_sample(_abi.MAT_DIELECTRIC, [0, 0, 0, 1.5], FLOOR_RAY, FLOOR_HIT, 1, 1)
    assert s[0, 8] == 1
    s = oracle.material_sample(_abi.MAT_DIFFUSE_LIGHT, [3, 3, 3, 0], FLOOR_RAY, FLOOR_HIT, 1, 1)
    assert s[0, 0] == 0  # emitters do not scatter
    s = oracle.material_sample(_abi.MAT_LAMBERT, [1, 1, 1, 0], FLOOR_RAY, FLOOR_HIT, 1, 1)
    assert s[0, 8] == 0


def test_dielectric_branches_and_fresnel_frequency():
    # normal incidence on ior 1.5: Schlick F = 0.04; reflect with probability F, weight 1
    s = oracle.material_sample(_abi.MAT_DIELECTRIC, [0, 0, 0, 1.5], FLOOR_RAY, FLOOR_HIT, 9, 20000)
    refl = s[:, 2] > 0
    assert abs(refl.mean() - 0.04) < 0.006
    np.testing.assert_allclose(s[refl, 4:7], np.ones((refl.sum(), 3)))
    # transmission scales by eta^2 = (1/1.5)^2 on entering
    np.testing.assert_allclose(s[~refl, 4:7], np.full(((~refl).sum(), 3), (1 / 1.5) ** 2), rtol=1e-12)
    # exiting (ray travels along +n): Beer-Lambert exp(-sigma t) applied, eta = 1.5
    ray_out = [0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    hit = [2.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    s = oracle.material_sample(_abi.MAT_DIELECTRIC, [0.1, 0.2, 0.3, 1.5], ray_out, hit, 9, 2000)
    tr = s[:, 2] > 0
    np.testing.assert_allclose(s[tr, 4:7], np.broadcast_to(np.exp(-np.array([0.1, 0.2, 0.3]) * 2.0) * 1.5 ** 2, (tr.sum(), 3)), rtol=1e-12)


def test_total_internal_reflection_draws_nothing():
    # grazing exit from glass: refraction impossible, reflection chosen without a random draw
    d = np.array([0.9, 0.2, 0.0]) / np.linalg.norm([0.9, 0.2, 0.0])
    ray = [0.0, -1.0, 0.0, *d, 0.0]
    hit = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    a = oracle.material_sample(_abi.MAT_DIELECTRIC, [0, 0, 0, 1.5], ray, hit, 3, 2)
    np.testing.assert_allclose(a[0, 1:4], a[1, 1:4])  # both draws identical: no RNG consumed
    np.testing.assert_allclose(a[0, 1:4], [d[0], -d[1], 0], atol=1e-12)


def test_subsurface_offsets_inside_and_weights():
    s = oracle.material_sample(_abi.MAT_SUBSURFACE, [0.8, 0.6, 0.4, 0.1], FLOOR_RAY, FLOOR_HIT, 4, 5000)
    assert (s[:, 0] == 1).all() and (s[:, 8] == 0).all()
    depth = -s[:, 10] + 1e-4 * s[:, 2]  # origin.y = -dist + eps * d.y
    assert (depth > 0).all()
    # E[exp(-dist/s)] with dist ~ Exp(mean s) is 1/2
    ratio = s[:, 4] / 0.8
    assert abs(ratio.mean() - 0.5) < 0.01
    np.testing.assert_allclose(s[:, 5] / 0.6, ratio, rtol=1e-12)
