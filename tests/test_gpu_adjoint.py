"""Path-replay adjoint (tpt_render_adjoint) against central finite differences.

With Russian roulette off (rr_start = max_bounces) the FP32 estimator is, for a fixed seed, a
smooth function of every albedo / F0 / absorption / emission parameter: none of them moves a
sampled direction or a branch.  Its central difference with common random numbers (h = 1e-3)
is then the exact derivative of the same realisation up to O(h^2), and the adjoint must match
it within 1%.  GGX roughness moves the sampled directions; the adjoint takes its gradient as a
per-path central difference with the same seeds, which sums to the render's own central
difference, so it is held to the same 1% (with Russian roulette on as well).  Agreement in
expectation with the finite differences of the plain FP64 render and of the oracle, roughness
included, is tested on the inverse-rendering problem in test_gpu_inverse_gradients.py.
"""
import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import _abi, scenes

from parity_scenes import material_zoo

pytestmark = pytest.mark.gpu

H = 1e-3


def _setup(gpu_ctx, sc, w, h):
    ds = T.DeviceScene(sc, gpu_ctx)
    base = np.array([m.param for m in sc.mats], dtype=np.float64)
    rng = np.random.default_rng(17)
    dl = rng.uniform(0.0, 1.0, (h, w, 3)).astype(np.float32)
    return ds, base, dl


def _loss(ds, cfg, dl, params):
    ds.update_materials(params)
    out = T.render(ds, cfg, want_sq=False)
    return float((dl.astype(np.float64) * (out.acc / cfg.spp)).sum())


def _fd(ds, cfg, dl, base, i, k):
    p = base.copy()
    p[i, k] += H
    lp = _loss(ds, cfg, dl, p)
    p[i, k] -= 2 * H
    lm = _loss(ds, cfg, dl, p)
    ds.update_materials(base)
    return (lp - lm) / (2 * H)


def _params(sc):
    """(material, param) pairs: every rgb channel (albedo / F0 / absorption / emission) and the
    roughness of every GGX material."""
    out = [(i, k) for i in range(len(sc.mats)) for k in range(3)]
    return out + [(i, 3) for i, m in enumerate(sc.mats) if m.kind == _abi.MAT_GGX]


# Paths of up to 9 bounces are summed and differentiated in one replay (their BSDF vertices are
# logged until the path total is known); deeper ones in two ("material_zoo_deep").
@pytest.mark.parametrize("name,rr,mb", [("material_zoo", 6, 6), ("cornell", 6, 6), ("material_zoo_rr", 2, 6),
                                        ("material_zoo_deep", 12, 12)])
def test_adjoint_matches_finite_differences(gpu_ctx, name, rr, mb):
    if name == "cornell":
        sc, w, h = scenes.cornell_box(32, 32), 32, 32
    else:
        sc, w, h = material_zoo(), 32, 32
    ds, base, dl = _setup(gpu_ctx, sc, w, h)
    cfg = T.RenderConfig(width=w, height=h, spp=64, max_bounces=mb, rr_start=rr, seed=21)
    if rr < mb:  # with Russian roulette only roughness is compared (its estimator is FD-consistent)
        pairs = [(i, 3) for i, m in enumerate(sc.mats) if m.kind == _abi.MAT_GGX]
    else:
        pairs = _params(sc)
    g, st = T.render_adjoint(ds, cfg, dl)
    assert st["paths"] == w * h * 64
    fd = np.array([_fd(ds, cfg, dl, base, i, k) for i, k in pairs])
    ad = np.array([g[i, k] for i, k in pairs])
    scale = np.abs(fd).max()
    for (i, k), a, f in zip(pairs, ad, fd):
        print(f"{name} mat {i} (kind {sc.mats[i].kind}) p{k}: adjoint {a:.6g} fd {f:.6g}", flush=True)
    sig = np.abs(fd) > 1e-3 * scale
    rel = np.abs(ad - fd)[sig] / np.abs(fd)[sig]
    assert sig.sum() >= min(6, len(pairs))
    assert rel.max() < 0.01, rel.max()
    # parameters with no influence get (close to) no gradient
    assert np.all(np.abs(ad[~sig]) <= 2e-3 * scale)


def test_adjoint_on_two_level_scene(gpu_ctx):
    """64 rotated instances: rays enter instances and the replay's traversal stack runs past its
    LDS part into the overflow area (the deepest instance trees need more than 16 entries)."""
    from parity_scenes import instanced_small
    sc = instanced_small()
    w = h = 24
    ds, base, dl = _setup(gpu_ctx, sc, w, h)
    cfg = T.RenderConfig(width=w, height=h, spp=32, max_bounces=4, rr_start=4, seed=8)
    g, st = T.render_adjoint(ds, cfg, dl)
    assert st["paths"] == w * h * 32
    pairs = [(i, k) for i in range(len(sc.mats)) for k in range(3)
             if sc.mats[i].kind in (_abi.MAT_LAMBERT, _abi.MAT_GGX, _abi.MAT_SUBSURFACE)][:12]
    fd = np.array([_fd(ds, cfg, dl, base, i, k) for i, k in pairs])
    ad = np.array([g[i, k] for i, k in pairs])
    for (i, k), a, f in zip(pairs, ad, fd):
        print(f"instanced mat {i} (kind {sc.mats[i].kind}) p{k}: adjoint {a:.6g} fd {f:.6g}", flush=True)
    sig = np.abs(fd) > 1e-3 * np.abs(fd).max()
    assert sig.sum() >= 4
    assert (np.abs(ad - fd)[sig] / np.abs(fd)[sig]).max() < 0.01


def test_adjoint_accumulates_into_grad(gpu_ctx):
    sc = scenes.cornell_box(16, 16)
    ds, base, dl = _setup(gpu_ctx, sc, 16, 16)
    cfg = T.RenderConfig(width=16, height=16, spp=8, max_bounces=4, rr_start=2, seed=3)
    g1, _ = T.render_adjoint(ds, cfg, dl)
    g2 = np.ones_like(g1)
    T.render_adjoint(ds, cfg, dl, g2)
    np.testing.assert_allclose(g2 - 1.0, g1, rtol=1e-5, atol=1e-9)


def test_adjoint_against_fp64_plain_and_oracle_fd(gpu_ctx):
    """The adjoint (FP32 path replay) against central differences (h = 1e-3) of the plain FP64
    GPU render and of the FP64 oracle -- the reference algorithm on the CPU, so the reference
    gradient -- in expectation over independent seeds: the three estimate the same gradient of
    the expected loss, each FD per seed with common random numbers.  Russian roulette is off, so
    no albedo / F0 / absorption / emission parameter moves a branch."""
    import dataclasses

    import oracle
    from parity_scenes import oracle_threads

    sc = material_zoo()
    w = h = 12
    spp = 128
    ds, base, dl = _setup(gpu_ctx, sc, w, h)
    kinds = [m.kind for m in sc.mats]
    light = kinds.index(_abi.MAT_DIFFUSE_LIGHT)
    pairs = [(kinds.index(_abi.MAT_LAMBERT), k) for k in range(3)] + \
            [(kinds.index(_abi.MAT_GGX), k) for k in range(3)] + \
            [(kinds.index(_abi.MAT_DIELECTRIC), k) for k in range(3)] + [(light, k) for k in range(3)]

    def loss_oracle(params, seed):
        mats = [T.Material(m.kind, tuple(float(x) for x in params[i])) for i, m in enumerate(sc.mats)]
        r = oracle.OracleScene(dataclasses.replace(sc, mats=mats)).render(
            w, h, spp, max_bounces=6, rr_start=6, seed=seed, threads=oracle_threads())
        return float((dl.astype(np.float64) * (r["acc"] / spp)).sum())

    ad, fd64, fdref = [], [], []
    for s in range(6):
        cfg = T.RenderConfig(width=w, height=h, spp=spp, max_bounces=6, rr_start=6, seed=1000 + s)
        g, _ = T.render_adjoint(ds, cfg, dl)
        ad.append([g[i, k] for i, k in pairs])
        cfg64 = T.RenderConfig(width=w, height=h, spp=spp, max_bounces=6, rr_start=6, seed=2000 + s,
                               precision=T.FP64_PLAIN)
        fd64.append([_fd(ds, cfg64, dl, base, i, k) for i, k in pairs])
        row = []
        for i, k in pairs:
            p = base.copy()
            p[i, k] += H
            lp = loss_oracle(p, 2000 + s)
            p[i, k] -= 2 * H
            row.append((lp - loss_oracle(p, 2000 + s)) / (2 * H))
        fdref.append(row)
    ad, fd64, fdref = np.array(ad), np.array(fd64), np.array(fdref)
    n = len(ad)

    def z(a, b):
        se = np.sqrt(a.var(0, ddof=1) / n + b.var(0, ddof=1) / n)
        return (a.mean(0) - b.mean(0)) / (se + 1e-12 * np.abs(b.mean(0)).max())

    for (i, k), a, f, r in zip(pairs, ad.mean(0), fd64.mean(0), fdref.mean(0)):
        print(f"mat {i} p{k}: adjoint {a:.6g}  fd fp64 plain {f:.6g}  fd oracle {r:.6g}", flush=True)
    # the plain FP64 form runs the reference's per-pixel streams: seed for seed its FD is the
    # oracle's except for the paths that a last-bit libm difference moves
    rel = np.abs(fd64 - fdref) / np.abs(fdref).max(0)
    print(f"plain FP64 vs oracle FD, seed for seed: median rel diff {np.median(rel):.2e}, max {rel.max():.2e}", flush=True)
    print(f"z adjoint vs oracle {np.round(z(ad, fdref), 2)}", flush=True)
    print(f"z plain vs oracle {np.round(z(fd64, fdref), 2)}", flush=True)
    assert np.median(rel) < 1e-2
    assert np.all(np.abs(z(fd64, fdref)) < 4.0)
    assert np.all(np.abs(z(ad, fdref)) < 4.0), z(ad, fdref)
    # the comparison resolves each parameter: its standard error is a small part of its gradient
    se = np.sqrt(ad.var(0, ddof=1) / n + fdref.var(0, ddof=1) / n)
    assert np.all(se < 0.1 * np.abs(fdref.mean(0))), se / np.abs(fdref.mean(0))
