"""The two heavy workloads at their full sizes, each scene built once per module.

Displaced icosphere (1,310,720 triangles, ground, 2048x1024 environment, 1920x1080): closest
hits of 100k render rays against the oracle, and a 32-spp frame against the oracle on a centre
crop.  4096 instances of four icosphere variants (5.2M unique triangles, 1920x1080): closest
hits of render rays through the centre, a 16-spp frame against the oracle on a centre crop, and
the adaptive configuration itself -- 64 to 4096 spp, threshold 0.01 relative standard error,
depth 8 -- over the whole frame, held to the sampler's invariants and, on an 8x8 centre crop,
to the oracle's adaptive render of the same pixels (render.rs:143-179).  The oracle's per-ray
loop over 4096 instances makes a full frame hours of CPU time, hence the crops.
"""
import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import scenes

from parity_scenes import compare, oracle_threads
from trace_checks import classify, path_rays, report

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
Z_FRAC_MAX = 0.012


class Heavy:
    def __init__(self, sc, gpu_ctx):
        self.sc = sc
        self.orc = oracle.OracleScene(sc)
        self.ds = T.DeviceScene(sc, gpu_ctx)


@pytest.fixture(scope="module")
def mesh(gpu_ctx):
    print("building the 1.31M-triangle mesh scene (host tables, oracle, device) ...", flush=True)
    return Heavy(scenes.mesh_scene(W, H), gpu_ctx)


@pytest.fixture(scope="module")
def instanced(gpu_ctx):
    print("building the 4096-instance scene (host tables, oracle, device) ...", flush=True)
    return Heavy(scenes.instanced_scene(W, H), gpu_ctx)


def _trace(name, hv, rays):
    rep = classify(hv.sc, rays, hv.ds.trace_closest(rays), threads=oracle_threads(), orc=hv.orc)
    report(name, rep)
    assert rep["agree"] >= 0.9999, rep["agree"]
    assert rep["t_rel_p999"] < 1e-5, rep["t_rel_p999"]


def _crop_parity(name, hv, spp, mb, rr, seed, crop, **kw):
    cfg = T.RenderConfig(width=W, height=H, spp=spp, max_bounces=mb, rr_start=rr, seed=seed, **kw)
    out = T.render(hv.ds, cfg)
    assert np.isfinite(out.acc).all() and np.isfinite(out.acc_sq).all()
    x0, y0, x1, y1 = crop
    ref = hv.orc.render(W, H, spp, max_bounces=mb, rr_start=rr, seed=seed, threads=oracle_threads(), crop=crop,
                        adaptive=bool(kw.get("adaptive_enabled")), adaptive_min_spp=kw.get("adaptive_min_spp", 8),
                        adaptive_threshold=kw.get("adaptive_threshold", 0.02))
    sl = (slice(y0, y1), slice(x0, x1))
    z, zm, ma, mb_, _ = compare(out.acc[sl], out.acc_sq[sl], out.acc_w[sl], ref["acc"][sl], ref["acc_sq"][sl],
                                ref["acc_w"][sl])
    frac = float((np.abs(z) > 3).mean())
    print(f"{name}: mean {ma:.6f} vs oracle {mb_:.6f}; |z|>3 fraction {frac:.4f}; image-mean z {np.round(zm, 2)}",
          flush=True)
    assert frac <= Z_FRAC_MAX, frac
    assert np.all(np.abs(zm) < 4.0), zm
    return out, ref, sl


def test_mesh_trace(mesh):
    rng = np.random.default_rng(13)
    _trace("mesh 1.31M path rays", mesh, path_rays(mesh.sc, rng, 100_000, threads=oracle_threads(), orc=mesh.orc))


def test_mesh_frame(mesh):
    out, _, _ = _crop_parity("mesh 1920x1080 32 spp (64x64 centre crop)", mesh, 32, 16, 3, 3, (928, 508, 992, 572))
    assert (out.acc_w == 32).all()
    print(f"mesh 1920x1080 32 spp: {out.stats}", flush=True)


def test_instanced_trace(instanced):
    rng = np.random.default_rng(17)
    rays = path_rays(instanced.sc, rng, 60_000, threads=oracle_threads(), crop=(640, 300, 1280, 780, W, H),
                     orc=instanced.orc)
    _trace("instanced 4096 path rays", instanced, rays)


def test_instanced_frame(instanced):
    out, _, _ = _crop_parity("instanced 1920x1080 16 spp (16x16 centre crop)", instanced, 16, 8, 3, 3,
                             (952, 532, 968, 548))
    assert (out.acc_w == 16).all()


def test_instanced_adaptive_full_config(instanced):
    """Adaptive 64..4096 spp, threshold 0.01, depth 8, rr 3 over the full frame."""
    out, ref, sl = _crop_parity("instanced 1920x1080 adaptive 64..4096 thr 0.01 (8x8 centre crop)", instanced, 4096,
                                8, 3, 3, (956, 536, 964, 544), adaptive_enabled=True, adaptive_min_spp=64,
                                adaptive_threshold=0.01)
    w = out.acc_w
    assert np.all(w == np.round(w)) and w.min() >= 64 and w.max() <= 4096, (w.min(), w.max())
    assert int(w.sum()) == out.stats["paths"], (int(w.sum()), out.stats["paths"])
    mg, mo = float(w[sl].mean()), float(ref["acc_w"][sl].mean())
    print(f"adaptive: frame mean spp {w.mean():.1f} (min {w.min():.0f}, max {w.max():.0f}), "
          f"{out.stats['paths']} paths in {out.stats['elapsed_ms'] / 1e3:.2f} s; crop mean spp {mg:.1f} vs oracle {mo:.1f}",
          flush=True)
    assert abs(mg - mo) <= 0.05 * mo, (mg, mo)
