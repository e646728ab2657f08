"""Tile batches (tpt_render_pixels) and the checkpointed render loop on the GPU.

A pixel's samples are seeded per pixel (and per sample in the FP32 form), so any subset of
pixels rendered in one call reproduces the full render's sums for those pixels: bit for bit
with the plain FP64 form and with the FP32 form at a fixed samples_per_pass (the FP64 sums are
then grouped alike); an adaptive FP32 render keeps the same samples (equal acc_w) and differs
only in FP64 summation order.  A render interrupted after a checkpoint and resumed equals the
uninterrupted one (render.rs:48-70, 248-272).
"""
import os

import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import output as O
from tinypt_amd import scenes

pytestmark = pytest.mark.gpu

W, H, SPP = 40, 24, 12


class Interrupted(RuntimeError):
    pass


def _cfg(**kw):
    base = dict(width=W, height=H, spp=SPP, max_bounces=6, rr_start=3, seed=5, samples_per_pass=4, tile=8)
    base.update(kw)
    return T.RenderConfig(**base)


@pytest.fixture(scope="module")
def cornell(gpu_ctx):
    return T.DeviceScene(scenes.cornell_box(W, H), gpu_ctx)


@pytest.mark.parametrize("prec", [T.FP32_WAVEFRONT, T.FP64_PLAIN])
def test_render_pixels_matches_full_render(cornell, prec):
    cfg = _cfg(precision=prec)
    print(f"render_pixels subset, precision {prec}", flush=True)
    full = T.render(cornell, cfg)
    rng = np.random.default_rng(1)
    pix = rng.permutation(W * H)[:301].astype(np.int32)
    s, sw, sq, st = T.render_pixels(cornell, cfg, pix, want_sq=True)
    assert st["paths"] == 301 * SPP
    assert np.array_equal(s, full.acc.reshape(-1, 3)[pix])
    assert np.array_equal(sw, full.acc_w.reshape(-1)[pix])
    assert np.array_equal(sq, full.acc_sq.reshape(-1, 3)[pix])


@pytest.mark.parametrize("prec", [T.FP32_WAVEFRONT, T.FP64_PLAIN])
def test_render_pixels_adaptive(cornell, prec):
    cfg = _cfg(precision=prec, adaptive_enabled=True, adaptive_min_spp=4, adaptive_threshold=0.3, spp=24)
    print(f"render_pixels adaptive, precision {prec}", flush=True)
    full = T.render(cornell, cfg)
    pix = np.arange(0, W * H, 7, dtype=np.int32)
    s, sw, _, _ = T.render_pixels(cornell, cfg, pix)
    assert np.array_equal(sw, full.acc_w.reshape(-1)[pix])
    assert (sw < 24).any() and (sw >= 4).all()
    if prec == T.FP64_PLAIN:
        assert np.array_equal(s, full.acc.reshape(-1, 3)[pix])
    else:
        assert np.allclose(s, full.acc.reshape(-1, 3)[pix], rtol=1e-12, atol=1e-12)


def test_render_pixels_checks_indices(cornell):
    cfg = _cfg()
    for bad in ([-1], [W * H], [0, 5, W * H + 3]):
        with pytest.raises(T.TptError):
            T.render_pixels(cornell, cfg, np.array(bad, np.int32))
    s, sw, _, st = T.render_pixels(cornell, cfg, np.zeros(0, np.int32))
    assert s.shape == (0, 3) and st["paths"] == 0


@pytest.mark.parametrize("prec", [T.FP32_WAVEFRONT, T.FP64_PLAIN])
def test_checkpointed_render_equals_full(cornell, tmp_path, prec):
    cfg = _cfg(precision=prec)
    print(f"checkpointed render, precision {prec}", flush=True)
    full = T.render(cornell, cfg)
    ck = str(tmp_path / O.ckpt_path(cfg.scene_hash))
    cfg.checkpoint_enabled = True
    cfg.checkpoint_every_tasks = 4
    out = T.render(cornell, cfg, ckpt_file=ck)
    assert np.array_equal(out.acc, full.acc) and np.array_equal(out.acc_w, full.acc_w)
    assert np.array_equal(out.acc_sq, full.acc_sq)
    n_tasks = len(O.tile_tasks(W, H, 8))
    nid, a, w = O.load_checkpoint(ck, cfg.scene_hash, W, H)
    assert nid == n_tasks and np.array_equal(a, full.acc) and np.array_equal(w, full.acc_w)


@pytest.mark.parametrize("prec", [T.FP32_WAVEFRONT, T.FP64_PLAIN])
def test_resume_after_interruption(cornell, tmp_path, monkeypatch, prec):
    cfg = _cfg(precision=prec)
    print(f"interrupted render, precision {prec}", flush=True)
    full = T.render(cornell, cfg)
    ck = str(tmp_path / "ck.bin")
    cfg.checkpoint_enabled = True
    cfg.checkpoint_every_tasks = 3
    monkeypatch.setattr(T, "_CKPT_BATCH_PATHS", 1)  # one checkpoint interval per GPU call
    real = T.render_pixels
    calls = []

    def flaky(*a, **k):
        calls.append(1)
        if len(calls) == 3:
            raise Interrupted("stopped after two batches")
        return real(*a, **k)

    monkeypatch.setattr(T, "render_pixels", flaky)
    with pytest.raises(Interrupted):
        T.render(cornell, cfg, ckpt_file=ck)
    nid, a, w = O.load_checkpoint(ck, cfg.scene_hash, W, H)
    assert nid == 6
    done = O.task_pixels(O.tile_tasks(W, H, 8)[:6], W)
    assert np.array_equal(w.reshape(-1)[done], np.full(done.size, float(SPP)))
    assert w.sum() == done.size * SPP
    monkeypatch.setattr(T, "render_pixels", real)
    out = T.render(cornell, cfg, ckpt_file=ck)
    assert out.acc_sq is None  # resumed: the file holds no squares
    assert np.array_equal(out.acc, full.acc) and np.array_equal(out.acc_w, full.acc_w)
    assert out.stats["paths"] == (W * H - done.size) * SPP


def test_foreign_checkpoint_is_ignored(cornell, tmp_path):
    cfg = _cfg()
    full = T.render(cornell, cfg)
    ck = str(tmp_path / "ck.bin")
    O.save_checkpoint(ck, cfg.scene_hash ^ 1, 5, np.ones((H, W, 3)), np.ones((H, W)))
    cfg.checkpoint_enabled = True
    cfg.checkpoint_every_tasks = 8
    out = T.render(cornell, cfg, ckpt_file=ck)
    assert np.array_equal(out.acc, full.acc)
    assert O.load_checkpoint(ck, cfg.scene_hash, W, H)[0] == len(O.tile_tasks(W, H, 8))


def test_write_rendered_image(cornell, tmp_path):
    cfg = _cfg()
    out = T.render(cornell, cfg)
    for name in ("c.ppm", "c.hdr", "c.exr"):
        p = str(tmp_path / name)
        O.write_image(p, out.acc, out.acc_w, exposure=cfg.exposure, tonemap=cfg.tonemap)
        assert os.path.getsize(p) > 0


def test_progress_render_equals_full(cornell, capsys):
    cfg = _cfg()
    full = T.render(cornell, cfg)
    out = T.render(cornell, cfg, progress=True)
    err = capsys.readouterr().err
    assert "Progress: 15/15 tiles (100.0%)" in err and "Done in" in err
    assert np.array_equal(out.acc, full.acc) and np.array_equal(out.acc_w, full.acc_w)
