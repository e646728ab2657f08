"""Scene files and checkpoint files against the oracle, end to end on the GPU.

* The reference's sample scenes (tests/golden/scenes: cornell.xml, default.xml, mesh.xml with
  cube.obj) loaded through tpt_scene_load_xml (mitsuba.rs:137-222), rendered by both GPU forms
  at a reduced size of the same aspect, against the oracle's render of the loaded scene.
* Checkpoint interchange (checkpoint.rs:48-138, render.rs:48-70): a HYPCKPT file written by the
  oracle's restatement after k of the image's tile tasks is resumed by the plain FP64 GPU form;
  the finished image keeps the file's pixels bit for bit and matches the oracle's uninterrupted
  render on the rest (the plain form runs the reference's per-pixel streams), and the final
  file the GPU render leaves is read back by the oracle's loader.
"""
import os

import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import output as O
from tinypt_amd import scenes
from tinypt_amd.loader import load_scene

from parity_scenes import compare, oracle_threads

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")


@pytest.mark.parametrize("name,w,h,spp", [("cornell.xml", 48, 48, 256), ("default.xml", 64, 36, 128),
                                          ("mesh.xml", 64, 36, 128)])
def test_loaded_sample_scene_matches_oracle(gpu_ctx, name, w, h, spp):
    cfg = T.RenderConfig()
    warns = []
    sc = load_scene(os.path.join(SCENES, name), cfg, warns)
    assert abs(cfg.width / cfg.height - w / h) < 1e-9  # the camera's aspect is the file's film
    mb, rr = cfg.max_bounces, cfg.rr_start
    ds = T.DeviceScene(sc, gpu_ctx)
    ref = oracle.OracleScene(sc).render(w, h, spp, max_bounces=mb, rr_start=rr, seed=9, threads=oracle_threads())
    for prec, label in ((T.FP32_WAVEFRONT, "fp32"), (T.FP64_PLAIN, "plain fp64")):
        out = T.render(ds, T.RenderConfig(width=w, height=h, spp=spp, max_bounces=mb, rr_start=rr, seed=9,
                                          precision=prec))
        assert (out.acc_w == spp).all()
        z, zm, ma, mb_, _ = compare(out.acc, out.acc_sq, out.acc_w, ref["acc"], ref["acc_sq"], ref["acc_w"])
        frac = float((np.abs(z) > 3).mean())
        extra = ""
        if prec == T.FP64_PLAIN:
            extra = f", bit-identical pixels {float((out.acc == ref['acc']).all(-1).mean()):.3f}"
        print(f"{name} {w}x{h} {spp} spp (depth {mb}, rr {rr}) {label}: mean {ma:.6f} vs oracle {mb_:.6f}; "
              f"|z|>3 fraction {frac:.4f}; image-mean z {np.round(zm, 2)}{extra}; warnings {len(warns)}", flush=True)
        assert frac <= 0.012, frac
        assert np.all(np.abs(zm) < 4.0), zm


def test_oracle_checkpoint_resumed_by_gpu(gpu_ctx, tmp_path):
    w, h, spp, tile, k = 48, 40, 32, 16, 4
    sc = scenes.cornell_box(w, h)
    cfg = T.RenderConfig(width=w, height=h, spp=spp, max_bounces=8, rr_start=3, seed=21, tile=tile,
                         precision=T.FP64_PLAIN, checkpoint_enabled=True, checkpoint_every_tasks=2,
                         scene_hash=0x4859503000C0FFEE)
    ref = oracle.OracleScene(sc).render(w, h, spp, max_bounces=8, rr_start=3, seed=21, tile=tile,
                                        threads=oracle_threads())
    tasks = oracle.task_tiles(w, h, tile)
    assert len(tasks) == 9
    done = np.zeros((h, w), bool)
    for x0, y0, x1, y1 in tasks[:k]:
        done[y0:y1, x0:x1] = True
    ck = str(tmp_path / O.ckpt_path(cfg.scene_hash))
    oracle.checkpoint_save(ck, cfg.scene_hash, k, np.where(done[..., None], ref["acc"], 0.0),
                           np.where(done, ref["acc_w"], 0.0))
    out = T.render(T.DeviceScene(sc, gpu_ctx), cfg, ckpt_file=ck)
    # the file's tasks are kept as written; the GPU rendered the others
    assert np.array_equal(out.acc[done], ref["acc"][done]) and np.array_equal(out.acc_w[done], ref["acc_w"][done])
    assert out.stats["paths"] == int((~done).sum()) * spp
    assert (out.acc_w == spp).all()
    rest = ~done
    same = float((out.acc[rest] == ref["acc"][rest]).all(-1).mean())
    close = float((np.abs(out.acc[rest] - ref["acc"][rest]) <= 1e-9 * np.maximum(np.abs(ref["acc"][rest]), 1e-3))
                  .all(-1).mean())
    print(f"resumed from an oracle checkpoint after {k} of {len(tasks)} tasks: rendered pixels bit-identical to the "
          f"oracle's {same:.3f}, within 1e-9 {close:.3f}", flush=True)
    assert close > 0.2, close
    # the render leaves a final checkpoint (render.rs:267-272) that the oracle's loader reads
    got = oracle.checkpoint_load(ck, cfg.scene_hash, w, h)
    assert got is not None and got[0] == len(tasks)
    assert np.array_equal(got[1], out.acc) and np.array_equal(got[2], out.acc_w)
    assert oracle.checkpoint_load(ck, cfg.scene_hash ^ 1, w, h) is None
