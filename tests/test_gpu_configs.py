"""The benchmark configurations at their full sizes.

Plain default scene (256x256, 64 spp, depth 8, rr 3, seed 1): the plain FP64 form and the FP32
form against the oracle over the whole image.  Cornell box 512x512 (depth 12, rr 4): the FP32
form against the oracle at 16 spp over the whole image.  The mesh and 4096-instance workloads are in
test_gpu_heavy_scenes.py.
"""
import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import scenes

from parity_scenes import compare, oracle_threads

pytestmark = pytest.mark.gpu

Z_FRAC_MAX = 0.012


def _report(name, z, zm, ma, mb):
    frac = float((np.abs(z) > 3).mean())
    print(f"{name}: mean {ma:.6f} vs oracle {mb:.6f}; |z|>3 fraction {frac:.4f}; image-mean z {zm}", flush=True)
    assert frac <= Z_FRAC_MAX, frac
    assert np.all(np.abs(zm) < 4.0), zm


def test_default_scene_full_size(gpu_ctx):
    sc = scenes.default_scene(256, 256, "black")
    ds = T.DeviceScene(sc, gpu_ctx)
    kw = dict(width=256, height=256, spp=64, max_bounces=8, rr_start=3, seed=1)
    ref = oracle.OracleScene(sc).render(256, 256, 64, max_bounces=8, rr_start=3, seed=1, threads=oracle_threads())
    for prec, name in ((T.FP64_PLAIN, "plain fp64"), (T.FP32_WAVEFRONT, "fp32")):
        out = T.render(ds, T.RenderConfig(precision=prec, **kw))
        assert (out.acc_w == 64).all()
        _report(f"default 256x256 {name}", *compare(out.acc, out.acc_sq, out.acc_w, ref["acc"], ref["acc_sq"],
                                                      ref["acc_w"])[:4])
        if prec == T.FP64_PLAIN:
            same = float((out.acc == ref["acc"]).all(-1).mean())
            print(f"default 256x256 plain: bit-identical pixels {same:.4f}", flush=True)


def test_cornell_full_size(gpu_ctx):
    sc = scenes.cornell_box(512, 512)
    out = T.render(T.DeviceScene(sc, gpu_ctx), T.RenderConfig(width=512, height=512, spp=16, max_bounces=12, rr_start=4,
                                                               seed=2))
    ref = oracle.OracleScene(sc).render(512, 512, 16, max_bounces=12, rr_start=4, seed=2, threads=oracle_threads())
    _report("cornell 512x512 fp32", *compare(out.acc, out.acc_sq, out.acc_w, ref["acc"], ref["acc_sq"], ref["acc_w"])[:4])
