"""Gradients of the inverse-rendering problem (Cornell box + 16 GGX spheres, F0 rgb and alpha
each: 64 parameters; 256x256, depth 8) against central finite differences.

The loss is the L2 image loss linearised at the initial parameters: dLoss/dpixel =
2 (I0 - target) / N with I0 a 1024-spp render at the initial parameters, so every estimator
below estimates the same gradient of the expected image.  Russian roulette is off (its
probability depends on the throughput, so the finite differences of a render with it jump
wherever a path's survival flips: their variance per seed was 10-100x larger).

* The adjoint (FP32 path replay) equals the central difference of the FP32 render with the same
  seeds: the exact derivative of the same realisation for F0, and for alpha the same per-path
  difference summed (the reference's VNDF sampler moves its samples with alpha).
* In expectation over independent seeds, the adjoint equals the central difference of the plain
  FP64 GPU render (h = 1e-3 for F0; 1e-2 for alpha, whose per-path differences jump with the
  sampled directions and have a variance near 1/h): z within 4 on every parameter, with enough
  seeds that the standard error of the comparison is below 1% of each F0 gradient.
* Seed for seed, the plain FP64 form's finite differences are the oracle's (the reference
  algorithm on the CPU, run on the same per-pixel random streams).
"""
import dataclasses

import numpy as np
import pytest
import torch

import oracle
import tinypt_amd as T
from tinypt_amd import _abi, resolve_pixels
from tinypt_amd.inverse import InverseProblem

from parity_scenes import oracle_threads

pytestmark = pytest.mark.gpu

MB = 8
RR = 8  # Russian roulette off
H_F0, H_ALPHA = 1e-3, 1e-2


class Problem:
    def __init__(self, gpu_ctx):
        self.prob = InverseProblem(ctx=gpu_ctx)
        self.ds = self.prob.dscene
        self.base = self.prob.table(torch.from_numpy(self.prob.init)).numpy()
        self.ds.update_materials(self.base)
        w, h = self.prob.width, self.prob.height
        self.w, self.h = w, h
        ref = T.render(self.ds, self.cfg(1024, 0xBEEF), want_sq=False)
        i0 = resolve_pixels(ref.acc, ref.acc_w)
        self.dl = (2.0 * (i0 - self.prob.target.numpy()) / i0.size).astype(np.float32)
        # parameters: per sphere the F0 channel with the largest gradient, for the six spheres
        # with the largest such gradients, and the two largest alpha gradients
        g, _ = T.render_adjoint(self.ds, self.cfg(4096, 77), self.dl)
        ids = self.prob.mat_ids
        f0 = sorted(((abs(g[m, :3]).max(), m, int(abs(g[m, :3]).argmax())) for m in ids), reverse=True)[:6]
        al = sorted(((abs(g[m, 3]), m) for m in ids), reverse=True)[:2]
        self.pairs = [(m, k) for _, m, k in f0] + [(m, 3) for _, m in al]

    def cfg(self, spp, seed, precision=T.FP32_WAVEFRONT):
        return T.RenderConfig(width=self.w, height=self.h, spp=spp, max_bounces=MB, rr_start=RR, seed=seed,
                              precision=precision)

    def loss(self, cfg, params):
        self.ds.update_materials(params)
        out = T.render(self.ds, cfg, want_sq=False)
        return float((self.dl.astype(np.float64) * (out.acc / cfg.spp)).sum())

    def fd(self, cfg, m, k, h):
        p = self.base.copy()
        p[m, k] += h
        lp = self.loss(cfg, p)
        p[m, k] -= 2 * h
        lm = self.loss(cfg, p)
        self.ds.update_materials(self.base)
        return (lp - lm) / (2 * h)


@pytest.fixture(scope="module")
def problem(gpu_ctx):
    return Problem(gpu_ctx)


def _h(k):
    return H_ALPHA if k == 3 else H_F0


def test_adjoint_is_the_derivative_of_its_render(problem):
    P = problem
    for seed in (5, 6):
        cfg = P.cfg(64, seed)
        P.ds.update_materials(P.base)
        g, st = T.render_adjoint(P.ds, cfg, P.dl)
        # pixels whose loss derivative is zero (the initial render equals the target there: black
        # pixels) are not replayed
        assert 0.8 * P.w * P.h * 64 <= st["paths"] <= P.w * P.h * 64
        for m, k in P.pairs:
            f = P.fd(cfg, m, k, 1e-3)  # the adjoint's own alpha step
            print(f"seed {seed} mat {m} p{k}: adjoint {g[m, k]:+.6e}  fd32 (same seeds) {f:+.6e}", flush=True)
            assert abs(g[m, k] - f) <= 1e-3 * abs(f) + 1e-12, (m, k, g[m, k], f)


def test_adjoint_matches_fp64_plain_fd_in_expectation(problem):
    P = problem
    n, spp = 36, 256
    ad, fd = [], []
    for s in range(n):
        P.ds.update_materials(P.base)
        g, _ = T.render_adjoint(P.ds, P.cfg(spp, 1000 + s), P.dl)
        ad.append([g[m, k] for m, k in P.pairs])
        fd.append([P.fd(P.cfg(spp, 5000 + s, T.FP64_PLAIN), m, k, _h(k)) for m, k in P.pairs])
    ad, fd = np.array(ad), np.array(fd)
    se = np.sqrt(ad.var(0, ddof=1) / n + fd.var(0, ddof=1) / n)
    z = (ad.mean(0) - fd.mean(0)) / se
    for j, (m, k) in enumerate(P.pairs):
        print(f"mat {m} p{k}: adjoint {ad[:, j].mean():+.5e}  fd fp64 plain {fd[:, j].mean():+.5e}  rel diff "
              f"{(ad[:, j].mean() - fd[:, j].mean()) / abs(fd[:, j].mean()):+.4f}  se/|g| {se[j] / abs(fd[:, j].mean()):.4f}"
              f"  z {z[j]:+.2f}", flush=True)
    assert np.all(np.abs(z) < 4.0), z
    f0 = np.array([k < 3 for _, k in P.pairs])
    assert np.all(se[f0] < 0.01 * np.abs(fd.mean(0)[f0])), se[f0] / np.abs(fd.mean(0)[f0])


def test_fp64_plain_fd_is_the_oracle_fd(problem):
    P = problem
    spp = 16
    mats = P.prob.scene.mats

    def loss_oracle(params, seed):
        ms = [T.Material(mt.kind, tuple(float(x) for x in params[i])) for i, mt in enumerate(mats)]
        sc = dataclasses.replace(P.prob.scene, mats=ms)
        r = oracle.OracleScene(sc).render(P.w, P.h, spp, max_bounces=MB, rr_start=RR, seed=seed,
                                          threads=oracle_threads())
        return float((P.dl.astype(np.float64) * (r["acc"] / spp)).sum())

    rel = []
    for seed in (31, 32):
        for m, k in P.pairs:
            h = _h(k)
            f64 = P.fd(P.cfg(spp, seed, T.FP64_PLAIN), m, k, h)
            p = P.base.copy()
            p[m, k] += h
            lp = loss_oracle(p, seed)
            p[m, k] -= 2 * h
            fo = (lp - loss_oracle(p, seed)) / (2 * h)
            rel.append(abs(f64 - fo) / max(abs(fo), 1e-30))
            print(f"seed {seed} mat {m} p{k}: fd fp64 plain {f64:+.6e}  fd oracle {fo:+.6e}", flush=True)
    print(f"plain FP64 vs oracle FD, seed for seed: median rel diff {np.median(rel):.2e}, max {max(rel):.2e}",
          flush=True)
    assert np.median(rel) < 1e-3, np.median(rel)
