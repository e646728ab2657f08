"""Inverse rendering of material parameters with the path-replay adjoint.

The problem: a Cornell box holding 16 GGX spheres, each with F0 rgb and roughness alpha (64
unknowns).  A target image is rendered at high spp from the true parameters; the optimiser starts
from parameters drawn U[0.2, 0.8] and minimises the L2 image loss with Adam, one forward render
(fresh seed per iteration) and one adjoint replay per step.

    prob = InverseProblem()
    hist = optimize(prob, steps=200)
    hist["s_per_iter"], hist["param_error"][-1]
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import DeviceScene, RenderConfig, render, resolve_pixels
from .autograd import render_image
from .scenes import inverse_material_ids, inverse_params, inverse_scene


@dataclass
class InverseProblem:
    width: int = 256
    height: int = 256
    n_spheres: int = 16
    init_seed: int = 5
    truth_seed: int = 1005
    target_spp: int = 1024
    spp: int = 16
    max_bounces: int = 8
    rr_start: int = 3
    radius: float = 0.12
    with_cubes: bool = True
    learn_alpha: bool = True
    ctx: object = None
    scene: object = field(init=False)
    dscene: DeviceScene = field(init=False)
    truth: np.ndarray = field(init=False)
    init: np.ndarray = field(init=False)
    mat_ids: list = field(init=False)
    base: np.ndarray = field(init=False)
    target: torch.Tensor = field(init=False)
    target_seconds: float = field(init=False)

    def __post_init__(self):
        self.truth = inverse_params(self.truth_seed, self.n_spheres)
        self.init = inverse_params(self.init_seed, self.n_spheres)
        if not self.learn_alpha:  # roughness known: start from the true values
            self.init[:, 3] = self.truth[:, 3]
        self.scene = inverse_scene(self.truth, self.width, self.height, self.n_spheres, self.radius, self.with_cubes)
        self.dscene = DeviceScene(self.scene, self.ctx)
        self.mat_ids = inverse_material_ids(self.n_spheres, self.with_cubes)
        self.base = np.array([m.param for m in self.scene.mats], dtype=np.float64)
        t0 = time.perf_counter()
        out = render(self.dscene, self.config(self.target_spp, seed=0x7A26E7), want_sq=False)
        self.target_seconds = time.perf_counter() - t0
        self.target = torch.from_numpy(resolve_pixels(out.acc, out.acc_w))

    def config(self, spp: int, seed: int) -> RenderConfig:
        return RenderConfig(width=self.width, height=self.height, spp=spp, max_bounces=self.max_bounces,
                            rr_start=self.rr_start, seed=seed)

    def table(self, sphere_params: torch.Tensor) -> torch.Tensor:
        """Full material table with the spheres' rows replaced (differentiable in sphere_params)."""
        t = torch.from_numpy(self.base.copy())
        idx = torch.tensor(self.mat_ids, dtype=torch.long)
        return t.index_put((idx,), sphere_params)


def image_error(prob: InverseProblem, params: np.ndarray, spp: int = 512, seed: int = 0x51CE) -> float:
    """Mean squared difference between the target and a render of the sphere parameters at
    ``spp`` (the two images' own Monte Carlo variance included: the value at the true parameters
    is the floor).  The 16-spp training loss is dominated by that variance; this measures the
    fit."""
    prob.dscene.update_materials(prob.table(torch.as_tensor(np.asarray(params, dtype=np.float64))).numpy())
    out = render(prob.dscene, prob.config(spp, seed), want_sq=False)
    prob.dscene.update_materials(prob.base)
    return float(((resolve_pixels(out.acc, out.acc_w) - prob.target.numpy()) ** 2).mean())


def optimize(prob: InverseProblem, steps: int = 200, lr: float = 0.02, seed0: int = 1, log=None,
             eval_spp: int = 0) -> dict:
    """Adam on the L2 image loss.  Returns per-step loss and mean absolute parameter error, and
    the mean wall time of a step (forward render + adjoint + update).  With ``eval_spp`` > 0 also
    the image error (``image_error``) of the initial, final and true parameters."""
    x = torch.tensor(prob.init, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([x], lr=lr)
    losses, errs, f0_errs, alpha_errs, times = [], [], [], [], []
    truth = torch.from_numpy(prob.truth)
    for it in range(steps):
        t0 = time.perf_counter()
        # independent paths for the adjoint: the L2 gradient is then unbiased
        img = render_image(prob.dscene, prob.config(prob.spp, seed0 + it), prob.table(x), decorrelate=True)
        loss = ((img - prob.target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        if not prob.learn_alpha:
            x.grad[:, 3] = 0.0
        opt.step()
        with torch.no_grad():
            x.clamp_(0.01, 1.0)
        times.append(time.perf_counter() - t0)
        losses.append(float(loss.detach()))
        err = (x.detach() - truth).abs()
        errs.append(float(err.mean()))
        f0_errs.append(float(err[:, :3].mean()))
        alpha_errs.append(float(err[:, 3].mean()))
        if log and (it % 20 == 0 or it == steps - 1):
            log(f"step {it}: loss {losses[-1]:.6g} param error {errs[-1]:.4f} ({times[-1] * 1e3:.1f} ms)")
    prob.dscene.update_materials(prob.base)
    init_err = np.abs(prob.init - prob.truth)
    img = {}
    if eval_spp > 0:
        img = {"image_error_init": image_error(prob, prob.init, eval_spp),
               "image_error_final": image_error(prob, x.detach().numpy(), eval_spp),
               "image_error_truth": image_error(prob, prob.truth, eval_spp), "image_error_spp": eval_spp}
    return {**img, "loss": losses, "param_error": errs, "f0_error": f0_errs, "alpha_error": alpha_errs,
            "s_per_iter": float(np.mean(times[1:] if steps > 1 else times)), "params": x.detach().numpy().copy(),
            "init_error": float(init_err.mean()), "init_f0_error": float(init_err[:, :3].mean()),
            "init_alpha_error": float(init_err[:, 3].mean())}
