"""torch.autograd bridge over the render and its path-replay adjoint.

    params = torch.tensor(base_table, dtype=torch.float64, requires_grad=True)   # (n_materials, 4)
    img = render_image(dscene, cfg, params)            # (h, w, 3) float64 = acc / acc_w
    loss = ((img - target) ** 2).mean()
    loss.backward()                                    # params.grad from tpt_render_adjoint

The forward pass uploads the material table (``tpt_scene_update_materials``) and renders with the
FP32 wavefront kernels; the backward pass replays paths from their seeds (``tpt_render_adjoint``).
By default it replays the very paths of the forward image (``decorrelate=False``): the gradient of
that image.  A loss that multiplies the image by itself (L2) gets a biased gradient that way,
because the image and its derivative share their Monte Carlo noise; ``decorrelate=True`` replays
an independent set of paths (seed XOR a fixed constant), which makes the L2 gradient unbiased.  Sampling pdfs and Russian-roulette
probabilities are held fixed, as in any detached-sampling path-replay estimator; the gradient is
unbiased for the expected image.  Tensors may live on the CPU or on the context's device; the
C ABI sees host or device pointers accordingly and torch only provides the memory.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import DeviceScene, RenderConfig, render, render_adjoint, render_device


DECORRELATE_SEED = 0x9E3779B97F4A7C15


class RenderFunction(torch.autograd.Function):
    """image = render(material table); d image / d table by path replay."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, dscene: DeviceScene, config: RenderConfig, decorrelate: bool = False):
        if config.adaptive_enabled:
            raise ValueError("the adjoint replays fixed-spp renders; disable adaptive sampling")
        table = params.detach().to("cpu", torch.float64).contiguous().numpy()
        if table.shape != (dscene.n_materials, 4):
            raise ValueError(f"params must have shape ({dscene.n_materials}, 4)")
        dscene.update_materials(table)
        h, w = int(config.height), int(config.width)
        if params.is_cuda:
            acc = torch.empty((h, w, 3), dtype=torch.float64, device=params.device)
            acc_w = torch.empty((h, w), dtype=torch.float64, device=params.device)
            # the library works on its own stream: let torch's pending work on this memory finish
            torch.cuda.current_stream(params.device).synchronize()
            render_device(dscene, config, acc, acc_w)
            img = acc / acc_w.clamp_min(1.0)[..., None]
        else:
            out = render(dscene, config, want_sq=False)
            img = torch.from_numpy(out.acc / np.maximum(out.acc_w, 1.0)[..., None])
        ctx.dscene, ctx.config = dscene, config
        ctx.decorrelate = decorrelate
        ctx.table = table
        ctx.device = params.device
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        dscene, config = ctx.dscene, ctx.config
        dscene.update_materials(ctx.table)  # the replay must see the forward pass's parameters
        g = grad_img.detach().to("cpu", torch.float32).contiguous().numpy()
        cfg = config
        if ctx.decorrelate:
            cfg = dataclasses.replace(config, seed=(int(config.seed) ^ DECORRELATE_SEED) & 0xFFFFFFFFFFFFFFFF)
        grad, _ = render_adjoint(dscene, cfg, g)
        return torch.from_numpy(grad).to(ctx.device), None, None, None


def render_image(dscene: DeviceScene, config: RenderConfig, params: torch.Tensor, decorrelate: bool = False) -> torch.Tensor:
    """Differentiable resolved image (acc / acc_w) of ``dscene`` with material table ``params``."""
    return RenderFunction.apply(params, dscene, config, decorrelate)
