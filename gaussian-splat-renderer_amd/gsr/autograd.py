"""torch.autograd bridge of the differentiable render (gsr_backward): SplatImage.apply(means,
log_scales, quats, opacity_logits, sh, view) -> (H, W, 3) float32 image, whose backward is the
library's render and preprocess backward on the GPU. No counterpart in the reference (it renders
forward only).

The parameters and gradients cross through host memory (the tensors may live on any torch
device): the library and torch each bring a HIP runtime with the same soname, and a process
initialises the GPU through the first one it loads, so the bridge does not ask torch for device
memory. The image and the gradients are float32 tensors on the parameters' device."""
from dataclasses import dataclass

import numpy as np
import torch

import gsr


@dataclass
class View:
    """What a render needs besides the parameters: the library context, the camera and the
    renderer settings (bounds, sort mode, list budget)."""

    ctx: gsr.Context
    camera: gsr.Camera
    sh_degree: int = 3
    bounds: str = "tight"
    sort_mode: str = "global"
    budget_mb: int = 64


def _np(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy())


class SplatImage(torch.autograd.Function):
    """image = render(means, exp(log_scales), quats, sigmoid(opacity_logits), sh) (black
    background); gradients for all five parameter tensors."""

    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, view: View):
        n = means.shape[0]
        m, ls, q, lg, s = (_np(t) for t in (means, log_scales, quats, opacity_logits, sh))
        alpha = (1.0 / (1.0 + np.exp(-lg.astype(np.float64)))).astype(np.float32)
        pos_alpha = np.concatenate([m.reshape(n, 3), alpha.reshape(n, 1)], 1)
        scales = np.concatenate([np.exp(ls.reshape(n, 3)), np.zeros((n, 1), np.float32)], 1).astype(np.float32)
        scene = gsr.Scene(view.ctx, pos_alpha, scales, q.reshape(n, 4), s.reshape(n, -1), view.sh_degree)
        r = gsr.Renderer(view.ctx, scene, bounds=view.bounds, target="rgba32f", sort_mode=view.sort_mode,
                         budget_mb=view.budget_mb, differentiable=True)
        img = r.render_to_host(view.camera)[..., :3]
        ctx.state = (r, scene)
        ctx.meta = [(t.shape, t.device) for t in (means, log_scales, quats, opacity_logits, sh)]
        return torch.from_numpy(np.ascontiguousarray(img)).to(means.device)

    @staticmethod
    def backward(ctx, grad_image):
        r, scene = ctx.state
        try:
            g = r.backward_to_host(_np(grad_image))
        finally:
            r.close()
            scene.close()
            ctx.state = None
        out = []
        for name, (shape, device) in zip(("means", "log_scales", "quats", "opacity_logits", "sh"), ctx.meta):
            out.append(torch.from_numpy(g[name]).reshape(shape).to(device))
        return (*out, None)


def render(means, log_scales, quats, opacity_logits, sh, view: View):
    return SplatImage.apply(means, log_scales, quats, opacity_logits, sh, view)
