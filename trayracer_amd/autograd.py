"""torch.autograd.Function over the differentiable render (test wrapper).

forward(params [M, 5]) renders one frame with params as both the evaluation and the
sampling parameters and returns the frame's mean image [H, W, 3]; backward runs the
path-replay adjoint through the C ABI (trt_backward). Tensors cross as host arrays.

The adjoint replays the renderer's last frame. When the renderer rendered another frame
after this node's forward (several render() calls before loss.backward()), backward
restores this node's parameters and settings and renders its frame again: setting the
parameters restarts accumulation, so the frame is rendered from sample 0 with the same
options, which reproduces it bit for bit.
"""
import copy

import numpy as np
import torch


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, renderer):
        p = params.detach().cpu().numpy().astype(np.float32)
        renderer.set_material_params(p, which=renderer.BOTH)
        renderer.RenderFrame()
        img = renderer.color()[..., :3].copy()
        ctx.renderer = renderer
        ctx.params = p
        ctx.settings = copy.copy(renderer.settings)
        ctx.frame = renderer.frame_id()
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_img):
        r = ctx.renderer
        if r.frame_id() != ctx.frame:  # another frame was rendered since: render this one again
            current = r.settings
            r.settings = copy.copy(ctx.settings)
            try:
                r.set_material_params(ctx.params, which=r.BOTH)
                r.RenderFrame()
                g = r.backward(grad_img.detach().cpu().numpy().astype(np.float32))
            finally:
                r.settings = current
        else:
            g = r.backward(grad_img.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(g), None


def render(params, renderer):
    return RenderFunction.apply(params, renderer)
