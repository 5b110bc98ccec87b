"""torch.autograd.Function over the differentiable render (test wrapper).

forward(params [M, 5]) renders one frame with params as both the evaluation and the
sampling parameters and returns the frame's mean image [H, W, 3]; backward runs the
path-replay adjoint through the C ABI (trt_backward). Tensors cross as host arrays.
"""
import numpy as np
import torch


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, renderer):
        renderer.set_material_params(params.detach().cpu().numpy().astype(np.float32), which=renderer.BOTH)
        renderer.RenderFrame()
        img = renderer.color()[..., :3].copy()
        ctx.renderer = renderer
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_img):
        g = ctx.renderer.backward(grad_img.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(g), None


def render(params, renderer):
    return RenderFunction.apply(params, renderer)
