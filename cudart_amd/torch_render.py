"""torch.autograd wrapper of the path tracer: differentiable rendering with respect
to per-material diffuse albedo (RGB) and emission scale (MTL Ka = e (1,1,1)).

forward renders the image with the HIP kernels (mean radiance of `spp` samples at the
renderer's seed); backward runs the path-replay adjoint kernel with the incoming
dLoss/dImage. Russian-roulette decisions are differentiated through their survival
probability; the loss change of such a decision is taken to second order with the
curvature `hess` (d2Loss/dI2 per element: 2/(H*W*3) for the mean squared error, 0 for
a first-order estimate). A renderer created with keep_samples=1 stores each sample's
radiance in the forward pass, and backward then only replays the paths.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ParallelRenderer


class PathTrace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo: torch.Tensor, emission: torch.Tensor, renderer: ParallelRenderer, spp: int,
                hess: float = 0.0):
        kd = albedo.detach().to("cpu", torch.float32).numpy()
        em = emission.detach().to("cpu", torch.float32).numpy()
        for m in range(kd.shape[0]):
            renderer.scene.set_material_params(m, kd[m], float(em[m]))
        renderer.reset()
        renderer.render(spp)
        img = torch.from_numpy(renderer.radiance()).to(albedo.device)
        ctx.renderer = renderer
        ctx.spp = spp
        ctx.hess = hess
        ctx.devices = (albedo.device, emission.device)
        ctx.params = (kd.copy(), em.copy())
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        g = np.ascontiguousarray(grad_img.detach().to("cpu", torch.float32).numpy())
        # The gradient is taken at the parameters this node's forward rendered with:
        # when another forward call has set other ones since, they are set back (the
        # renderer's kept radiance then no longer matches the scene's materials, and
        # the adjoint traces each sample's primal path at these parameters again).
        scene = ctx.renderer.scene
        kd, em = ctx.params
        if not scene.has_material_params(kd, em):
            for m in range(kd.shape[0]):
                scene.set_material_params(m, kd[m], float(em[m]))
        d_alb, d_em = ctx.renderer.adjoint_grad(g, ctx.spp, ctx.hess)
        return (torch.from_numpy(d_alb).to(ctx.devices[0]), torch.from_numpy(d_em).to(ctx.devices[1]), None, None,
                None)


def render(albedo: torch.Tensor, emission: torch.Tensor, renderer: ParallelRenderer, spp: int,
           hess: float = 0.0) -> torch.Tensor:
    """(H, W, 3) float32 image; differentiable in albedo (M, 3) and emission (M,)."""
    return PathTrace.apply(albedo, emission, renderer, spp, hess)
