"""torch.autograd over the path-replay adjoint: images as differentiable
functions of a scene's textures (GenericMaterial / DiffuseMaterial base colour,
GenericMaterial roughness).

The forward pass uploads the textures and renders iterations
first_iteration .. first_iteration + num_iterations - 1 on the GPU; the backward
pass replays the paths of iterations backward_first_iteration .. (+num) with the
incoming dLoss/dImage and returns dLoss/dtexture. Use a backward_first_iteration
different from first_iteration whenever dLoss/dImage depends on the image
itself (an L2 loss, say): with the same paths, the two estimates are
correlated and their product is biased.

Sampling decisions (pdfs, MIS weights, Russian roulette) are held fixed in the
backward pass; only BSDF values are differentiated. Both passes run entirely
in the native library; the torch side only moves texels and image gradients.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import abi
from .integrator import GpuScene


def _to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to(device="cpu", dtype=torch.float32).contiguous().numpy()


class _RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gscene: GpuScene, settings: abi.PtSettings, first_iteration: int, num_iterations: int,
                backward_first_iteration: int, indices: tuple, *textures: torch.Tensor):
        for idx, tex in zip(indices, textures):
            gscene.write_texture(idx, _to_host(tex))
        image = gscene.render_mean(settings, first_iteration, num_iterations)
        out_device = textures[0].device if textures else torch.device("cpu")
        ctx.gscene = gscene
        ctx.settings = settings
        ctx.num_iterations = num_iterations
        ctx.backward_first_iteration = backward_first_iteration
        ctx.indices = indices
        ctx.save_for_backward(*textures)
        return torch.from_numpy(image).to(out_device)

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        gs: GpuScene = ctx.gscene
        textures = ctx.saved_tensors
        # the texels of the forward pass (a later forward may have replaced them)
        for idx, tex in zip(ctx.indices, textures):
            gs.write_texture(idx, _to_host(tex))
        gs.clear_gradients()
        g = grad_image.detach()
        # a gradient on the GPU goes to the library by its device pointer, without a host round trip
        g = g.to(torch.float32).contiguous() if g.is_cuda else _to_host(g)
        gs.render_adjoint(ctx.settings, g, ctx.backward_first_iteration, ctx.num_iterations)
        grads = []
        for idx, tex, need in zip(ctx.indices, textures, ctx.needs_input_grad[6:]):
            grads.append(torch.from_numpy(gs.gradient(idx)).to(tex.device) if need else None)
        return (None, None, None, None, None, None, *grads)


def render(gscene: GpuScene, settings: abi.PtSettings, textures: Sequence[tuple[int, torch.Tensor]],
           first_iteration: int = 0, num_iterations: int | None = None,
           backward_first_iteration: int | None = None) -> torch.Tensor:
    """Differentiable render of `gscene` as an (H, W, 3) float32 tensor.

    textures: (device texture index, tensor of the texture's shape) pairs; see
    GpuScene.texture_index. num_iterations defaults to settings.total_spp.
    """
    n = int(settings.total_spp if num_iterations is None else num_iterations)
    bfi = first_iteration + n if backward_first_iteration is None else backward_first_iteration
    shapes = gscene.texture_shapes()
    indices = []
    tensors = []
    for idx, tex in textures:
        if tuple(tex.shape) != tuple(shapes[idx]):
            raise ValueError(f"texture {idx}: shape {tuple(tex.shape)} != {tuple(shapes[idx])}")
        indices.append(int(idx))
        tensors.append(tex)
    return _RenderFunction.apply(gscene, settings, int(first_iteration), n, int(bfi), tuple(indices), *tensors)
