"""Differentiable forward projection for PyTorch.

``project(volume, axes)`` is the full spatial projector of the reference's
forwardProject (``MultiGPUGridder_Matlab_Class.m:198-243``): Vol->CAS
(``utils/CASFromVol_Gridder.m``), the Kaiser-Bessel gather
(``gpuForwardProject.cu``) and the inverse image chain
(``gpuGridder.cpp:432-518``), each a HIP kernel of ``libMultiGPUGridder.so``.
Its backward applies the exact transpose, stage by stage in reverse order:

    grad_volume = m^3 CASToVol_keep0( Gather^T( ZeroIdx0( ImgsToCAS(grad_images) ) / m^2 ) )

where Gather^T is the hand-written adjoint scatter of the gather (the same
(pixel, voxel, weight) triples), CASToVol_keep0 is volFromCAS without the
index-0 planes dropped, and ZeroIdx0 clears the index-0 row / column of every
CAS image, which the forward chain's FromCAS discards.  The gradient is taken
with respect to the volume only; the directions are treated as constants.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops


def _sizes(n, interp_factor, extra_padding):
    m = int(np.float32(n) * np.float32(interp_factor))
    return m, m + 2 * extra_padding


class ForwardProjection(torch.autograd.Function):
    """images (P, N, N) = P(volume (N, N, N)) for directions (P, 9)."""

    @staticmethod
    def forward(ctx, volume, axes, interp_factor=2.0, extra_padding=3):
        if volume.dtype != torch.float32 or not volume.is_cuda:
            raise TypeError("volume must be a float32 CUDA tensor")
        n = volume.shape[-1]
        m, _ = _sizes(n, interp_factor, extra_padding)
        axes = axes.to(device=volume.device, dtype=torch.float32).reshape(-1, 9).contiguous()
        cas = ops.volume_to_cas_volume(volume.detach().contiguous(), interp_factor, extra_padding)
        cas_imgs = ops.forward_project(cas, axes, m)
        imgs = ops.cas_images_to_images(cas_imgs, n, clobber=True)
        ctx.save_for_backward(axes)
        ctx.geom = (n, interp_factor, extra_padding)
        return imgs

    @staticmethod
    def backward(ctx, grad_imgs):
        (axes,) = ctx.saved_tensors
        n, interp_factor, extra_padding = ctx.geom
        m, v = _sizes(n, interp_factor, extra_padding)
        g = grad_imgs.to(torch.float32).contiguous()
        cas_imgs = ops.cas_images_to_images_transpose(g, m)
        cas_vol = ops.back_project_adjoint(cas_imgs, axes, v)
        grad_vol = ops.volume_to_cas_volume_transpose(cas_vol, n, interp_factor, extra_padding)
        return grad_vol, None, None, None


def project(volume: torch.Tensor, axes, interp_factor: float = 2.0, extra_padding: int = 3) -> torch.Tensor:
    """Differentiable forward projection of ``volume`` along ``axes`` (P x 9)."""
    if not torch.is_tensor(axes):
        axes = torch.from_numpy(np.ascontiguousarray(axes, dtype=np.float32))
    return ForwardProjection.apply(volume, axes, interp_factor, extra_padding)


def project_transpose(images: torch.Tensor, axes, n: int, interp_factor: float = 2.0,
                      extra_padding: int = 3) -> torch.Tensor:
    """The exact transpose P^T applied to ``images`` (what backward computes)."""
    if not torch.is_tensor(axes):
        axes = torch.from_numpy(np.ascontiguousarray(axes, dtype=np.float32))
    axes = axes.to(device=images.device, dtype=torch.float32).reshape(-1, 9).contiguous()
    m, v = _sizes(n, interp_factor, extra_padding)
    cas_imgs = ops.cas_images_to_images_transpose(images.contiguous(), m)
    cas_vol = ops.back_project_adjoint(cas_imgs, axes, v)
    return ops.volume_to_cas_volume_transpose(cas_vol, n, interp_factor, extra_padding)
