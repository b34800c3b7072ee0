"""torch.autograd operators over the HIP projection kernels.

gratopy's backprojection B is the adjoint of the forward projection R in
*weighted* inner products, <R x, y>_sino = <x, B y>_img with
<u, v>_img = dx^2 sum u v and <g, h>_sino = ds sum_a w_a sum_s g h
(doc/source/getting_started.rst:91-131).  Autograd needs Euclidean
transposes:

    R^T g = c * B(W^-1 g),      B^T x = (1/c) * W R(x),      c = dx^2/ds,

with W = diag(angle weights).  For the parallel beam the forward kernel
scales by float(dx)^2/float(ds) even in double precision (radon.cl:82-83),
so c uses those float-rounded values and the transposes stay exact.

Both backward passes are themselves projections on the GPU (one
backprojection or one forward projection plus an angle weighting and a
scale), so gradients cost one extra transform each.
"""

from __future__ import annotations

import numpy as np

from . import _device as dev
from . import _hip
from .gratopy import backprojection, forwardprojection, weight_sinogram

torch = dev.torch


def transpose_scale(projectionsetting) -> float:
    """c with R^T = c B W^-1 for this geometry (float-rounded for Radon)."""
    ps = projectionsetting
    if ps.is_parallel:
        dx = float(np.float32(ps.delta_x))
        ds = float(np.float32(ps.delta_s))
        return dx * dx / ds
    return ps.delta_x**2 / ps.delta_s


def _scale_(t, a, ps):
    stream = ps.queue.torch_stream()
    _hip.call("gt_scale", stream.cuda_stream, _hip.dtype_code(t.dtype), t.data_ptr(), float(a), t.numel())
    return t


def _contig(t):
    """Inputs of the kernels must be C or F contiguous."""
    if dev.is_c_contiguous(t) or dev.is_f_contiguous(t):
        return t
    return t.contiguous()


def _apply_transpose_forward(grad_sino, ps):
    """R^T g = c * B(W^-1 g)."""
    g = weight_sinogram(_contig(grad_sino), ps, divide=True)
    out = backprojection(g, ps)
    return _scale_(out, transpose_scale(ps), ps)


def _apply_transpose_back(grad_img, ps):
    """B^T x = W R(x) / c."""
    s = forwardprojection(_contig(grad_img), ps)
    weight_sinogram(s, ps, s)
    return _scale_(s, 1.0 / transpose_scale(ps), ps)


class ForwardProjection(torch.autograd.Function):
    """sino = R img; d/d img = R^T (Euclidean)."""

    @staticmethod
    def forward(ctx, img, projectionsetting):
        ctx.ps = projectionsetting
        return forwardprojection(_contig(img), projectionsetting)

    @staticmethod
    def backward(ctx, grad):
        return _apply_transpose_forward(grad, ctx.ps), None


class BackProjection(torch.autograd.Function):
    """img = B sino; d/d sino = B^T (Euclidean)."""

    @staticmethod
    def forward(ctx, sino, projectionsetting):
        ctx.ps = projectionsetting
        return backprojection(_contig(sino), projectionsetting)

    @staticmethod
    def backward(ctx, grad):
        return _apply_transpose_back(grad, ctx.ps), None


def project(img, projectionsetting):
    """Differentiable forward projection."""
    return ForwardProjection.apply(img, projectionsetting)


def backproject(sino, projectionsetting):
    """Differentiable backprojection."""
    return BackProjection.apply(sino, projectionsetting)


def tv_eps(x, eps: float):
    """Smoothed isotropic total variation sum sqrt(|grad x|^2 + eps^2) with the
    forward differences and Neumann boundary of gratopy's TV solver
    (total_variation.cl:54-93), per slice (torch autograd)."""
    gx = torch.zeros_like(x)
    gy = torch.zeros_like(x)
    gx[:-1] = x[1:] - x[:-1]
    gy[:, :-1] = x[:, 1:] - x[:, :-1]
    return torch.sqrt(gx * gx + gy * gy + eps * eps).sum()
