"""Differentiable unrolled TV primal-dual reconstruction with a hand-written
adjoint and checkpointing.

``tv_unrolled(sino, ps, mu, number_iterations, ...)`` runs the iteration of
:func:`gratopy_amd.total_variation` (gratopy/gratopy.py:2554-2762) for a fixed
number of steps and returns u_N as a torch tensor that carries a gradient
with respect to ``sino``.

Forward: the iteration runs on the HIP kernels; the full state
(u, u_bar, v, lambda) is stored every ``checkpoint`` iterations.
Backward: for each segment, last to first, the forward steps are recomputed
from the segment's checkpoint (keeping every state of that segment), then the
vector-Jacobian product of each step is applied in reverse with the adjoint
kernels of the C ABI (gt_tv_adj_*), one forward projection (B^T) and one
backprojection (K^T) per step.  The step map is piecewise linear: the
primal clamp max(0, .) contributes the mask [u_{i+1} > 0] and the dual ball
projection its Jacobian (I - a a^T/|a|^2)/|a| where it was active; at the
measure-zero kinks one side's derivative is taken.
"""

from __future__ import annotations

import numpy as np

from . import _device as dev
from . import _hip
from .autograd import transpose_scale
from .gratopy import backprojection, forwardprojection, normest

torch = dev.torch


class _Params:
    """The scalars of the iteration, rounded to float32 as the kernels take them."""

    def __init__(self, ps, mu, norm_estimate, slice_thickness, stepsize_weighting):
        sigma = 1.0 / np.sqrt(13) * stepsize_weighting
        tau = 1.0 / np.sqrt(13) / stepsize_weighting
        mu = mu * norm_estimate**2 * ps.delta_x
        mu = mu / (sigma + mu)
        f32 = lambda v: float(np.float32(v))
        self.sigma = f32(sigma)
        self.sigma_p = f32(sigma / norm_estimate)
        self.c = f32(mu)
        self.tau = f32(tau)
        self.n = f32(1.0 / norm_estimate)
        self.zd = f32(slice_thickness)
        self.cT = transpose_scale(ps)


def _dims(ps, sino):
    Nx, Ny = ps.img_shape
    Nz = int(sino.shape[2]) if sino.dim() > 2 else 1
    return Nx, Ny, Nz


def _step(ps, prm, f, state, order):
    """One primal-dual step; returns the new state (fresh tensors)."""
    u, ubar, v, lam = state
    s = ps.queue.torch_stream()
    dcode = _hip.dtype_code(f.dtype)
    ocode = dev.order_code(order)
    Nx, Ny, Nz = _dims(ps, f)
    Ns, Na = ps.sinogram_shape
    with torch.cuda.stream(s):
        v2 = v.clone()
        lam2 = lam.clone(memory_format=torch.preserve_format)
        u_new = torch.empty_like(u)
        u_ext = u.clone(memory_format=torch.preserve_format)
    _hip.call("gt_tv_dual", s.cuda_stream, dcode, v2.data_ptr(), ubar.data_ptr(), prm.sigma, prm.zd,
              Nx, Ny, Nz, ocode)
    ku = forwardprojection(ubar, ps)
    _hip.call("gt_tv_update_lambda", s.cuda_stream, dcode, lam2.data_ptr(), ku.data_ptr(), f.data_ptr(),
              prm.sigma_p, prm.c, Ns, Na, Nz, ocode)
    kstar = backprojection(lam2, ps)
    _hip.call("gt_tv_primal", s.cuda_stream, dcode, u_new.data_ptr(), u_ext.data_ptr(), v2.data_ptr(),
              kstar.data_ptr(), prm.tau, prm.n, prm.zd, Nx, Ny, Nz, ocode)
    return (u_new, u_ext, v2, lam2)


def _adjoint_step(ps, prm, f, state, u_next, cot, fb, order):
    """Cotangents of the step's inputs from those of its outputs.

    state = (u, ubar, v, lam) before the step, u_next the step's new u,
    cot = (ub', ubarb', vb', lamb'); fb accumulates the cotangent of f.
    """
    ub_n, ubarb_n, vb_n, lamb_n = cot
    u, ubar, v, lam = state
    s = ps.queue.torch_stream()
    dcode = _hip.dtype_code(f.dtype)
    ocode = dev.order_code(order)
    Nx, Ny, Nz = _dims(ps, f)
    Ns, Na = ps.sinogram_shape
    dtype = dev.to_np_dtype(f.dtype)
    ps.ensure_dtype(dtype)
    w = ps.angle_weights_buf[dtype]
    with torch.cuda.stream(s):
        wb = torch.empty_like(u)
        ub = torch.empty_like(u)
        lamb = torch.empty_like(lam)
        winv = torch.empty_like(lam)
        ab = torch.empty_like(v)
        ubarb = torch.empty_like(u)
    _hip.call("gt_tv_adj_primal", s.cuda_stream, dcode, wb.data_ptr(), ub.data_ptr(), ub_n.data_ptr(),
              ubarb_n.data_ptr(), u_next.data_ptr(), u.numel())
    rwb = forwardprojection(wb, ps)
    _hip.call("gt_tv_adj_lambda", s.cuda_stream, dcode, lamb.data_ptr(), fb.data_ptr(), winv.data_ptr(),
              lamb_n.data_ptr(), rwb.data_ptr(), w.data_ptr(), prm.tau * prm.n / prm.cT, prm.c,
              prm.sigma_p, Ns, Na, Nz, ocode)
    bterm = backprojection(winv, ps)
    _hip.call("gt_tv_adj_dual", s.cuda_stream, dcode, ab.data_ptr(), vb_n.data_ptr(), wb.data_ptr(),
              v.data_ptr(), ubar.data_ptr(), prm.sigma, prm.tau, prm.zd, Nx, Ny, Nz, ocode)
    _hip.call("gt_tv_adj_div", s.cuda_stream, dcode, ubarb.data_ptr(), bterm.data_ptr(),
              prm.sigma_p * prm.cT, ab.data_ptr(), prm.sigma, prm.zd, Nx, Ny, Nz, ocode)
    return (ub, ubarb, ab, lamb)


def _zero_state(ps, f, order):
    Nx, Ny, Nz = _dims(ps, f)
    shape = (Nx, Ny) + ((Nz,) if f.dim() > 2 else ())
    dt = dev.to_np_dtype(f.dtype)
    u = dev.zeros(shape, dt, order=order, device=f.device)
    ubar = dev.zeros(shape, dt, order=order, device=f.device)
    v = torch.zeros(4 * u.numel(), dtype=f.dtype, device=f.device)
    lam = dev.zeros(tuple(f.shape), dt, order=order, device=f.device)
    return (u, ubar, v, lam)


class _UnrolledTV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, ps, prm, n_iter, checkpoint):
        order = "C" if dev.is_c_contiguous(sino) else "F"
        f = sino.detach()
        state = _zero_state(ps, f, order)
        checkpoints = {}
        for i in range(n_iter):
            if i % checkpoint == 0:
                checkpoints[i] = state
            state = _step(ps, prm, f, state, order)
        ctx.ps, ctx.prm, ctx.n_iter, ctx.checkpoint, ctx.order = ps, prm, n_iter, checkpoint, order
        ctx.checkpoints = checkpoints
        ctx.save_for_backward(f)
        return state[0]

    @staticmethod
    def backward(ctx, grad_u):
        (f,) = ctx.saved_tensors
        ps, prm, order = ctx.ps, ctx.prm, ctx.order
        n_iter, k = ctx.n_iter, ctx.checkpoint
        s = ps.queue.torch_stream()
        with torch.cuda.stream(s):
            fb = torch.zeros_like(f)
            ub = grad_u.clone(memory_format=torch.preserve_format)
            if dev.order_of(ub) != order:
                ub = dev.to_device(dev.to_host(ub), f.device, order=order)
            zeros = _zero_state(ps, f, order)
            cot = (ub, zeros[1], zeros[2], zeros[3])
        for start in reversed(range(0, n_iter, k)):
            stop = min(start + k, n_iter)
            states = [ctx.checkpoints[start]]
            for _ in range(start, stop):
                states.append(_step(ps, prm, f, states[-1], order))
            for i in reversed(range(start, stop)):
                cot = _adjoint_step(ps, prm, f, states[i - start], states[i - start + 1][0], cot, fb, order)
        ctx.checkpoints = None
        return fb, None, None, None, None


def tv_unrolled(sino, projectionsetting, mu, number_iterations=100, *, norm_estimate=None,
                slice_thickness=0.0, stepsize_weighting=10.0, checkpoint=10):
    """Differentiable ``number_iterations``-step TV primal-dual reconstruction.

    Same iteration and parameters as :func:`gratopy_amd.total_variation`
    (slices of a 3-D ``sino`` coupled through ``slice_thickness``; 0 treats
    them as a batch of independent problems).  Returns u_N; gradients flow
    to ``sino``.
    """
    ps = projectionsetting
    if norm_estimate is None:
        norm_estimate = normest(ps)
    if sino.dim() == 2:
        slice_thickness = 0.0
    prm = _Params(ps, mu, norm_estimate, slice_thickness, stepsize_weighting)
    if not (dev.is_c_contiguous(sino) or dev.is_f_contiguous(sino)):
        sino = sino.contiguous()
    return _UnrolledTV.apply(sino, ps, prm, int(number_iterations), int(checkpoint))
