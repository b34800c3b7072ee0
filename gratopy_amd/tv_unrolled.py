"""Differentiable unrolled TV primal-dual reconstruction with a hand-written
adjoint and checkpointing.

``tv_unrolled(sino, ps, mu, number_iterations, ...)`` runs the iteration of
:func:`gratopy_amd.total_variation` (gratopy/gratopy.py:2554-2762) for a fixed
number of steps and returns u_N as a torch tensor that carries a gradient
with respect to ``sino``.

Forward: the iteration runs in place on the HIP kernels; the full state
(u, u_bar, v, lambda) is copied into a checkpoint slot every ``checkpoint``
iterations.
Backward: for each segment, last to first, the forward steps are recomputed
from the segment's checkpoint into a pool of states reused by every segment,
then the vector-Jacobian product of each step is applied in reverse, in
place on one set of cotangents, with the adjoint kernels of the C ABI
(gt_tv_adj_*), one forward projection (B^T) and one backprojection (K^T) per
step.  Each pass allocates its buffers once, before its first step.  The step map is piecewise linear: the
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


class _Buffers:
    """Every device buffer of one unrolled run, allocated once per pass: the
    working state (u, u_bar, v, lambda), checkpoint states, a segment pool
    for the backward recomputation, the cotangent state and the projection
    scratch.  Each step then only launches kernels."""

    def __init__(self, ps, f, order):
        self.ps, self.f, self.order = ps, f, order
        Nx, Ny, Nz = _dims(ps, f)
        self.img_shape = (Nx, Ny) + ((Nz,) if f.dim() > 2 else ())
        self.dtype = dev.to_np_dtype(f.dtype)
        self.ku = self.sino()
        self.kstar = self.img()

    def img(self):
        return dev.zeros(self.img_shape, self.dtype, order=self.order, device=self.f.device)

    def sino(self):
        return dev.zeros(tuple(self.f.shape), self.dtype, order=self.order, device=self.f.device)

    def dual(self):
        return torch.zeros(4 * int(np.prod(self.img_shape)), dtype=self.f.dtype, device=self.f.device)

    def state(self):
        return [self.img(), self.img(), self.dual(), self.sino()]


def _copy_state(dst, src):
    for d, s in zip(dst, src):
        d.copy_(s)


def _step(ps, prm, f, src, dst, buf):
    """One primal-dual step from state `src` = [u, u_bar, v, lambda] into
    `dst` (dst is src: in place, with u and u_bar swapping buffers)."""
    s = ps.queue.torch_stream()
    dcode = _hip.dtype_code(f.dtype)
    ocode = dev.order_code(buf.order)
    Nx, Ny, Nz = _dims(ps, f)
    Ns, Na = ps.sinogram_shape
    u, ubar, v, lam = src
    if dst is not src:
        with torch.cuda.stream(s):
            dst[2].copy_(v)
            dst[3].copy_(lam)
            dst[1].copy_(u)
        v, lam = dst[2], dst[3]
    _hip.call("gt_tv_dual", s.cuda_stream, dcode, v.data_ptr(), ubar.data_ptr(), prm.sigma, prm.zd,
              Nx, Ny, Nz, ocode)
    forwardprojection(ubar, ps, sino=buf.ku)
    _hip.call("gt_tv_update_lambda", s.cuda_stream, dcode, lam.data_ptr(), buf.ku.data_ptr(),
              f.data_ptr(), prm.sigma_p, prm.c, Ns, Na, Nz, ocode)
    backprojection(lam, ps, img=buf.kstar)
    if dst is src:
        # u_bar (consumed above) receives the new u, u the extrapolation
        _hip.call("gt_tv_primal", s.cuda_stream, dcode, ubar.data_ptr(), u.data_ptr(), v.data_ptr(),
                  buf.kstar.data_ptr(), prm.tau, prm.n, prm.zd, Nx, Ny, Nz, ocode)
        src[0], src[1] = ubar, u
    else:
        _hip.call("gt_tv_primal", s.cuda_stream, dcode, dst[0].data_ptr(), dst[1].data_ptr(),
                  v.data_ptr(), buf.kstar.data_ptr(), prm.tau, prm.n, prm.zd, Nx, Ny, Nz, ocode)


def _adjoint_step(ps, prm, f, state, u_next, cot, fb, buf, scratch):
    """Replace the cotangents `cot` = [ub', ubarb', vb', lamb'] of a step's
    outputs by those of its inputs: in place, since the adjoint kernels read
    and write each element at one index, except the dual cotangent, which
    alternates with scratch[4] (gt_tv_adj_dual takes no aliased output).
    state = [u, u_bar, v, lambda] before the step, u_next the step's new u;
    fb accumulates the cotangent of f."""
    ub, ubarb, vb, lamb = cot
    u, ubar, v, lam = state
    wb, rwb, winv, bterm, ab = scratch
    s = ps.queue.torch_stream()
    dcode = _hip.dtype_code(f.dtype)
    ocode = dev.order_code(buf.order)
    Nx, Ny, Nz = _dims(ps, f)
    Ns, Na = ps.sinogram_shape
    w = ps.angle_weights_buf[buf.dtype]
    _hip.call("gt_tv_adj_primal", s.cuda_stream, dcode, wb.data_ptr(), ub.data_ptr(), ub.data_ptr(),
              ubarb.data_ptr(), u_next.data_ptr(), u.numel())
    forwardprojection(wb, ps, sino=rwb)
    _hip.call("gt_tv_adj_lambda", s.cuda_stream, dcode, lamb.data_ptr(), fb.data_ptr(), winv.data_ptr(),
              lamb.data_ptr(), rwb.data_ptr(), w.data_ptr(), prm.tau * prm.n / prm.cT, prm.c,
              prm.sigma_p, Ns, Na, Nz, ocode)
    backprojection(winv, ps, img=bterm)
    _hip.call("gt_tv_adj_dual", s.cuda_stream, dcode, ab.data_ptr(), vb.data_ptr(), wb.data_ptr(),
              v.data_ptr(), ubar.data_ptr(), prm.sigma, prm.tau, prm.zd, Nx, Ny, Nz, ocode)
    _hip.call("gt_tv_adj_div", s.cuda_stream, dcode, ubarb.data_ptr(), bterm.data_ptr(),
              prm.sigma_p * prm.cT, ab.data_ptr(), prm.sigma, prm.zd, Nx, Ny, Nz, ocode)
    cot[2], scratch[4] = ab, vb


class _UnrolledTV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, ps, prm, n_iter, checkpoint):
        order = "C" if dev.is_c_contiguous(sino) else "F"
        f = sino.detach()
        ps.ensure_dtype(dev.to_np_dtype(f.dtype))
        s = ps.queue.torch_stream()
        with torch.cuda.stream(s):
            buf = _Buffers(ps, f, order)
            state = buf.state()
            checkpoints = [buf.state() for _ in range(0, n_iter, checkpoint)]
        for i in range(n_iter):
            if i % checkpoint == 0:
                with torch.cuda.stream(s):
                    _copy_state(checkpoints[i // checkpoint], state)
            _step(ps, prm, f, state, state, buf)
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
            buf = _Buffers(ps, f, order)
            fb = buf.sino()
            cot = buf.state()
            cot[0].copy_(grad_u)  # element by element, whatever the strides of grad_u
            pool = [buf.state() for _ in range(min(k, n_iter))]  # states 1..k of a segment
            scratch = [buf.img(), buf.sino(), buf.sino(), buf.img(), buf.dual()]
        for start in reversed(range(0, n_iter, k)):
            stop = min(start + k, n_iter)
            states = [ctx.checkpoints[start // k]] + pool[: stop - start]
            for j in range(stop - start):
                _step(ps, prm, f, states[j], states[j + 1], buf)
            for i in reversed(range(stop - start)):
                _adjoint_step(ps, prm, f, states[i], states[i + 1][0], cot, fb, buf, scratch)
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
