"""The differentiation configuration at its real size, in float64: Nx = 512,
360 angles, 768 detectors, Radon and fanbeam (detector width 1024, RE 400,
R 800, image width 512 for fanbeam).

* The unrolled TV primal-dual solver (100 iterations, checkpoints every 10)
  on a batch of 16 slices inside the configuration's loss
  L(f) = 0.5 ||R x - y||^2 + lam TV_eps(x), x = TV_100(f), lam = 1e-3,
  eps = 1e-4 (TV's mu = 1e-2 for fanbeam, the same data weight for Radon):
  the gradient with respect to the measured sinograms f from the
  hand-written adjoint (gt_tv_adj_* kernels) against finite differences
  (five-point stencil, step 2e-9 max|f|) in 32 random directions, on phantoms with a
  constant background of 0.25 so that the reconstruction stays off the clamp
  max(0, .).  The iteration is piecewise linear (that clamp, unit-ball
  projections): a +-h step that crosses a kink in some pixel changes the
  difference quotient slightly, so the median relative error is held to 1e-6
  and the largest to 1e-4.
* One slice against the gradient assembled from the oracle's operators:
  the same iteration in torch on the CPU, with R and its backprojection
  from the oracle (the reference kernels restated in C) as autograd
  functions whose transposes are the oracle's own B and R
  (R^T g = c B(W^-1 g), B^T x = W R(x) / c), differentiated by torch
  autograd -- no HIP kernel and no hand-written adjoint on that side.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd.autograd import project, transpose_scale, tv_eps
from gratopy_amd.phantom import add_noise, random_ellipses
from gratopy_amd.tv_unrolled import _Params, tv_unrolled
from oracle import solvers as osolv
from tests.helpers import record_parity, rel_err

pytestmark = pytest.mark.gpu

F64 = np.dtype("float64")
LAM, EPS, ITERS, CHECKPOINT = 1e-3, 1e-4, 100, 10
# TV's data weight mu ||R||^2 dx (gratopy.py:2738-2741): mu = 1e-2 gives 67 for
# the fanbeam geometry (||R|| = 81.8, dx = 1); the parallel beam's pixels are
# 2/512 wide (||R|| = 2.45), where mu = 1e-2 would weight the data by 2e-4 and
# leave a reconstruction clamped to zero almost everywhere, so mu is chosen
# per geometry for the same weight
DATA_WEIGHT = 67.0


def _mu(ps, norm):
    return DATA_WEIGHT / (norm**2 * ps.delta_x)


def _ps(geometry):
    if geometry == "radon":
        return gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (512, 512), 360, 768)
    return gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (512, 512), 360, 768, R=800.0, RE=400.0,
                                      detector_width=1024.0, image_width=512.0)


def _norm(ps):
    np.random.seed(0)
    return float(gratopy.normest(ps, number_iterations=50, dtype="float64"))


def _batch(ps, nz, seed, background=0.0):
    """Phantoms (plus a constant background), their clean sinograms y (the
    loss's target) and noisy sinograms f (the solver's input), (Ns, Na, nz)
    float64 F order."""
    rng = np.random.default_rng(seed)
    imgs = np.stack([random_ellipses(512, 15, seed=seed * 100 + z, dtype=F64) + background
                     for z in range(nz)], axis=-1)
    y = gratopy.forwardprojection(dev.to_device(np.asfortranarray(imgs), "cuda:0"), ps)
    f = dev.to_device(np.asfortranarray(add_noise(dev.to_host(y), 1.0, rng)), "cuda:0")
    return y, f


@pytest.mark.timeout(900)
@pytest.mark.parametrize("geometry", ["fan", "radon"])
def test_config4_unrolled_tv_finite_differences_batch16(geometry):
    print(f"config4 unrolled TV(100) batch 16, 32 FD directions, {geometry}", flush=True)
    ps = _ps(geometry)
    norm = _norm(ps)
    mu = _mu(ps, norm)
    # a positive background keeps the reconstruction off the clamp max(0, .):
    # with a zero background about half of its pixels sit exactly at that
    # kink, and +-h steps then straddle it in many pixels at once
    y, f0 = _batch(ps, 16, 41 if geometry == "fan" else 42, background=0.25)

    def loss(f):
        x = tv_unrolled(f, ps, mu, ITERS, norm_estimate=norm, checkpoint=CHECKPOINT)
        return 0.5 * ((project(x, ps) - y) ** 2).sum() + LAM * sum(tv_eps(x[:, :, z], EPS)
                                                                 for z in range(x.shape[2]))

    f = f0.clone().requires_grad_(True)
    loss(f).backward()
    grad = f.grad.detach()
    # the step is relative to the data: 1e-6 on the fanbeam sinograms (values
    # up to ~500 in pixel units); the parallel-beam ones are ~250x smaller
    # (pixels 2/512 wide), where an absolute 1e-6 would cross ~250x more
    # kinks of the iteration
    h = 2e-9 * float(f0.abs().max())
    errs = []
    with torch.no_grad():
        for k in range(32):
            d = torch.tensor(np.random.default_rng(700 + k).standard_normal(f0.shape), device="cuda:0")
            # five-point stencil: TV_eps has curvature ~1/eps where the
            # reconstruction is flat, which puts ~1e-4 of error into a
            # central difference at h = 1e-6; this one's error is O(h^4)
            L = [float(loss(f0 + m * h * d)) for m in (-2, -1, 1, 2)]
            fd = (L[0] - 8 * L[1] + 8 * L[2] - L[3]) / (12 * h)
            an = float((grad * d).sum())
            errs.append(abs(fd - an) / abs(an))
    errs = np.array(errs)
    record_parity("config4", f"unrolled_tv100_fd_{geometry}", batch=16, directions=32, mu=mu, h=h,
                  median_rel_err=float(np.median(errs)), max_rel_err=float(errs.max()))
    assert np.median(errs) < 1e-6 and errs.max() < 1e-4, errs


# ---- the same iteration on the CPU over the oracle's operators ---------------------


class _OracleR(torch.autograd.Function):
    """sino = R x by the oracle; x^bar = R^T g = c B(W^-1 g), c = dx^2/ds."""

    @staticmethod
    def forward(ctx, x, g, c):
        ctx.g, ctx.c = g, c
        return torch.from_numpy(g.fwd(np.asfortranarray(x.numpy())))

    @staticmethod
    def backward(ctx, gs):
        w = ctx.g.angle_weights(F64)
        return torch.from_numpy(ctx.c * ctx.g.back(np.asfortranarray(gs.numpy() / w[None, :]))), None, None


class _OracleB(torch.autograd.Function):
    """img = B s by the oracle; s^bar = B^T g = W R(g) / c."""

    @staticmethod
    def forward(ctx, s, g, c):
        ctx.g, ctx.c = g, c
        return torch.from_numpy(g.back(np.asfortranarray(s.numpy())))

    @staticmethod
    def backward(ctx, gx):
        w = ctx.g.angle_weights(F64)
        return torch.from_numpy(w[None, :] * ctx.g.fwd(np.asfortranarray(gx.numpy())) / ctx.c), None, None


def _grad3(x):
    g = torch.zeros((3,) + x.shape, dtype=x.dtype)
    g[0, :-1] = x[1:] - x[:-1]
    g[1, :, :-1] = x[:, 1:] - x[:, :-1]
    return g


def _div3(p):
    d = torch.zeros(p.shape[1:], dtype=p.dtype)
    d[0] = p[0, 0]
    d[1:-1] += p[0, 1:-1] - p[0, :-2]
    d[-1] += -p[0, -2]
    d[:, 0] += p[1, :, 0]
    d[:, 1:-1] += p[1, :, 1:-1] - p[1, :, :-2]
    d[:, -1] += -p[1, :, -2]
    return d


def _oracle_unrolled(f, g, prm, n_iter, c):
    """The TV primal-dual iteration of gratopy.py:2744-2759 (slice_thickness
    0) in torch ops on the CPU, projections by the oracle."""
    u = torch.zeros(g.img_shape, dtype=f.dtype)
    ubar = torch.zeros_like(u)
    v = torch.zeros((3,) + g.img_shape, dtype=f.dtype)
    lam = torch.zeros_like(f)
    for _ in range(n_iter):
        a = v + prm.sigma * _grad3(ubar)
        norm = torch.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)
        v = a / torch.where(norm > 1, norm, torch.ones_like(norm))
        lam = (lam + prm.sigma_p * (_OracleR.apply(ubar, g, c) - f)) * prm.c
        u_new = torch.relu(u + prm.tau * (_div3(v) - prm.n * _OracleB.apply(lam, g, c)))
        ubar = 2 * u_new - u
        u = u_new
    return u


@pytest.mark.timeout(900)
def test_config4_unrolled_gradient_against_oracle_operators():
    print("config4 unrolled TV(100) gradient vs oracle-assembled, radon, 1 slice", flush=True)
    ps = _ps("radon")
    g = osolv.OracleGeometry.from_settings(ps)
    norm = _norm(ps)
    mu = _mu(ps, norm)
    _, f_dev = _batch(ps, 1, 43)
    f_host = dev.to_host(f_dev)[:, :, 0]
    target = np.random.default_rng(44).random(ps.img_shape)

    f1 = dev.to_device(np.asfortranarray(f_host), "cuda:0").requires_grad_(True)
    u1 = tv_unrolled(f1, ps, mu, ITERS, norm_estimate=norm, checkpoint=CHECKPOINT)
    (0.5 * ((u1 - torch.tensor(target, device="cuda:0")) ** 2).sum()).backward()

    prm = _Params(ps, mu, norm, 0.0, 10.0)
    f2 = torch.tensor(f_host).requires_grad_(True)
    u2 = _oracle_unrolled(f2, g, prm, ITERS, transpose_scale(ps))
    (0.5 * ((u2 - torch.tensor(target)) ** 2).sum()).backward()

    e_u = rel_err(u1.detach().cpu().numpy(), u2.detach().numpy())
    e_g = rel_err(f1.grad.cpu().numpy(), f2.grad.numpy())
    record_parity("config4", "unrolled_tv100_grad_vs_oracle_radon", mu=mu, e_u=e_u, e_grad=e_g,
                  clamped_fraction=float((u2.detach().numpy() == 0).mean()))
    assert e_u < 1e-11, e_u
    assert e_g < 1e-9, e_g
