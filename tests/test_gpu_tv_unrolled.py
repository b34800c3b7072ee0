"""Hand-written adjoint of the unrolled TV primal-dual solver (float64):
against torch autograd through the same iteration written with the
differentiable projection operators, and against central finite differences
(h = 1e-6, 32 random directions, counting directions that cross a kink)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd.autograd import backproject, project
from gratopy_amd.tv_unrolled import _Params, tv_unrolled
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

# ||R||_E of the geometries below (float64 power iteration), so the
# primal-dual step sizes satisfy sigma*tau*||K||^2 <= 1
NORM = {("radon", 48): 2.455, ("fan", 48): 24.04, ("radon", 40): 2.4545, ("fan", 40): 21.94}


def _ps(geometry, N, Na, Ns):
    if geometry == "radon":
        return gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), Na, Ns)
    return gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, N), Na, Ns, R=800.0, RE=400.0,
                                      detector_width=2.0 * N, image_width=float(N))


def _torch_unrolled(f, ps, prm, n_iter):
    """The iteration in torch ops (gradient via torch autograd)."""
    def grad3(x):
        g = torch.zeros((3,) + x.shape, dtype=x.dtype, device=x.device)
        g[0, :-1] = x[1:] - x[:-1]
        g[1, :, :-1] = x[:, 1:] - x[:, :-1]
        return g

    def div3(p):
        d = torch.zeros(p.shape[1:], dtype=p.dtype, device=p.device)
        d[0] = p[0, 0]
        d[1:-1] += p[0, 1:-1] - p[0, :-2]
        d[-1] += -p[0, -2]
        d[:, 0] += p[1, :, 0]
        d[:, 1:-1] += p[1, :, 1:-1] - p[1, :, :-2]
        d[:, -1] += -p[1, :, -2]
        return d

    shape = tuple(ps.img_shape) + tuple(f.shape[2:])
    u = torch.zeros(shape, dtype=f.dtype, device=f.device)
    ubar = torch.zeros_like(u)
    v = torch.zeros((3,) + shape, dtype=f.dtype, device=f.device)
    lam = torch.zeros_like(f)
    for _ in range(n_iter):
        a = v + prm.sigma * grad3(ubar)
        norm = torch.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)
        v = a / torch.where(norm > 1, norm, torch.ones_like(norm))
        lam = (lam + prm.sigma_p * (project(ubar, ps) - f)) * prm.c
        u_new = torch.relu(u + prm.tau * (div3(v) - prm.n * backproject(lam, ps)))
        ubar = 2 * u_new - u
        u = u_new
    return u


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_adjoint_matches_torch_autograd(geometry):
    print(f"tv unrolled vs torch {geometry}", flush=True)
    ps = _ps(geometry, 48, 40, 70)
    rng = np.random.default_rng(0)
    f0 = torch.tensor(rng.random(ps.sinogram_shape + (4,)), device="cuda:0")
    target = torch.tensor(rng.random(ps.img_shape + (4,)), device="cuda:0")
    n_iter, mu, norm = 30, 40.0, NORM[(geometry, 48)]

    f1 = f0.clone().requires_grad_(True)
    u1 = tv_unrolled(f1, ps, mu, n_iter, norm_estimate=norm, checkpoint=7)
    (0.5 * ((u1 - target) ** 2).sum()).backward()

    prm = _Params(ps, mu, norm, 0.0, 10.0)
    f2 = f0.clone().requires_grad_(True)
    u2 = torch.stack([_torch_unrolled(f2[:, :, z], ps, prm, n_iter) for z in range(4)], dim=-1)
    (0.5 * ((u2 - target) ** 2).sum()).backward()

    assert rel_err(u1.detach().cpu().numpy(), u2.detach().cpu().numpy()) < 1e-11
    assert rel_err(f1.grad.cpu().numpy(), f2.grad.cpu().numpy()) < 1e-9


def test_unrolled_forward_equals_total_variation():
    ps = _ps("fan", 40, 30, 60)
    f = gratopy.forwardprojection(dev.to_device(np.asfortranarray(np.random.default_rng(3).random((40, 40, 2))),
                                                "cuda:0"), ps)
    u_a = tv_unrolled(f, ps, 25.0, 17, norm_estimate=NORM[("fan", 40)])
    u_b = gratopy.total_variation(f, ps, 25.0, 17, slice_thickness=0.0, norm_estimate=NORM[("fan", 40)])
    assert torch.equal(u_a.cpu(), u_b.cpu())


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_adjoint_finite_differences(geometry):
    print(f"tv unrolled fd {geometry}", flush=True)
    ps = _ps(geometry, 40, 30, 60)
    rng = np.random.default_rng(1)
    f0 = torch.tensor(rng.random(ps.sinogram_shape + (3,)), device="cuda:0")
    target = torch.tensor(rng.random(ps.img_shape + (3,)), device="cuda:0")
    n_iter, mu, norm = 25, 30.0, NORM[(geometry, 40)]

    def loss(f):
        return float(0.5 * ((tv_unrolled(f, ps, mu, n_iter, norm_estimate=norm) - target) ** 2).sum())

    f = f0.clone().requires_grad_(True)
    u = tv_unrolled(f, ps, mu, n_iter, norm_estimate=norm, checkpoint=10)
    (0.5 * ((u - target) ** 2).sum()).backward()
    g = f.grad
    h = 1e-6
    bad = 0
    for k in range(32):
        d = torch.tensor(np.random.default_rng(50 + k).standard_normal(f0.shape), device="cuda:0")
        fd = (loss(f0 + h * d) - loss(f0 - h * d)) / (2 * h)
        an = float((g * d).sum())
        if abs(fd - an) > 1e-5 * max(abs(an), 1e-12):
            bad += 1  # a direction whose +-h segment crosses a kink of max(0,.) or the ball
    assert bad <= 2, bad
