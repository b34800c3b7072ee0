"""Autograd operators (float64, Nx=512, 360 angles, 768 detectors, Radon and
fanbeam): gradients of 0.5||Rx - y||^2 + lam*TV_eps(x) (lam=1e-3, eps=1e-4)
against central finite differences (h=1e-6, 32 random directions) and the
data-term gradient against the one assembled from the oracle's R and B."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd.autograd import backproject, project, transpose_scale, tv_eps
from oracle import solvers as osolv
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _ps(geometry, N=512):
    if geometry == "radon":
        return gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), 360, 768)
    return gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, N), 360, 768, R=800.0, RE=400.0,
                                      detector_width=1024.0, image_width=float(N))


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_gradient_matches_oracle_adjoint(geometry):
    print(f"autograd oracle {geometry}", flush=True)
    ps = _ps(geometry)
    g = osolv.OracleGeometry.from_settings(ps)
    rng = np.random.default_rng(0)
    x0 = rng.random(ps.img_shape)
    y0 = rng.random(ps.sinogram_shape)
    x = dev.to_device(np.asfortranarray(x0), "cuda:0").requires_grad_(True)
    y = dev.to_device(np.asfortranarray(y0), "cuda:0")
    loss = 0.5 * ((project(x, ps) - y) ** 2).sum()
    loss.backward()
    r = g.fwd(np.asfortranarray(x0)) - y0
    w = ps.angle_weights
    ref = transpose_scale(ps) * g.back(np.asfortranarray(r / w[None, :]))
    assert rel_err(x.grad.cpu().numpy(), ref) < 1e-12

    # backprojection's gradient: B^T (B s - t) = W R(B s - t) / c
    s = dev.to_device(np.asfortranarray(y0), "cuda:0").requires_grad_(True)
    t = dev.to_device(np.asfortranarray(x0), "cuda:0")
    loss2 = 0.5 * ((backproject(s, ps) - t) ** 2).sum()
    loss2.backward()
    resid = g.back(np.asfortranarray(y0)) - x0
    ref2 = w[None, :] * g.fwd(np.asfortranarray(resid)) / transpose_scale(ps)
    assert rel_err(s.grad.cpu().numpy(), ref2) < 1e-12


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_gradient_finite_differences(geometry):
    print(f"autograd fd {geometry}", flush=True)
    ps = _ps(geometry)
    rng = np.random.default_rng(1)
    x0 = torch.tensor(rng.random(ps.img_shape), device="cuda:0", dtype=torch.float64)
    y = gratopy.forwardprojection(dev.to_device(np.asfortranarray(rng.random(ps.img_shape)), "cuda:0"), ps)
    lam, eps = 1e-3, 1e-4

    def loss(x):
        return 0.5 * ((project(x, ps) - y) ** 2).sum() + lam * tv_eps(x, eps)

    x = x0.clone().requires_grad_(True)
    L = loss(x)
    L.backward()
    grad = x.grad.detach()
    h = 1e-6
    worst = 0.0
    for k in range(32):
        d = torch.tensor(np.random.default_rng(100 + k).standard_normal(ps.img_shape), device="cuda:0")
        with torch.no_grad():
            fd = (loss(x0 + h * d) - loss(x0 - h * d)) / (2 * h)
        an = float((grad * d).sum())
        worst = max(worst, abs(float(fd) - an) / max(abs(an), 1e-30))
    assert worst < 1e-5, worst


def test_adjoint_pairing_of_autograd_operators():
    """<R x, y>_E = <x, R^T y>_E computed through autograd (float64)."""
    ps = _ps("fan", 96)
    rng = np.random.default_rng(2)
    x = torch.tensor(rng.random(ps.img_shape), device="cuda:0", dtype=torch.float64, requires_grad=True)
    yv = torch.tensor(rng.random(ps.sinogram_shape), device="cuda:0", dtype=torch.float64)
    (project(x, ps) * yv).sum().backward()
    lhs = float((project(x.detach(), ps) * yv).sum())
    rhs = float((x.detach() * x.grad).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-13
