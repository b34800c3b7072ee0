"""Parity at the benchmark's full size (2048^2 image, 1440 angles, 2560
detectors; fanbeam as in bench.py's default workload and parallel beam as in
its radon workload), float32, the library's automatic launch shapes:

* the forward projection of every 16th angle against the oracle run on those
  angles alone (angles are independent in the forward), float64 oracle as the
  truth, within 1e-5 or 1.5x the reference algorithm's own float32 distance;
* the reference's adjointness pairing (tests/test_fanbeam.py:358-449,
  tests/test_radon.py:340-424) on random full-size pairs;
* linearity, R(2x + 3y) = 2Rx + 3Ry, and the same for the backprojection;
* one slice of a 2-slice stack bitwise equal to that slice projected alone
  (one launch variant pinned)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd import autotune
from tests.test_gpu_projection import _oracle

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")
N, NA, NS = 2048, 1440, 2560


def _settings(geometry, angles=NA):
    if geometry == "fan":
        return gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, N), angles, NS, R=4096.0,
                                          RE=2048.0, image_width=2048.0, detector_width=4096.0)
    return gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), angles, NS, image_width=2048.0,
                                      detector_width=2048.0 * 2 ** 0.5)


@pytest.fixture
def automatic():
    before = autotune.current_variant()
    autotune.set_variant(automatic=1)
    yield
    autotune.set_variant(**before)


@pytest.mark.parametrize("geometry", ["fan", "radon"])
def test_full_size_forward_against_the_oracle(geometry, automatic):
    print(f"full size forward {geometry}", flush=True)
    ps = _settings(geometry)
    img = np.asfortranarray(np.random.default_rng(11).random((N, N)).astype(F32))
    got = dev.to_host(gratopy.forwardprojection(dev.to_device(img, "cuda:0"), ps))[:, ::16]
    sub = _settings(geometry, angles=np.asarray(ps.angles)[::16])
    truth = _oracle(sub, img.astype(F64), F64, False)
    scale = np.abs(truth).max()
    e_ref = np.abs(_oracle(sub, img, F32, False) - truth).max() / scale
    e_hip = np.abs(got - truth).max() / scale
    assert e_hip <= max(1e-5, 1.5 * e_ref), (e_hip, e_ref)


@pytest.mark.parametrize("geometry", ["fan", "radon"])
def test_full_size_adjointness_and_linearity(geometry, automatic):
    print(f"full size adjointness {geometry}", flush=True)
    ps = _settings(geometry)
    rng = np.random.default_rng(12)
    w = torch.tensor(ps.angle_weights, device="cuda:0")
    xs = [dev.to_device(np.asfortranarray(rng.random((N, N)).astype(F32)), "cuda:0") for _ in range(2)]
    ys = [dev.to_device(np.asfortranarray(rng.random((NS, NA)).astype(F32)), "cuda:0") for _ in range(2)]
    for x, y in zip(xs, ys):
        Rx = gratopy.forwardprojection(x, ps)
        Ry = gratopy.backprojection(y, ps)
        lhs = float((Rx.double() * y.double() * w[None, :]).sum()) * ps.delta_s
        rhs = float((x.double() * Ry.double()).sum()) * ps.delta_x ** 2
        assert abs(lhs - rhs) / min(abs(lhs), abs(rhs)) < 1e-5
    fx = [gratopy.forwardprojection(x, ps) for x in xs]
    comb = gratopy.forwardprojection(2 * xs[0] + 3 * xs[1], ps)
    ref = 2 * fx[0] + 3 * fx[1]
    assert float((comb - ref).abs().max()) <= 1e-5 * float(ref.abs().max())
    by = [gratopy.backprojection(y, ps) for y in ys]
    comb = gratopy.backprojection(2 * ys[0] + 3 * ys[1], ps)
    ref = 2 * by[0] + 3 * by[1]
    assert float((comb - ref).abs().max()) <= 1e-5 * float(ref.abs().max())


@pytest.mark.parametrize("geometry", ["fan", "radon"])
def test_full_size_stack_slice_equals_single_slice(geometry):
    print(f"full size stack {geometry}", flush=True)
    before = autotune.current_variant()
    try:
        autotune.set_variant(fwd_group=16, fwd_bands=8, fwd_chunk=4, back_rows=8, back_bands=1,
                             back_stage=16, back_tile=32)
        ps = _settings(geometry)
        rng = np.random.default_rng(13)
        stack = dev.to_device(np.asfortranarray(rng.random((N, N, 2)).astype(F32)), "cuda:0")
        one = dev.to_device(np.asfortranarray(dev.to_host(stack)[:, :, 1]), "cuda:0")
        s2 = gratopy.forwardprojection(stack, ps)
        s1 = gratopy.forwardprojection(one, ps)
        assert torch.equal(s2[:, :, 1], s1)
        b2 = gratopy.backprojection(s2, ps)
        b1 = gratopy.backprojection(s1, ps)
        torch.cuda.synchronize()
        assert torch.equal(b2[:, :, 1], b1)
    finally:
        autotune.set_variant(**before)
