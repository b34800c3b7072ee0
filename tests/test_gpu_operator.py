"""The operator interface on the GPU: Radon/Fanbeam application equals the
function API, host arrays are moved to the device, outputs are reused, and
composites apply right to left (the reference's tests/test_operator.py)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd.operator import Fanbeam, Radon
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def test_radon_operator_equals_function_api():
    N, A = 64, 40
    R = Radon(image_domain=N, angles=A)
    ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), A, R.detectors.number)
    img = np.random.default_rng(0).random((N, N)).astype(np.float32)
    dimg = dev.to_device(np.asfortranarray(img), "cuda:0")
    sino_op = R.apply_to(dimg)
    sino_fn = gratopy.forwardprojection(dimg, ps)
    assert rel_err(dev.to_host(sino_op), dev.to_host(sino_fn)) < 1e-6
    back_op = R.T * sino_op
    back_fn = gratopy.backprojection(sino_fn, ps)
    assert rel_err(dev.to_host(back_op), dev.to_host(back_fn)) < 1e-6
    gram = (R.T * R).apply_to(dimg)
    assert rel_err(dev.to_host(gram), dev.to_host(back_fn)) < 1e-6
    scaled = (2.5 * R).apply_to(dimg)
    assert rel_err(dev.to_host(scaled), 2.5 * dev.to_host(sino_fn)) < 1e-6


def test_numpy_coercion_and_queue_reuse():
    R = Radon(image_domain=16, angles=10)
    sino = R.apply_to(np.zeros((16, 16), np.float32), queue="cuda:0")
    assert isinstance(sino, torch.Tensor) and tuple(sino.shape) == (R.detectors.number, 10)
    assert float(sino.abs().max()) == 0.0
    assert R.projection_settings is not None
    sino2 = R.apply_to(np.ones((16, 16), np.float32))  # reuses the remembered queue
    assert float(sino2.abs().max()) > 0


def test_outputs_are_reused():
    R = Radon(image_domain=16, angles=10)
    img = torch.zeros(16, 16, device="cuda:0")
    out = torch.full(R.output_shape, 42.0, device="cuda:0")
    res = R.apply_to(img, output=out)
    assert res.data_ptr() == out.data_ptr() and float(out.abs().max()) == 0.0
    out2 = torch.full((16, 16), 42.0, device="cuda:0")
    res2 = (R.T * R).apply_to(img, output=out2)
    assert res2.data_ptr() == out2.data_ptr() and float(out2.abs().max()) == 0.0
    ones = torch.ones(16, 16, device="cuda:0")
    expected = (R + R).apply_to(ones)
    out3 = torch.full(R.output_shape, 42.0, device="cuda:0")
    res3 = (R + R).apply_to(ones, output=out3)
    assert res3.data_ptr() == out3.data_ptr()
    assert rel_err(out3.cpu().numpy(), expected.cpu().numpy()) < 1e-6
    with pytest.raises(ValueError, match="Output shape mismatch"):
        R.apply_to(img, output=torch.zeros(3, 3, device="cuda:0"))
    with pytest.raises(ValueError, match="Input shape mismatch"):
        R.apply_to(torch.zeros(15, 16, device="cuda:0"))


def test_operator_3d_and_float64():
    R = Radon(image_domain=(20, 28), angles=15, detectors=33)
    img = torch.rand(20, 28, 3, dtype=torch.float64, device="cuda:0")
    sino = R * img
    assert tuple(sino.shape) == (33, 15, 3) and sino.dtype == torch.float64
    for z in range(3):
        assert rel_err((R * img[:, :, z].contiguous()).cpu().numpy(), sino[:, :, z].cpu().numpy()) < 1e-13


def test_fanbeam_operator_equals_function_api():
    F = Fanbeam((800.0, 400.0), 48, 60, 70)
    ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (48, 48), 60, 70, R=800.0, RE=400.0,
                                    detector_width=2.0)
    img = torch.rand(48, 48, device="cuda:0")
    assert rel_err((F * img).cpu().numpy(), gratopy.forwardprojection(img, ps).cpu().numpy()) < 1e-6
    sino = torch.rand(70, 60, device="cuda:0")
    assert rel_err((F.T * sino).cpu().numpy(), gratopy.backprojection(sino, ps).cpu().numpy()) < 1e-6
