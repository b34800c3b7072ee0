"""Operator kernels supplied as HIP sources (the reference's
tests/test_custom_kernels.py:17-110): a plugin is compiled at first use,
its forward and adjoint kernels are chosen by the naming ABI for every pair
of array orders, and compiled plugins are cached by source signature, so two
specs with the same base name never share kernels.  A brute-force Radon
plugin run through ``Radon(kernel_spec=...)`` must reproduce the built-in
transform, which checks that the geometry tables reach plugin kernels in the
reference's argument order."""

from __future__ import annotations

import shutil
from copy import copy
from pathlib import Path

import numpy as np
import pytest

from gratopy_amd import _device as dev
from gratopy_amd.operator import HipKernelSpec, Radon, invalidate_kernel_cache
from gratopy_amd.operator.hip import _PROGRAM_CACHE, _HipOperator
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

KERNELS = Path(__file__).parent / "kernels"
AFFINE = KERNELS / "affine.hip"
AFFINE_ALT = KERNELS / "affine_alt.hip"
RADON_BRUTEFORCE = KERNELS / "radon_bruteforce.hip"


class AffineOperator(_HipOperator):
    """Elementwise test operator: 2x+1 forward, 3x-1 adjoint (affine.hip)."""

    def __init__(self, shape, adjoint=False, kernel_spec=None):
        super().__init__(name="AffineOperator", state={"adjoint": adjoint}, input_shape=shape,
                         output_shape=shape, kernel_spec=kernel_spec)

    def _default_kernel_spec(self):
        return HipKernelSpec.from_path(AFFINE, base_name="affine")

    @property
    def adjoint(self):
        return self.state["adjoint"]

    @property
    def T(self):
        op = copy(self)
        op.state = copy(self.state)
        op.state["adjoint"] = not self.state["adjoint"]
        return op


def test_custom_kernel_spec_end_to_end(hip_device):
    shape = (4, 5)
    spec = HipKernelSpec.from_path(AFFINE, base_name="affine")
    op = AffineOperator(shape=shape, kernel_spec=spec)
    expected = 2 * np.arange(np.prod(shape), dtype=np.float32).reshape(shape) + 1
    for input_order in "CF":
        for output_order in "CF":
            host = np.require(np.arange(np.prod(shape), dtype=np.float32).reshape(shape),
                              requirements=input_order)
            dinput = dev.to_device(host, hip_device)
            doutput = dev.zeros(shape, np.float32, order=output_order, device=hip_device)
            result = op.apply_to(dinput, output=doutput)
            assert result is doutput
            assert dev.order_of(result) == output_order
            np.testing.assert_allclose(dev.to_host(result), expected)


def test_custom_kernel_spec_double_and_three_axes(hip_device):
    shape = (70, 5, 3)  # more than one block along the first axis
    op = AffineOperator(shape=shape)
    host = np.random.default_rng(1).random(shape)
    for order in "CF":
        result = op.apply_to(dev.to_device(np.require(host, requirements=order), hip_device))
        assert dev.order_of(result) == order
        np.testing.assert_array_equal(dev.to_host(result), 2 * host + 1)


def test_custom_kernel_spec_adjoint_kernel_is_used(hip_device):
    shape = (4, 5)
    op = AffineOperator(shape=shape)
    host = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    result = op.T.apply_to(host, queue=hip_device)
    np.testing.assert_allclose(dev.to_host(result), 3 * host - 1)


def test_custom_kernel_spec_scalar_and_composition(hip_device):
    shape = (4, 5)
    op = AffineOperator(shape=shape)
    host = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    np.testing.assert_allclose(dev.to_host((0.5 * op).apply_to(host, queue=hip_device)),
                               0.5 * (2 * host + 1))
    np.testing.assert_allclose(dev.to_host((op.T * op).apply_to(host, queue=hip_device)),
                               3 * (2 * host + 1) - 1)


def test_custom_kernel_spec_cache_isolation_by_source_signature(hip_device):
    shape = (4, 5)
    host = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    op1 = AffineOperator(shape=shape, kernel_spec=HipKernelSpec.from_path(AFFINE, base_name="affine"))
    op2 = AffineOperator(shape=shape,
                         kernel_spec=HipKernelSpec.from_path(AFFINE_ALT, base_name="affine"))
    before = dev.to_host(op1.apply_to(host, queue=hip_device))
    second = dev.to_host(op2.apply_to(host, queue=hip_device))
    after = dev.to_host(op1.apply_to(host, queue=hip_device))
    np.testing.assert_allclose(before, 2 * host + 1)
    np.testing.assert_allclose(second, 5 * host - 3)
    np.testing.assert_allclose(after, 2 * host + 1)


def test_changed_source_is_recompiled_and_invalidate_clears(hip_device, tmp_path):
    shape = (3, 4)
    host = np.arange(12, dtype=np.float32).reshape(shape)
    path = tmp_path / "affine.hip"
    shutil.copy(AFFINE, path)
    spec = HipKernelSpec.from_path(path, base_name="affine")
    op = AffineOperator(shape=shape, kernel_spec=spec)
    np.testing.assert_allclose(dev.to_host(op.apply_to(host, queue=hip_device)), 2 * host + 1)
    shutil.copy(AFFINE_ALT, path)  # same path, new content: a new signature
    np.testing.assert_allclose(dev.to_host(op.apply_to(host, queue=hip_device)), 5 * host - 3)
    assert len(_PROGRAM_CACHE) >= 2
    invalidate_kernel_cache()
    assert len(_PROGRAM_CACHE) == 0
    np.testing.assert_allclose(dev.to_host(op.apply_to(host, queue=hip_device)), 5 * host - 3)


def test_missing_kernel_variant_is_a_key_error(hip_device, tmp_path):
    # a plugin that defines only the forward kernels
    path = tmp_path / "forward_only.hip"
    text = AFFINE.read_text()
    path.write_text(text[: text.index("GT_EXPORT_VARIANTS(affine_ad")])
    op = AffineOperator(shape=(2, 2), kernel_spec=HipKernelSpec.from_path(path, base_name="affine"))
    host = np.ones((2, 2), np.float32)
    np.testing.assert_allclose(dev.to_host(op.apply_to(host, queue=hip_device)), 3.0)
    with pytest.raises(KeyError, match="affine_ad_float"):
        op.T.apply_to(host, queue=hip_device)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_bruteforce_radon_plugin_matches_builtin(hip_device, dtype, tol):
    N, A, Nz = 37, 23, 2
    spec = HipKernelSpec.from_path(RADON_BRUTEFORCE, base_name="radon")
    R = Radon(image_domain=N, angles=A)
    P = Radon(image_domain=N, angles=A, kernel_spec=spec)
    rng = np.random.default_rng(7)
    img = rng.random((N, N, Nz)).astype(dtype)
    sino = rng.random((R.detectors.number, A, Nz)).astype(dtype)
    for in_order in "CF":
        dimg = dev.to_device(np.require(img, requirements=in_order), hip_device)
        dsino = dev.to_device(np.require(sino, requirements=in_order), hip_device)
        for out_order in "CF":
            out = dev.zeros(tuple(R.output_shape) + (Nz,), dtype, order=out_order,
                            device=hip_device)
            got = dev.to_host(P.apply_to(dimg, output=out))
            want = dev.to_host(R.apply_to(dimg))
            assert rel_err(got, want) < tol, (in_order, out_order)
            out = dev.zeros((N, N, Nz), dtype, order=out_order, device=hip_device)
            got = dev.to_host(P.T.apply_to(dsino, output=out))
            want = dev.to_host(R.T.apply_to(dsino))
            assert rel_err(got, want) < tol, (in_order, out_order)
