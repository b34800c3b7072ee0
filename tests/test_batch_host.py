"""Host-side validation of the batch API (no device needed: every case here
is rejected before the library is called)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd.batch import BatchPlan


def _ps(N=8, Na=5, Ns=12):
    return gratopy.ProjectionSettings(None, gratopy.RADON, (N, N), Na, Ns)


def _img(shape, dtype=np.float32):
    return dev.to_device(np.asfortranarray(np.zeros(shape, dtype)), "cpu", order="F")


def test_lengths_must_match():
    ps = _ps()
    with pytest.raises(ValueError, match="equal length"):
        BatchPlan([ps, ps], [_img((8, 8))], [_img((12, 5))])
    with pytest.raises(ValueError, match="equal length"):
        BatchPlan([], [], [])


def test_slices_must_be_2d_fortran():
    ps = _ps()
    three_d = torch.zeros((8, 8, 2), dtype=torch.float32)
    with pytest.raises(ValueError, match="2-D Fortran"):
        BatchPlan([ps], [three_d], [_img((12, 5))])
    non_f = torch.zeros((8, 6), dtype=torch.float32)  # C-contiguous, not F-contiguous
    with pytest.raises(ValueError, match="2-D Fortran"):
        BatchPlan([gratopy.ProjectionSettings(None, gratopy.RADON, (8, 6), 5, 12)], [non_f],
                  [_img((12, 5))])


def test_shapes_and_dtypes_must_match_settings():
    ps = _ps()
    with pytest.raises(ValueError, match="shapes do not match"):
        BatchPlan([ps], [_img((8, 9))], [_img((12, 5))])
    with pytest.raises(ValueError, match="one dtype"):
        BatchPlan([ps], [_img((8, 8))], [_img((12, 5), np.float64)])


def test_every_operand_is_checked_before_the_library_is_called():
    ps = _ps()
    img, sino = _img((8, 8)), _img((12, 5))
    # well-formed but on the CPU: rejected, no host pointer reaches a kernel
    with pytest.raises(ValueError, match="HIP device"):
        BatchPlan([ps], [img], [sino])
    # measured sinogram with fewer angles than the settings
    with pytest.raises(ValueError, match="data shapes do not match"):
        BatchPlan([ps], [img], [sino], datas=[_img((12, 4))])
    # a float64 second image in a float32 batch
    with pytest.raises(ValueError, match="one dtype"):
        BatchPlan([ps], [img], [sino], imgs2=[_img((8, 8), np.float64)])
    # TV dual variable too small for the slice
    with pytest.raises(ValueError, match="4 \\* Nx \\* Ny"):
        BatchPlan([ps], [img], [sino], fields=[torch.zeros(4 * 8 * 8 - 1)])
    with pytest.raises(ValueError, match="one entry per slice"):
        BatchPlan([ps], [img], [sino], datas=[])
