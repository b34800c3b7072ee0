"""The compacted banded forward (GT_TUNE_FWD_COMPACT = 1: a prologue lists,
per row band and angle, the bin groups whose strips can meet the band, and
only those are launched) gives the uncompacted result bit for bit: a group
left out of a band would have added exact zeros there, so any group the
prologue drops wrongly shows up as a difference.  Geometries: the benchmark
stacks' geometries (2048^2, 2560 detectors, fewer angles), shifted,
reversed-detector, non-square and ragged ones, fan bins wide enough that
bin groups fall back to per-bin rays, and more bands than image rows."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd import autotune

pytestmark = pytest.mark.gpu

CASES = [
    ("radon", (2048, 2048), 45, 2560, dict(detector_width=2048 * 2 ** 0.5, image_width=2048.0)),
    ("fan", (2048, 2048), 45, 2560, dict(R=4096.0, RE=2048.0, detector_width=4096.0, image_width=2048.0)),
    ("radon", (300, 170), 37, 400, dict(detector_width=2.5, midpoint_shift=(0.3, -0.2), detector_shift=0.4)),
    ("fan", (300, 170), 37, 400, dict(R=900.0, RE=500.0, detector_width=700.0, detector_shift=20.0,
                                      midpoint_shift=(15.0, -7.0), reverse_detector=True)),
    ("radon", (40, 33), 17, 61, {}),
    ("fan", (40, 33), 17, 61, dict(R=160.0, RE=70.0, detector_width=120.0)),
    ("fan", (5, 29), 3, 1, dict(R=800.0, RE=400.0, detector_width=300.0)),
    ("fan", (17, 3), 9, 40, dict(R=90.0, RE=40.0, detector_width=60.0, detector_shift=3.0)),
]

VARIANTS = [dict(fwd_group=k, fwd_bands=b, fwd_chunk=c)
            for k, b, c in ((2, 8, 1), (4, 32, 1), (8, 8, 4), (16, 8, 4), (16, 16, 4), (16, 32, 4),
                            (32, 8, 4))]


@pytest.fixture
def restore_variant():
    before = autotune.current_variant()
    yield
    autotune.set_variant(**before)


@pytest.mark.parametrize("variant", VARIANTS, ids=[autotune.variant_key(v) for v in VARIANTS])
def test_compacted_forward_is_bitwise_the_uncompacted_one(variant, restore_variant):
    print(f"compact vs full, {autotune.variant_key(variant)}", flush=True)
    for geometry, shape, na, nd, kw in CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        nz = 1 if shape[0] >= 2048 else 3
        host = np.random.default_rng(5).random(shape + (nz,))
        for dtype in (np.float32, np.float64):
            if shape[0] >= 2048 and dtype == np.float64:
                continue
            img = dev.to_device(np.asfortranarray(host.astype(dtype)), "cuda:0")
            out = []
            for compact in (0, 1):
                autotune.set_variant(**variant, fwd_compact=compact)
                out.append(gratopy.forwardprojection(img, ps))
            torch.cuda.synchronize()
            assert torch.equal(out[0], out[1]), (geometry, shape, np.dtype(dtype).name,
                                                 float((out[0] - out[1]).abs().max()))
