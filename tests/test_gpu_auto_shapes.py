"""The automatic launch-shape choice (the library's default) across its
thresholds: stacks whose geometry, work Nx*Ny*Na*Nz and pixel count put each
direction on every branch (forward: bands only, banded 16-byte reads with 8 or
32 row bands, K = 16; backprojection: unstaged without (tiny calls) and with
angle bands, LDS-staged with and without angle
bands, tiles 16 and 32 wide) meet the parity bar slice by slice, and batch
plans in automatic mode agree with the single-slice path to rounding (each
picks its own shape, so they are not bitwise equal)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd import autotune
from gratopy_amd.batch import batch_backprojection, batch_forwardprojection
from tests.helpers import rel_err
from tests.test_gpu_projection import TOL, _oracle

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")


@pytest.fixture
def automatic():
    before = autotune.current_variant()
    autotune.set_variant(automatic=1)
    yield
    autotune.set_variant(**before)


# (geometry, N, angles, detectors, slices) -> forward / backprojection shape
STACKS = [
    ("radon", 64, 30, 90, 1),      # 1.2e5: bands, one-pixel reads / unstaged, no bands
    ("fan", 128, 90, 192, 16),     # work 2.4e7: bands only / unstaged, angle bands
    ("radon", 128, 120, 192, 8),   # 1.6e7: banded 16-byte reads / unstaged, angle bands
    ("fan", 256, 360, 384, 4),     # 9.4e7: banded 16-byte reads / staged, angle bands, 32 wide
    ("fan", 512, 360, 768, 1),     # 9.4e7, one 512^2 slice: 32 row bands / staged, angle bands
    ("radon", 256, 360, 384, 50),  # 1.2e9: banded 16-byte reads / staged, angle bands, 16 wide
    ("radon", 1024, 60, 1536, 9),  # 5.7e8, 9.4e6 pixels: staged without bands, 16 wide
    ("fan", 512, 720, 768, 85),    # 1.6e10: K = 16 / staged without bands, 32 wide
]


def _settings(geometry, N, na, nd):
    if geometry == "fan":
        return gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, N), na, nd, R=2.0 * N,
                                          RE=1.0 * N, image_width=float(N), detector_width=2.0 * N)
    return gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), na, nd)


@pytest.mark.parametrize("geometry,N,na,nd,nz", STACKS)
def test_automatic_shapes_meet_the_bar_slice_by_slice(geometry, N, na, nd, nz, automatic):
    print(f"auto {geometry} {N}^2 x {nz}, {na} angles, {nd} det: work {N * N * na * nz:.2e}", flush=True)
    ps = _settings(geometry, N, na, nd)
    rng = np.random.default_rng(9)
    img = np.asfortranarray(rng.random((N, N, nz)).astype(F32))
    sino = gratopy.forwardprojection(dev.to_device(img, "cuda:0"), ps)
    back = gratopy.backprojection(sino, ps)
    torch.cuda.synchronize()
    sino_h, back_h = dev.to_host(sino), dev.to_host(back)
    for z in (0, nz // 2, nz - 1):
        ref_s = _oracle(ps, img[:, :, z], F64, False)
        e_ref = rel_err(_oracle(ps, img[:, :, z], F32, False), ref_s)
        assert rel_err(sino_h[:, :, z], ref_s) <= max(TOL[F32], 1.2 * e_ref), ("forward", z)
        ref_b = _oracle(ps, sino_h[:, :, z].astype(F64), F64, True)
        e_ref = rel_err(_oracle(ps, sino_h[:, :, z], F32, True), ref_b)
        assert rel_err(back_h[:, :, z], ref_b) <= max(TOL[F32], 1.2 * e_ref), ("back", z)


def test_batch_plans_in_automatic_mode_agree_with_single_slices(automatic):
    rng = np.random.default_rng(4)
    settings, imgs = [], []
    for i, (N, na) in enumerate([(64, 30), (96, 61), (128, 90), (48, 17)]):
        settings.append(_settings("fan" if i % 2 else "radon", N, na, int(1.5 * N)))
        imgs.append(dev.to_device(np.asfortranarray(rng.random((N, N)).astype(F32)), "cuda:0"))
    sinos = batch_forwardprojection(imgs, settings)
    backs = batch_backprojection(sinos, settings)
    torch.cuda.synchronize()
    for ps, img, s, b in zip(settings, imgs, sinos, backs):
        ref_s = gratopy.forwardprojection(img, ps)
        assert rel_err(dev.to_host(s), dev.to_host(ref_s)) < 1e-5
        assert rel_err(dev.to_host(b), dev.to_host(gratopy.backprojection(s, ps))) < 1e-5
