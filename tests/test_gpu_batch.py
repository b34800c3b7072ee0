"""Batched uneven slices (gt_batch_*): every slice's result equals, bit for
bit, the single-slice function with the same settings (same kernel bodies,
same arithmetic), for projections, Landweber and TV, eager and graph-captured,
under plain launches and under chunked forward reads with the LDS-staged
backprojection; the single-slice path itself is pinned to the oracle
elsewhere."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd import _hip
from gratopy_amd.batch import (BatchPlan, batch_backprojection, batch_forwardprojection,
                               batch_landweber, batch_total_variation)

pytestmark = pytest.mark.gpu

# launch variants a plan can be built with (batches take no angle bands)
BATCH_VARIANTS = {
    "K4-RY1": dict(fwd_group=4, back_rows=1, fwd_chunk=1, back_stage=0, back_tile=16),
    "K8xC4-RY4xS16T32": dict(fwd_group=8, back_rows=4, fwd_chunk=4, back_stage=16, back_tile=32),
    "K16xC4-RY8xS8T16": dict(fwd_group=16, back_rows=8, fwd_chunk=4, back_stage=8, back_tile=16),
    "K8xB8xC4-RY8xS16T32": dict(fwd_group=8, fwd_bands=8, back_rows=8, fwd_chunk=4, back_stage=16,
                                back_tile=32),
    "K4xB4-RY2": dict(fwd_group=4, fwd_bands=4, back_rows=2, fwd_chunk=1, back_stage=0, back_tile=16),
}


@pytest.fixture(params=list(BATCH_VARIANTS))
def variant(request):
    from gratopy_amd import autotune

    before = autotune.current_variant()
    autotune.set_variant(**{"fwd_bands": 1, "back_bands": 1, **BATCH_VARIANTS[request.param]})
    yield request.param
    autotune.set_variant(**before)


def _mix(dtype, n=7, seed=0):
    """n slices alternating Radon / fanbeam, uneven sizes, angles, detectors."""
    rng = np.random.default_rng(seed)
    settings, imgs = [], []
    for i in range(n):
        N = int(rng.choice([17, 32, 48, 64, 96]))
        M = N if i % 3 else int(N * 3 // 4)
        Na = int(rng.choice([5, 12, 30, 61]))
        Ns = int(np.ceil(1.5 * N))
        if i % 2 == 0:
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, M), Na, Ns)
        else:
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, M), Na, Ns, R=2.0 * N, RE=1.0 * N,
                                            image_width=float(N), detector_width=2.0 * N)
        settings.append(ps)
        imgs.append(dev.to_device(np.asfortranarray(rng.random((N, M)).astype(dtype)), "cuda:0"))
    return settings, imgs


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_projections_equal_single_slice(dtype, variant):
    print(f"batch projections {np.dtype(dtype).name} {variant}", flush=True)
    settings, imgs = _mix(dtype)
    sinos = batch_forwardprojection(imgs, settings)
    backs = batch_backprojection(sinos, settings)
    torch.cuda.synchronize()
    for ps, img, s, b in zip(settings, imgs, sinos, backs):
        ref_s = gratopy.forwardprojection(img, ps)
        ref_b = gratopy.backprojection(ref_s, ps)
        assert torch.equal(s.cpu(), ref_s.cpu())
        assert torch.equal(b.cpu(), ref_b.cpu())


@pytest.mark.parametrize("graph", [False, True])
def test_batch_landweber_equals_single_slice(graph, variant):
    print(f"batch landweber graph={graph} {variant}", flush=True)
    settings, imgs = _mix(np.float32, seed=1)
    sinos = [gratopy.forwardprojection(img, ps) for ps, img in zip(settings, imgs)]
    norms = [gratopy.normest(ps) for ps in settings]
    got = batch_landweber(sinos, settings, 12, w=0.8, norm_estimates=norms, graph=graph)
    for ps, f, u, nrm in zip(settings, sinos, got, norms):
        ref = gratopy.landweber(f, ps, 12, w=0.8, norm_estimate=nrm)
        assert torch.equal(u.cpu(), ref.cpu())


@pytest.mark.parametrize("n_iter", [1, 6, 7])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_tv_equals_single_slice(n_iter, dtype, variant):
    print(f"batch tv {n_iter} {np.dtype(dtype).name} {variant}", flush=True)
    settings, imgs = _mix(dtype, seed=2)
    sinos = [gratopy.forwardprojection(img, ps) for ps, img in zip(settings, imgs)]
    norms = [gratopy.normest(ps) for ps in settings]
    for graph in (False, True):
        got = batch_total_variation(sinos, settings, 20.0, n_iter, norm_estimates=norms, graph=graph)
        for ps, f, u, nrm in zip(settings, sinos, got, norms):
            ref = gratopy.total_variation(f, ps, 20.0, n_iter, slice_thickness=0.0, norm_estimate=nrm)
            assert torch.isfinite(ref).all()
            assert torch.equal(u.cpu(), ref.cpu())


def test_batch_tv_on_a_plain_plan_is_rejected():
    settings, imgs = _mix(np.float32, n=2, seed=3)
    sinos = [gratopy.forwardprojection(img, ps) for ps, img in zip(settings, imgs)]
    plan = BatchPlan(settings, imgs, sinos)
    with pytest.raises(ValueError, match="TV needs"):
        plan.launch(_hip.BATCH_TV_EVEN, 1.0)
    with pytest.raises(ValueError, match="Landweber needs"):
        plan.launch(_hip.BATCH_LANDWEBER)
    plan.destroy()
