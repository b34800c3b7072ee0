"""Solvers replayed as captured HIP graphs equal the eager launches bit for
bit (same kernels, same order, deterministic reductions).  Each solver is run
eagerly and graphed several times in a row, with eager work in between, and
every replay is compared."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev

pytestmark = pytest.mark.gpu


# ||R||_E of the two geometries below (float64 power iteration), so the
# iterations stay bounded and the bitwise comparisons see finite values
NORM = {"radon": 2.28, "fan": 27.9}


def _setup(geometry, dtype=np.float32, nz=2):
    if geometry == "radon":
        ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (64, 48), 40, 80)
    else:
        ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (64, 48), 50, 80, R=400.0, RE=200.0,
                                        detector_width=150.0)
    rng = np.random.default_rng(7)
    img = dev.to_device(np.asfortranarray(rng.random((64, 48, nz)).astype(dtype)), "cuda:0")
    sino = gratopy.forwardprojection(img, ps)
    return ps, sino


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_landweber_graph_bitwise(geometry):
    print(f"graph landweber {geometry}", flush=True)
    ps, sino = _setup(geometry)
    ref = gratopy.landweber(sino, ps, 13, norm_estimate=NORM[geometry]).cpu()
    assert torch.isfinite(ref).all()
    for _ in range(3):
        got = gratopy.landweber(sino, ps, 13, norm_estimate=NORM[geometry], graph=True).cpu()
        assert torch.equal(got, ref)
        gratopy.backprojection(sino, ps)  # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("geometry", ["radon", "fan"])
@pytest.mark.parametrize("n", [1, 10, 11])
def test_tv_graph_bitwise(geometry, n):
    print(f"graph tv {geometry} {n}", flush=True)
    ps, sino = _setup(geometry)
    ref = gratopy.total_variation(sino, ps, 30.0, n, slice_thickness=0.0,
                                  norm_estimate=NORM[geometry]).cpu()
    assert torch.isfinite(ref).all()
    for _ in range(2):
        got = gratopy.total_variation(sino, ps, 30.0, n, slice_thickness=0.0,
                                      norm_estimate=NORM[geometry], graph=True).cpu()
        assert torch.equal(got, ref)
        gratopy.forwardprojection(gratopy.backprojection(sino, ps), ps)
        torch.cuda.synchronize()


@pytest.mark.parametrize("n", [6, 7])
def test_cg_graph_bitwise(n):
    print(f"graph cg {n}", flush=True)
    ps, sino = _setup("fan", np.float64)
    ref = gratopy.conjugate_gradients(sino, ps, n).cpu()
    for _ in range(2):
        got = gratopy.conjugate_gradients(sino, ps, n, graph=True).cpu()
        assert torch.equal(got, ref)


def test_graph_on_explicit_stream():
    s = torch.cuda.Stream()
    ps = gratopy.ProjectionSettings(s, gratopy.RADON, (32, 32), 20, 40)
    img = dev.to_device(np.asfortranarray(np.random.default_rng(1).random((32, 32)).astype(np.float32)),
                        "cuda:0")
    with torch.cuda.stream(s):
        sino = gratopy.forwardprojection(img, ps)
        ref = gratopy.landweber(sino, ps, 6, norm_estimate=2.45)
        got = gratopy.landweber(sino, ps, 6, norm_estimate=2.45, graph=True)
    s.synchronize()
    assert torch.equal(got.cpu(), ref.cpu())
