"""gt_projection_noalloc (the projections with the library's own device
workspace) against the entry points that take a caller workspace
(gt_{radon,fan}_{fwd,back}, reached through forwardprojection /
backprojection), through the C ABI: two calls in a row, the second smaller,
so the second runs on what the first left in the shared array."""

from __future__ import annotations

import numpy as np
import pytest

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from gratopy_amd import _hip

pytestmark = pytest.mark.gpu

# (img_shape, Nz, angles, detectors, order) of the first and the second call
CALLS = [((96, 80), 3, 90, 120, "F"), ((40, 33), 1, 25, 50, "C")]


def _settings(geometry, img_shape, na, ns, device):
    kw = {} if geometry == gratopy.RADON else dict(R=300.0, RE=150.0, detector_width=2.2 * img_shape[0])
    return gratopy.ProjectionSettings(device, geometry, img_shape, na, ns, **kw)


def _noalloc(ps, adjoint, out, inp, dtype):
    import torch

    lib = _hip.load_library()
    ps.ensure_dtype(dtype)
    Nx, Ny = ps.img_shape
    Ns, Na = ps.sinogram_shape
    Nz = int(out.shape[2]) if out.dim() > 2 else 1
    sdpd = ps.sdpd_buf[dtype].data_ptr() if ps.is_fan else None
    rc = lib.gt_projection_noalloc(
        torch.cuda.current_stream().cuda_stream, _hip.GEOM_FAN if ps.is_fan else _hip.GEOM_RADON,
        int(adjoint), _hip.dtype_code(dtype), out.data_ptr(), inp.data_ptr(),
        ps.ofs_buf[dtype].data_ptr(), sdpd, ps.geometry_information[dtype].data_ptr(),
        Nx, Ny, Ns, Na, Nz, dev.order_code(dev.order_of(out)), dev.order_code(dev.order_of(inp)))
    _hip.check(rc, "gt_projection_noalloc")


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("geometry", [gratopy.RADON, gratopy.FANBEAM])
def test_noalloc_matches_workspace_entry_points(hip_device, geometry, dtype):
    import torch

    rng = np.random.default_rng(7)
    for img_shape, nz, na, ns, order in CALLS:
        print(f"noalloc {geometry.value} {dtype} img {img_shape}x{nz} angles {na} det {ns} {order}",
              flush=True)
        ps = _settings(geometry, img_shape, na, ns, hip_device)
        shape = img_shape + ((nz,) if nz > 1 else ())
        img = dev.to_device(rng.random(shape).astype(dtype), hip_device, order=order)
        sino_ref = gratopy.forwardprojection(img, ps)
        sino_in = dev.to_device(rng.random(ps.sinogram_shape + shape[2:]).astype(dtype), hip_device,
                                order=order)
        back_ref = gratopy.backprojection(sino_in, ps)
        torch.cuda.synchronize()
        sino = torch.full_like(sino_ref, float("nan"))
        back = torch.full_like(back_ref, float("nan"))
        before = torch.cuda.memory_allocated()
        _noalloc(ps, False, sino, img, np.dtype(dtype))
        _noalloc(ps, True, back, sino_in, np.dtype(dtype))
        torch.cuda.synchronize()
        assert torch.cuda.memory_allocated() == before
        # same kernels and launch shapes as the workspace entry points: bitwise equal
        assert torch.equal(sino, sino_ref)
        assert torch.equal(back, back_ref)

