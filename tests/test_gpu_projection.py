"""Parity of the HIP projection kernels with the CPU oracle (needs a GPU).

Bar: normwise max|hip - oracle| / max|oracle| <= 1e-5 in float32 and
<= 1e-12 in float64, every memory order, 2-D and 3-D; the reference's
control numbers (rtol 1e-3) through the drop-in API; adjointness in the
weighted inner products.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
import oracle
from gratopy_amd import _device as dev
from tests.helpers import assert_control, create_phantoms, phantom, rel_err

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")
TOL = {F32: 1e-5, F64: 1e-12}


def _ps(geometry, img_shape, angles, n_det, **kw):
    return gratopy.ProjectionSettings("cuda:0", geometry, img_shape, angles, n_det, **kw)


def _oracle(ps, x, dtype, adjoint, order="F"):
    t = ps.struct
    x = np.require(x, dtype, order)
    if ps.is_fan:
        if adjoint:
            return oracle.fan_back(x, t["ofs_dict"][dtype], t["sdpd_dict"][dtype], t["geo_dict"][dtype],
                                   ps.img_shape, img_order=order)
        return oracle.fan_fwd(x, t["ofs_dict"][dtype], t["sdpd_dict"][dtype], t["geo_dict"][dtype],
                              *ps.sinogram_shape, sino_order=order)
    if adjoint:
        return oracle.radon_back(x, t["ofs_dict"][dtype], t["geo_dict"][dtype], ps.img_shape, img_order=order)
    return oracle.radon_fwd(x, t["ofs_dict"][dtype], t["geo_dict"][dtype], *ps.sinogram_shape, sino_order=order)


def _hip(ps, x, dtype, adjoint, in_order="F", out_order="F"):
    xd = dev.to_device(np.require(x, dtype, in_order), "cuda:0", order=in_order)
    z = (x.shape[2],) if x.ndim > 2 else ()
    if adjoint:
        out = dev.zeros(ps.img_shape + z, dtype, order=out_order, device="cuda:0")
        gratopy.backprojection(xd, ps, img=out)
    else:
        out = dev.zeros(ps.sinogram_shape + z, dtype, order=out_order, device="cuda:0")
        gratopy.forwardprojection(xd, ps, sino=out)
    torch.cuda.synchronize()
    assert dev.order_of(out) == out_order or min(out.shape) == 1
    return dev.to_host(out)


GEOMETRIES = {
    "radon": (gratopy.RADON, {}),
    "radon_shift": (gratopy.RADON, dict(detector_width=2.5, midpoint_shift=(0.1, -0.2), detector_shift=0.1)),
    "fan": (gratopy.FANBEAM, dict(R=800.0, RE=400.0, detector_width=300.0)),
    "fan_shift": (gratopy.FANBEAM, dict(R=8.0, RE=3.0, detector_width=6.0, image_width=2.0,
                                        midpoint_shift=(0.2, 0.1), detector_shift=0.3, reverse_detector=True)),
}


@pytest.mark.parametrize("name", list(GEOMETRIES))
@pytest.mark.parametrize("dtype", [F32, F64])
def test_parity_all_orders(name, dtype):
    geometry, kw = GEOMETRIES[name]
    print(f"parity {name} {dtype}", flush=True)
    ps = _ps(geometry, (96, 72), 61, 110, **kw)
    rng = np.random.default_rng(1)
    img = rng.random((96, 72))
    sino = rng.random(ps.sinogram_shape)
    ref_f = _oracle(ps, img, dtype, False)
    ref_b = _oracle(ps, sino, dtype, True)
    for o_in in ("F", "C"):
        for o_out in ("F", "C"):
            assert rel_err(_hip(ps, img, dtype, False, o_in, o_out), ref_f) < TOL[dtype], (o_in, o_out)
            assert rel_err(_hip(ps, sino, dtype, True, o_in, o_out), ref_b) < TOL[dtype], (o_in, o_out)


@pytest.mark.parametrize("geometry", ["radon", "fan"])
@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("order", ["F", "C"])
def test_parity_3d_stack(geometry, dtype, order):
    g, kw = GEOMETRIES[geometry]
    print(f"3d {geometry} {dtype} {order}", flush=True)
    ps = _ps(g, (64, 80), 45, 90, **kw)
    rng = np.random.default_rng(2)
    img = rng.random((64, 80, 3))
    sino = rng.random(ps.sinogram_shape + (3,))
    assert rel_err(_hip(ps, img, dtype, False, order, order), _oracle(ps, img, dtype, False, order)) < TOL[dtype]
    assert rel_err(_hip(ps, sino, dtype, True, order, order), _oracle(ps, sino, dtype, True, order)) < TOL[dtype]


@pytest.mark.parametrize("dtype", [F32, F64])
def test_parity_fanbeam_config2(dtype):
    """The 512^2, 360-angle, 768-detector fanbeam workload (RE=400, R=800,
    detector width 2*Nx, image width Nx) on a random-ellipse phantom."""
    from gratopy_amd.phantom import random_ellipses

    print(f"config2 {dtype}", flush=True)
    ps = _ps(gratopy.FANBEAM, (512, 512), 360, 768, R=800.0, RE=400.0, detector_width=1024.0,
             image_width=512.0)
    img = random_ellipses(512, 15, seed=2, dtype=dtype)
    ref = _oracle(ps, img, dtype, False)
    got = _hip(ps, img, dtype, False)
    ref_b = _oracle(ps, ref, dtype, True)
    got_b = _hip(ps, ref, dtype, True)
    if dtype == F64:
        assert rel_err(got, ref) < TOL[dtype]
        assert rel_err(got_b, ref_b) < TOL[dtype]
        return
    # float32: at this size the reference's own float32 arithmetic is
    # ~4e-5 (forward) away from exact, so two float32 implementations cannot
    # agree to 1e-5.  Bar: the HIP float32 result is within 1e-5 of the
    # float64 oracle, or no farther from it than the reference algorithm in
    # float32 (the oracle in float32) is.
    truth = _oracle(ps, img, F64, False)
    truth_b = _oracle(ps, ref, F64, True)
    for name, g, r, t in (("forward", got, ref, truth), ("backward", got_b, ref_b, truth_b)):
        e_hip, e_ref = rel_err(g, t), rel_err(r, t)
        print(f"config2 float32 {name}: error vs float64 oracle: hip {e_hip:.3e}, "
              f"reference algorithm {e_ref:.3e}; hip vs float32 oracle {rel_err(g, r):.3e}", flush=True)
        assert e_hip <= max(TOL[F32], e_ref)


@pytest.mark.parametrize("shape,angles,nd", [((1, 1), 1, 1), ((3, 1), 2, 3), ((1, 5), 7, 2),
                                             ((17, 3), 1, 40), ((5, 29), 3, 1)])
@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_parity_ragged_sizes(shape, angles, nd, geometry):
    g, kw = GEOMETRIES[geometry]
    print(f"ragged {geometry} {shape} {angles} {nd}", flush=True)
    ps = _ps(g, shape, angles, nd, **kw)
    rng = np.random.default_rng(3)
    img = rng.random(shape)
    sino = rng.random(ps.sinogram_shape)
    for dtype in (F32, F64):
        ref_f = _oracle(ps, img, dtype, False)
        ref_b = _oracle(ps, sino, dtype, True)
        got_f = _hip(ps, img, dtype, False)
        got_b = _hip(ps, sino, dtype, True)
        # a projection that is identically zero is compared on the scale of its input
        scale_f = max(np.abs(ref_f).max(), 1e-3)
        scale_b = max(np.abs(ref_b).max(), 1e-3)
        assert np.abs(got_f - ref_f).max() <= TOL[dtype] * scale_f
        assert np.abs(got_b - ref_b).max() <= TOL[dtype] * scale_b


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_adjointness_gpu(geometry):
    """The reference's adjointness experiment (tests/test_radon.py:340-424,
    tests/test_fanbeam.py:358-449): 100 random pairs, fewer than 10 may exceed
    a relative pairing gap of 1e-5 in float32; float64 must hold to 1e-12."""
    print(f"adjoint {geometry}", flush=True)
    if geometry == "radon":
        ps = _ps(gratopy.RADON, (400, 400), 180, 230)
    else:
        ps = _ps(gratopy.FANBEAM, (400, 400), 360, 230, detector_width=83.0, R=900.0, RE=300.0)
    rng = np.random.default_rng(4)
    w = torch.tensor(ps.angle_weights, device="cuda:0")
    errors = 0
    for dtype in (F32, F64):
        trials = 100 if dtype == F32 else 5
        for _ in range(trials):
            x = dev.to_device(np.asfortranarray(rng.random(ps.img_shape).astype(dtype)), "cuda:0")
            y = dev.to_device(np.asfortranarray(rng.random(ps.sinogram_shape).astype(dtype)), "cuda:0")
            Rx = gratopy.forwardprojection(x, ps)
            Ry = gratopy.backprojection(y, ps)
            lhs = float((Rx.double() * y.double() * w[None, :]).sum()) * ps.delta_s
            rhs = float((x.double() * Ry.double()).sum()) * ps.delta_x**2
            if not ps.is_fan and dtype == F64:
                dxf, dsf = float(np.float32(ps.delta_x)), float(np.float32(ps.delta_s))
                lhs *= (ps.delta_x**2 / ps.delta_s) / (dxf * dxf / dsf)
            gap = abs(lhs - rhs) / min(abs(lhs), abs(rhs))
            if dtype == F32:
                errors += gap > 1e-5
            else:
                assert gap < 1e-12
    assert errors < 10


# ---- the reference's control numbers through the drop-in API -------------------


def _run(ps, img_np, order="F"):
    img = dev.to_device(img_np, "cuda:0", order=order)
    sino = gratopy.forwardprojection(img, ps)
    back = gratopy.backprojection(sino, ps)
    return dev.to_host(sino), dev.to_host(back)


def test_control_numbers_radon():
    print("control radon", flush=True)
    N = 1200
    img = create_phantoms(N, F32)
    ps = _ps(gratopy.RADON, img.shape, 360, 600, image_width=4.0, detector_width=4.0)
    sino, back = _run(ps, img)
    dims = (N, N, 600, 360, 2)
    assert_control(sino, dims, 4737.367, "sino", "sinogram")
    assert_control(back, dims, 7427.70, "img", "backprojection")
    for dtype in (F32, F64):
        p = phantom(600, dtype)
        ps = _ps(gratopy.RADON, (600, 600), 180, 300)
        for o1 in "FC":
            for o2 in "FC":
                sino_d = dev.zeros((300, 180), dtype, order=o2, device="cuda:0")
                img_d = dev.to_device(np.require(p, dtype, o1), "cuda:0", order=o1)
                gratopy.forwardprojection(img_d, ps, sino=sino_d)
                back_d = dev.zeros((600, 600), dtype, order=o1, device="cuda:0")
                gratopy.backprojection(sino_d, ps, img=back_d)
                dims = (600, 600, 300, 180, 1)
                assert_control(dev.to_host(sino_d), dims, 6.69419, "sino", f"sinogram {dtype}{o1}{o2}")
                assert_control(dev.to_host(back_d), dims, 20.2049, "img", f"back {dtype}{o1}{o2}")


def test_control_numbers_fan():
    print("control fan", flush=True)
    N = 1200
    img = create_phantoms(N, F32)
    ps = _ps(gratopy.FANBEAM, img.shape, 360, 600, detector_width=400.0, RE=200.0, R=752.0)
    sino, back = _run(ps, img)
    dims = (N, N, 600, 360, 2)
    assert_control(sino, dims, 94031.43, "sino", "sinogram")
    assert_control(back, dims, 1482240.72, "img", "backprojection")
    for dtype in (F32, F64):
        p = phantom(600, dtype)
        ps = _ps(gratopy.FANBEAM, (600, 600), 180, 300, detector_width=400.0, RE=200.0, R=752.0)
        for o1 in "FC":
            for o2 in "FC":
                sino_d = dev.zeros((300, 180), dtype, order=o2, device="cuda:0")
                img_d = dev.to_device(np.require(p, dtype, o1), "cuda:0", order=o1)
                gratopy.forwardprojection(img_d, ps, sino=sino_d)
                back_d = dev.zeros((600, 600), dtype, order=o1, device="cuda:0")
                gratopy.backprojection(sino_d, ps, img=back_d)
                dims = (600, 600, 300, 180, 1)
                assert_control(dev.to_host(sino_d), dims, 416.107, "sino", f"sinogram {dtype}{o1}{o2}")
                assert_control(dev.to_host(back_d), dims, 7851.41, "img", f"back {dtype}{o1}{o2}")


def test_control_numbers_limited_shift_nonquadratic():
    print("control limited/shift/nonquadratic", flush=True)
    N = 1200
    img = create_phantoms(N, F32)
    Ns, na = int(0.3 * N), 180
    delta = np.pi * 3 / 4 / (na - 1)
    incorrect = np.linspace(0, np.pi * 3 / 4.0, na) + np.pi / 8
    correct = (incorrect, np.pi / 8 - delta / 2, np.pi * 7 / 8 + delta / 2)
    ps_c = _ps(gratopy.RADON, img.shape, correct, Ns, detector_width=2.0)
    sino, back = _run(ps_c, img)
    dims = (N, N, Ns, 3, 2)
    assert_control(sino, dims, 340.93, "sino", "limited sinogram")
    assert_control(back, dims, 3385.137, "img", "limited backprojection")

    ps = _ps(gratopy.FANBEAM, img.shape, 360, 600, image_width=2.0, R=5.0, RE=3.0,
             detector_width=6.0, midpoint_shift=[0, 0.5])
    sino, back = _run(ps, img)
    dims = (N, N, 600, 360, 2)
    assert_control(sino, dims, 297.880, "sino", "shifted fan sinogram")
    assert_control(back, dims, 12217.92, "img", "shifted fan backprojection")

    N2 = int(N * 2 / 3.0)
    img2 = np.asfortranarray(img[:, 0:N2, :])
    ps = _ps(gratopy.RADON, img2.shape, 360, 600)
    sino, back = _run(ps, img2)
    dims = (N, N2, 600, 360, 2)
    assert_control(sino, dims, 2401.26, "sino", "non-square sinogram")
    assert_control(back, dims, 3310.3464, "img", "non-square backprojection")


def test_single_line_sparse_matrix_gpu():
    print("sparse matrix", flush=True)
    Nx, nd, angles = 40, 30, 12
    for geometry, kw in ((gratopy.RADON, {}), (gratopy.FANBEAM, dict(detector_width=400.0, R=752.0, RE=200.0))):
        ps = _ps(geometry, (Nx, Nx), angles, nd, **kw)
        for dtype in (F32, F64):
            M = ps.create_sparse_matrix(dtype=dtype, order="F").tocsr()
            img = phantom(Nx, F64)
            sino = (M @ img.reshape(-1, order="F")).reshape(nd, angles, order="F")
            assert rel_err(sino, _oracle(ps, img, dtype, False)) < TOL[dtype] * 10
