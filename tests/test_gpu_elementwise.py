"""The one-to-one elementwise entry points of the C ABI on the GPU, called
directly through ctypes on buffers in the reference's layout, against the
oracle's restatement of the matching reference kernels:

* gt_tv_update_v      <- update_v_*            (gratopy/total_variation.cl:54-93)
* gt_tv_update_normv  <- update_NormV_unchor_* (gratopy/total_variation.cl:150-167)
* gt_tv_update_lambda <- update_lambda_L2_*    (gratopy/total_variation.cl:95-106)
* gt_tv_update_u      <- update_u_*            (gratopy/total_variation.cl:108-148)
* gt_tv_update_extra  <- update_extra_         (gratopy/gratopy.py:2708-2717)
* gt_axpy / gt_axpby  <- equ_mul_add_* / mul_add_add_* (gratopy/utilities.cl:69-80)
* gt_weight_sino      <- multiply_* / divide_* (gratopy/utilities.cl:38-67)

and the fused gt_tv_dual / gt_tv_primal against the same kernels in sequence.
Both orders, both element types, 2-D and 3-D (with and without slice
coupling).  Bars: normwise max|d|/max|ref| <= 1e-6 (float32) and 1e-14
(float64), a few units in the last place: the kernels evaluate the same
expressions, and only contraction into FMAs, hypot and division may round
differently.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from gratopy_amd import _hip
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

BAR = {np.dtype("float32"): 1e-6, np.dtype("float64"): 1e-14}
SHAPES = [((37, 29), 0.0), ((21, 17, 5), 1.3), ((21, 17, 5), 0.0)]


def _dev(a):
    """Device copy of a numpy array that keeps its memory layout (flat data)."""
    flat = np.ravel(a, order="K")
    return torch.from_numpy(flat.copy()).to("cuda:0")


def _host(t, like):
    """Numpy array shaped and ordered like `like` from the flat device data."""
    order = "F" if like.flags.f_contiguous and not like.flags.c_contiguous else "C"
    return t.cpu().numpy().reshape(like.shape, order=order)


def _order(a):
    return 1 if a.flags.c_contiguous else 0


def _dims(shape):
    return (shape + (1,))[:3] if len(shape) == 2 else shape


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(got, ref, dtype, label):
    e = rel_err(got, ref)
    print(f"{label}: rel err {e:.2e}", flush=True)
    assert e <= BAR[np.dtype(dtype)], (label, e)


def _inputs(shape, dtype, order, seed):
    rng = np.random.default_rng(seed)
    u = np.require(rng.standard_normal(shape), dtype, order)
    # dual variable: 4 reals per pixel at the pixel's position in the image order
    v = np.require(rng.standard_normal((int(np.prod(shape)), 4)) * 0.8, dtype, "C")
    return rng, u, v


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("shape,zd", SHAPES)
def test_tv_update_v_and_normv(hip_device, shape, zd, order, dtype):
    lib = _hip.load_library()
    print(f"tv_update_v/normv {shape} zd={zd} {order} {dtype}", flush=True)
    _, u, v = _inputs(shape, dtype, order, 1)
    Nx, Ny, Nz = _dims(shape)
    ref = v.copy()
    oracle.tv_update_v(ref, u, 0.7, zd)
    v_d, u_d = _dev(v), _dev(u)
    _hip.check(lib.gt_tv_update_v(_stream(), _hip.dtype_code(dtype), v_d.data_ptr(), u_d.data_ptr(),
                                  0.7, zd, Nx, Ny, Nz, _order(u)), "gt_tv_update_v")
    torch.cuda.synchronize()
    got = v_d.cpu().numpy().reshape(v.shape)
    # the pad entry (index 3) is left alone by both
    _check(got, ref, dtype, "tv_update_v")
    oracle.tv_update_normv(ref, shape, order)
    n_d = torch.empty(0, device="cuda:0")
    _hip.check(lib.gt_tv_update_normv(_stream(), _hip.dtype_code(dtype), v_d.data_ptr(),
                                      n_d.data_ptr(), Nx, Ny, Nz, _order(u)), "gt_tv_update_normv")
    torch.cuda.synchronize()
    _check(v_d.cpu().numpy().reshape(v.shape), ref, dtype, "tv_update_normv")
    # fused dual step from the same start
    f_d = _dev(v)
    _hip.check(lib.gt_tv_dual(_stream(), _hip.dtype_code(dtype), f_d.data_ptr(), u_d.data_ptr(), 0.7,
                              zd, Nx, Ny, Nz, _order(u)), "gt_tv_dual")
    torch.cuda.synchronize()
    _check(f_d.cpu().numpy().reshape(v.shape), ref, dtype, "tv_dual")


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("shape,zd", SHAPES)
def test_tv_update_u_and_extra(hip_device, shape, zd, order, dtype):
    lib = _hip.load_library()
    print(f"tv_update_u/extra {shape} zd={zd} {order} {dtype}", flush=True)
    rng, u_old, v = _inputs(shape, dtype, order, 2)
    kstar = np.require(rng.standard_normal(shape), dtype, order)
    Nx, Ny, Nz = _dims(shape)
    tau, norming = 0.3, 0.45
    ref_u = np.zeros_like(u_old)
    oracle.tv_update_u(ref_u, u_old, v, kstar, tau, norming, zd)
    # u_new must have non-trivial positive and clamped entries
    assert (ref_u > 0).mean() > 0.2 and (ref_u == 0).mean() > 0.2
    un_d, uo_d, v_d, k_d = (_dev(np.zeros_like(u_old)), _dev(u_old), _dev(v), _dev(kstar))
    _hip.check(lib.gt_tv_update_u(_stream(), _hip.dtype_code(dtype), un_d.data_ptr(),
                                  uo_d.data_ptr(), v_d.data_ptr(), k_d.data_ptr(), tau, norming, zd,
                                  Nx, Ny, Nz, _order(u_old)), "gt_tv_update_u")
    torch.cuda.synchronize()
    _check(_host(un_d, u_old), ref_u, dtype, "tv_update_u")
    # extragradient: u_old <- 2 u_new - u_old
    ref_extra = u_old.copy(order="K")
    oracle.tv_update_extra(ref_u, ref_extra)
    _hip.check(lib.gt_tv_update_extra(_stream(), _hip.dtype_code(dtype), un_d.data_ptr(),
                                      uo_d.data_ptr(), u_old.size), "gt_tv_update_extra")
    torch.cuda.synchronize()
    _check(_host(uo_d, u_old), ref_extra, dtype, "tv_update_extra")
    _check(_host(un_d, u_old), ref_u, dtype, "tv_update_extra leaves u_ alone")
    # fused primal step from the same start
    pn_d, po_d = _dev(np.zeros_like(u_old)), _dev(u_old)
    _hip.check(lib.gt_tv_primal(_stream(), _hip.dtype_code(dtype), pn_d.data_ptr(), po_d.data_ptr(),
                                v_d.data_ptr(), k_d.data_ptr(), tau, norming, zd, Nx, Ny, Nz,
                                _order(u_old)), "gt_tv_primal")
    torch.cuda.synchronize()
    _check(_host(pn_d, u_old), ref_u, dtype, "tv_primal u_new")
    _check(_host(po_d, u_old), ref_extra, dtype, "tv_primal u_bar")


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("sino_shape", [(45, 31), (19, 13, 4)])
def test_tv_update_lambda(hip_device, sino_shape, order, dtype):
    lib = _hip.load_library()
    print(f"tv_update_lambda {sino_shape} {order} {dtype}", flush=True)
    rng = np.random.default_rng(3)
    lam, ku, f = (np.require(rng.standard_normal(sino_shape), dtype, order) for _ in range(3))
    Ns, Na, Nz = _dims(sino_shape)
    sigma, sp1 = 0.37, 0.81
    ref = lam.copy(order="K")
    oracle.tv_update_lambda(ref, ku, f, sigma, sp1)
    l_d, k_d, f_d = _dev(lam), _dev(ku), _dev(f)
    _hip.check(lib.gt_tv_update_lambda(_stream(), _hip.dtype_code(dtype), l_d.data_ptr(),
                                       k_d.data_ptr(), f_d.data_ptr(), sigma, sp1, Ns, Na, Nz,
                                       _order(lam)), "gt_tv_update_lambda")
    torch.cuda.synchronize()
    _check(_host(l_d, lam), ref, dtype, "tv_update_lambda")


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("n", [1, 1000, 300_001])
def test_axpy_axpby(hip_device, n, dtype):
    lib = _hip.load_library()
    print(f"axpy/axpby n={n} {dtype}", flush=True)
    rng = np.random.default_rng(4)
    rhs, x, y = (rng.standard_normal(n).astype(dtype) for _ in range(3))
    a = -0.731
    ref = rhs.copy(order="K")
    oracle.equ_mul_add(ref, a, x)
    r_d, x_d, y_d = _dev(rhs), _dev(x), _dev(y)
    _hip.check(lib.gt_axpy(_stream(), _hip.dtype_code(dtype), r_d.data_ptr(), a, x_d.data_ptr(), n),
               "gt_axpy")
    torch.cuda.synchronize()
    _check(r_d.cpu().numpy(), ref, dtype, "axpy")
    ref2 = rhs.copy(order="K")
    oracle.mul_add_add(ref2, a, x, y)
    r2_d = _dev(rhs)
    _hip.check(lib.gt_axpby(_stream(), _hip.dtype_code(dtype), r2_d.data_ptr(), a, x_d.data_ptr(),
                            y_d.data_ptr(), n), "gt_axpby")
    torch.cuda.synchronize()
    _check(r2_d.cpu().numpy(), ref2, dtype, "axpby")


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("divide", [0, 1])
@pytest.mark.parametrize("sino_shape", [(45, 31), (19, 13, 4)])
def test_weight_sino(hip_device, sino_shape, divide, order, dtype):
    lib = _hip.load_library()
    print(f"weight_sino {sino_shape} divide={divide} {order} {dtype}", flush=True)
    rng = np.random.default_rng(6)
    sino = np.require(rng.standard_normal(sino_shape), dtype, order)
    w = (0.5 + rng.random(sino_shape[1])).astype(dtype)
    ref = oracle.weight_sino(sino, w, divide=bool(divide))
    Ns, Na, Nz = _dims(sino_shape)
    s_d, w_d, o_d = _dev(sino), _dev(w), _dev(np.zeros_like(sino))
    _hip.check(lib.gt_weight_sino(_stream(), _hip.dtype_code(dtype), o_d.data_ptr(), s_d.data_ptr(),
                                  w_d.data_ptr(), Ns, Na, Nz, _order(sino), divide), "gt_weight_sino")
    torch.cuda.synchronize()
    _check(_host(o_d, sino), ref, dtype, "weight_sino")
    # in place (out aliases in)
    _hip.check(lib.gt_weight_sino(_stream(), _hip.dtype_code(dtype), s_d.data_ptr(), s_d.data_ptr(),
                                  w_d.data_ptr(), Ns, Na, Nz, _order(sino), divide), "gt_weight_sino")
    torch.cuda.synchronize()
    _check(_host(s_d, sino), ref, dtype, "weight_sino in place")
