"""Many 2-D slices of uneven geometry in one launch per step.

The reference reconstructs a set of slices with different geometries by a
host loop over ``forwardprojection`` / ``backprojection`` / ``landweber`` /
``total_variation`` (gratopy/gratopy.py:126-255, 2380-2446, 2554-2762), one
``ProjectionSettings`` per geometry.  Here a :class:`BatchPlan` binds every
slice (its settings, image, sinogram and solver buffers) once; each step of
a solver is then a few kernel launches over tile lists that cover all slices
(``gt_batch_launch`` of ``include/gratopy_hip.h``), and the iteration loop
can be captured in a HIP graph.  Per slice the results are bitwise those of
the single-slice functions with the same settings.

Slices are 2-D float32 or float64 device tensors in Fortran order.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _device as dev
from . import _hip
from .gratopy import _iterate, _stream_of, normest

torch = dev.torch


def _f_order(t, name):
    if t.dim() != 2 or not dev.is_f_contiguous(t):
        raise ValueError(f"{name}: batch slices must be 2-D Fortran-ordered device tensors")
    return t


class BatchPlan:
    """Device plan of a batch (gt_batch_plan_create / gt_batch_plan_destroy).

    ``settings[i]`` describes slice i; ``imgs[i]`` (Nx x Ny) and
    ``sinos[i]`` (Ns x Na) are its operands.  Optional per-slice buffers:
    ``datas`` (measured sinograms), ``imgs2`` and ``fields`` (TV state) and
    ``params`` (six scalars, see gt_batch_item).  The plan keeps references to
    every tensor it points at.
    """

    def __init__(self, settings, imgs, sinos, datas=None, imgs2=None, fields=None, params=None):
        n = len(settings)
        if not (n and len(imgs) == n and len(sinos) == n):
            raise ValueError("settings, imgs and sinos must be non-empty lists of equal length")
        for name, extra in (("datas", datas), ("imgs2", imgs2), ("fields", fields), ("params", params)):
            if extra is not None and len(extra) != n:
                raise ValueError(f"{name} must have one entry per slice")
        dtype = dev.to_np_dtype(imgs[0].dtype)
        device = imgs[0].device
        self.dtype = dtype
        self.n = n
        self.settings = list(settings)
        self._keep = [imgs, sinos, datas, imgs2, fields]
        items = (_hip.BatchItem * n)()
        for i, ps in enumerate(settings):
            img, sino = _f_order(imgs[i], "img"), _f_order(sinos[i], "sino")
            # every pointer handed to the kernels must be a tensor of this
            # batch's dtype, on its HIP device, covering what the slice's
            # geometry reads or writes
            operands = [("img", img, ps.img_shape), ("sino", sino, ps.sinogram_shape)]
            if datas is not None:
                operands.append(("data", _f_order(datas[i], "data"), ps.sinogram_shape))
            if imgs2 is not None:
                operands.append(("img2", _f_order(imgs2[i], "img2"), ps.img_shape))
            for name, t, shape in operands:
                if dev.to_np_dtype(t.dtype) != dtype:
                    raise ValueError("all slices of a batch must have one dtype")
                if tuple(t.shape) != tuple(shape):
                    raise ValueError(f"slice {i}: {name} shapes do not match its ProjectionSettings")
            if fields is not None:
                fld = fields[i]
                if dev.to_np_dtype(fld.dtype) != dtype or not fld.is_contiguous():
                    raise ValueError(f"slice {i}: field must be a contiguous tensor of the batch dtype")
                if fld.numel() < 4 * ps.img_shape[0] * ps.img_shape[1]:
                    raise ValueError(f"slice {i}: field needs 4 * Nx * Ny entries")
                operands.append(("field", fld, None))
            for name, t, _ in operands:
                if t.device.type != "cuda" or t.device != device:
                    raise ValueError(f"slice {i}: {name} must live on the batch's HIP device "
                                     f"({device}), got {t.device}")
            if ps.queue.device is not None and dev.torch.device(ps.queue.device) != device:
                raise ValueError(f"slice {i}: its ProjectionSettings live on {ps.queue.device}, "
                                 f"the batch on {device}")
            ps.ensure_dtype(dtype)
            tables = [ps.ofs_buf[dtype], ps.geometry_information[dtype]]
            if ps.is_fan:
                tables.append(ps.sdpd_buf[dtype])
            # the plan holds raw pointers to the tables: keep these tensors
            # alive even if the settings later replace them (set_angle_weights)
            self._keep.append(tables)
            it = items[i]
            it.geometry = _hip.GEOM_FAN if ps.is_fan else _hip.GEOM_RADON
            it.Nx, it.Ny = ps.img_shape
            it.Ns, it.Na = ps.sinogram_shape
            it.ofs = ps.ofs_buf[dtype].data_ptr()
            it.sdpd = ps.sdpd_buf[dtype].data_ptr() if ps.is_fan else None
            it.geo = ps.geometry_information[dtype].data_ptr()
            it.img = img.data_ptr()
            it.sino = sino.data_ptr()
            if datas is not None:
                it.data = datas[i].data_ptr()
            if imgs2 is not None:
                it.img2 = imgs2[i].data_ptr()
            if fields is not None:
                it.field = fields[i].data_ptr()
            if params is not None:
                for k, v in enumerate(params[i]):
                    it.p[k] = float(v)
        lib = _hip.load_library()
        handle = ctypes.c_void_p()
        _hip.check(lib.gt_batch_plan_create(ctypes.byref(handle), _hip.dtype_code(dtype), n,
                                            ctypes.cast(items, ctypes.c_void_p)),
                   "gt_batch_plan_create")
        self._handle = handle
        self.bytes = int(lib.gt_batch_plan_bytes(handle))

    def launch(self, op, sigma=0.0, stream=None):
        if self._handle is None:
            raise RuntimeError("batch plan already destroyed")
        s = stream if stream is not None else _stream_of(self.settings[0])
        _hip.call("gt_batch_launch", self._handle, s.cuda_stream, int(op), float(sigma))

    def destroy(self):
        if self._handle is not None:
            _stream_of(self.settings[0]).synchronize()
            _hip.call("gt_batch_plan_destroy", self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.destroy()
        except Exception:
            pass


def _device_of(tensors):
    return tensors[0].device


def batch_forwardprojection(imgs, settings):
    """[R_i img_i] for all slices, one launch set."""
    with torch.cuda.stream(_stream_of(settings[0])):
        sinos = [dev.empty(ps.sinogram_shape, dev.to_np_dtype(img.dtype), order="F", device=img.device)
                 for ps, img in zip(settings, imgs)]
    plan = BatchPlan(settings, imgs, sinos)
    plan.launch(_hip.BATCH_FORWARD)
    plan.destroy()
    return sinos


def batch_backprojection(sinos, settings):
    """[B_i sino_i] for all slices, one launch set."""
    with torch.cuda.stream(_stream_of(settings[0])):
        imgs = [dev.empty(ps.img_shape, dev.to_np_dtype(s.dtype), order="F", device=s.device)
                for ps, s in zip(settings, sinos)]
    plan = BatchPlan(settings, imgs, sinos)
    plan.launch(_hip.BATCH_BACKWARD)
    plan.destroy()
    return imgs


def _norms(settings, norm_estimates):
    if norm_estimates is not None:
        return [float(v) for v in norm_estimates]
    cache = {}
    out = []
    for ps in settings:
        if id(ps) not in cache:
            cache[id(ps)] = normest(ps)
        out.append(cache[id(ps)])
    return out


def batch_landweber(sinos, settings, number_iterations=100, w=1.0, *, norm_estimates=None,
                    graph=True):
    """Landweber for every slice (gratopy.py:2380-2446): u_i <- u_i -
    w/||R_i||^2 B_i(R_i u_i - f_i), u_i = 0 at the start.  One iteration of
    all slices is one batched forward and one batched backprojection with the
    residual and the update fused into their epilogues."""
    n = len(sinos)
    norms = _norms(settings, norm_estimates)
    dtype = dev.to_np_dtype(sinos[0].dtype)
    with torch.cuda.stream(_stream_of(settings[0])):
        us = [dev.zeros(ps.img_shape, dtype, order="F", device=sinos[0].device) for ps in settings]
        res = [dev.empty(ps.sinogram_shape, dtype, order="F", device=sinos[0].device) for ps in settings]
    params = [(-float(dtype.type(w / norms[i] ** 2)), 0, 0, 0, 0, 0) for i in range(n)]
    plan = BatchPlan(settings, us, res, datas=sinos, params=params)

    def body(_):
        plan.launch(_hip.BATCH_LANDWEBER)

    _iterate([body], number_iterations, settings[0], graph, "Batched Landweber reconstruction")
    plan.destroy()
    return us


def batch_total_variation(sinos, settings, mu, number_iterations=1000, stepsize_weighting=10.0, *,
                          norm_estimates=None, graph=True):
    """The TV primal-dual reconstruction of :func:`gratopy_amd.total_variation`
    (gratopy.py:2554-2762, slice_thickness 0) for every slice, one
    Chambolle-Pock step of all slices per five launches: dual update,
    image copy, forward projection with the lambda update in its epilogue,
    sinogram staging, backprojection with the primal and extragradient
    updates in its epilogue."""
    n = len(sinos)
    norms = _norms(settings, norm_estimates)
    dtype = dev.to_np_dtype(sinos[0].dtype)
    sigma = 1.0 / np.sqrt(13) * stepsize_weighting
    tau = 1.0 / np.sqrt(13) / stepsize_weighting
    assert mu > 0, "Regularization parameter mu must be positive"
    device = sinos[0].device
    with torch.cuda.stream(_stream_of(settings[0])):
        us = [dev.zeros(ps.img_shape, dtype, order="F", device=device) for ps in settings]
        us2 = [dev.zeros(ps.img_shape, dtype, order="F", device=device) for ps in settings]
        fields = [torch.zeros(4 * ps.img_shape[0] * ps.img_shape[1], dtype=dev.to_torch_dtype(dtype),
                              device=device) for ps in settings]
        lambdas = [dev.zeros(ps.sinogram_shape, dtype, order="F", device=device) for ps in settings]
    params = []
    for i, ps in enumerate(settings):
        m = mu * norms[i] ** 2 * ps.delta_x
        m = m / (sigma + m)
        params.append((0.0, sigma / norms[i], m, tau, 1.0 / norms[i], 0.0))
    plan = BatchPlan(settings, us, lambdas, datas=sinos, imgs2=us2, fields=fields, params=params)

    def even(_):
        plan.launch(_hip.BATCH_TV_EVEN, sigma)

    def odd(_):
        plan.launch(_hip.BATCH_TV_ODD, sigma)

    k = _iterate([even, odd], number_iterations, settings[0], graph, "Batched TV reconstruction")
    plan.destroy()
    return us if k % 2 == 0 else us2
