"""Drop-in counterpart of gratopy's ``gratopy/gratopy.py`` on MI355X.

Same public names, arguments, return values and error types as the
reference: ``ProjectionSettings``, ``forwardprojection``, ``backprojection``,
``radon``/``radon_ad``/``fanbeam``/``fanbeam_ad``, ``radon_struct``,
``fanbeam_struct``, ``read_angles``, ``weight_sinogram``, ``normest``,
``landweber``, ``conjugate_gradients`` and ``total_variation``.

Differences in kind, not in results:

* device arrays are torch tensors on a HIP device (C or F contiguous, as
  pyopencl arrays are), and ``queue`` is a device/stream (see
  :class:`gratopy_amd._device.Queue`);
* every transform and update is a hand-written gfx950 kernel reached through
  the C ABI of ``libgratopy_hip.so`` (include/gratopy_hip.h); nothing falls
  back to the CPU;
* ``wait_for`` accepts ``torch.cuda.Event`` objects (work is stream-ordered);
  functions that return an event in the reference return a
  ``torch.cuda.Event`` recorded after the launch.
"""

from __future__ import annotations

import contextlib

import numbers
import sys
from typing import Literal, TypeAlias

import numpy as np

from . import _device as dev
from . import _hip
from ._version import __version__
from .utilities import GeometryType

AngularRangeSection: TypeAlias = tuple

VERSION = __version__

PARALLEL = RADON = GeometryType.RADON
FANBEAM = FAN = GeometryType.FANBEAM


# ---------------------------------------------------------------------------
# launch plumbing
# ---------------------------------------------------------------------------
def _stream_of(projectionsetting):
    return projectionsetting.queue.torch_stream()


def _wait(stream, wait_for):
    torch = dev.torch
    for ev in wait_for or ():
        if isinstance(ev, torch.cuda.Event):
            stream.wait_event(ev)


def _record(stream):
    if dev.torch.cuda.is_current_stream_capturing():
        return None  # no event nodes inside a captured graph
    ev = dev.torch.cuda.Event()
    ev.record(stream)
    return ev


def _dims(arr):
    return int(arr.shape[2]) if arr.dim() > 2 else 1


def _require_device_array(a, name):
    if not dev.is_device_array(a):
        raise TypeError(f"{name} must be a device array (torch.Tensor on a HIP device), got {type(a)}")
    if a.device.type != "cuda":
        raise TypeError(f"{name} must live on a HIP device, got {a.device}")


def check_compatibility(img, sino, projectionsetting):
    """Assert that img, sino and projectionsetting agree in dtype and shape
    (gratopy.py:85-123; failures raise AssertionError as there)."""
    assert sino.dtype == img.dtype, (
        f"sinogram and image do not share common data type: {sino.dtype} and {img.dtype}"
    )
    assert tuple(sino.shape[0:2]) == tuple(projectionsetting.sinogram_shape), (
        f"The dimensions of the sinogram {tuple(sino.shape)} "
        f"do not match the projectionsetting's {projectionsetting.sinogram_shape}"
    )
    assert tuple(img.shape[0:2]) == tuple(projectionsetting.img_shape), (
        f"The dimensions of the image {tuple(img.shape)} do not match the "
        f"projectionsetting's {projectionsetting.img_shape}"
    )
    if len(sino.shape) > 2 and sino.shape[2] > 1:
        assert len(img.shape) > 2, "The sinogram has a third dimension but the image does not."
        assert sino.shape[2] == img.shape[2], (
            f"The third dimension (z-direction) of the sinogram is {sino.shape[2]} and the "
            f"image's is {img.shape[2]}, they do not coincide."
        )
    if len(img.shape) > 2 and img.shape[2] > 1:
        assert len(sino.shape) > 2, " The sinogram has a third dimension but the image does not."


def launch_projection(is_fan, adjoint, out, inp, tables, img_shape, sinogram_shape, stream,
                      stream_is_current=False):
    """Enqueue one projection on ``stream``: out = R inp (adjoint=False) or
    out = R^# inp.  ``tables`` maps 'ofs', 'geo' (and 'sdpd' for fanbeam) to
    device arrays of the element type of ``out``.  Shared by the function API
    and the operator API.  ``stream_is_current``: ``stream`` is torch's
    current stream of ``out``'s device, so the workspace is allocated there
    without switching streams (a few microseconds per call on small problems)."""
    torch = dev.torch
    Nx, Ny = img_shape
    Ns, Na = sinogram_shape
    Nz = _dims(out)
    in_order = dev.order_of(inp)
    out_order = dev.order_of(out)
    dcode = _hip.dtype_code(out.dtype)
    geom = _hip.GEOM_FAN if is_fan else _hip.GEOM_RADON
    lib = _hip.load_library()
    nbytes = lib.gt_projection_workspace_bytes(geom, int(adjoint), dcode, Nx, Ny, Ns, Na, Nz,
                                               dev.order_code(in_order))
    with contextlib.nullcontext() if stream_is_current else torch.cuda.stream(stream):
        ws = dev.workspace(nbytes, out.device)
        head = (stream.cuda_stream, dcode, out.data_ptr(), inp.data_ptr(), tables["ofs"].data_ptr())
        tail = (Nx, Ny, Ns, Na, Nz, dev.order_code(out_order), dev.order_code(in_order),
                ws.data_ptr(), nbytes)
        if is_fan:
            name = "gt_fan_back" if adjoint else "gt_fan_fwd"
            _hip.call(name, *head, tables["sdpd"].data_ptr(), tables["geo"].data_ptr(), *tail)
        else:
            name = "gt_radon_back" if adjoint else "gt_radon_fwd"
            _hip.call(name, *head, tables["geo"].data_ptr(), *tail)


def _project(out, inp, projectionsetting, adjoint, wait_for):
    """Launch one projection: out = R inp (adjoint=False) or R^# inp."""
    ps = projectionsetting
    if adjoint:
        img, sino = out, inp
    else:
        sino, img = out, inp
    _require_device_array(img, "img")
    _require_device_array(sino, "sino")
    check_compatibility(img, sino, ps)
    dtype = dev.to_np_dtype(sino.dtype)
    ps.ensure_dtype(dtype)
    stream = _stream_of(ps)
    _wait(stream, wait_for)
    tables = {"ofs": ps.ofs_buf[dtype], "geo": ps.geometry_information[dtype]}
    if ps.is_fan:
        tables["sdpd"] = ps.sdpd_buf[dtype]
    launch_projection(ps.is_fan, adjoint, out, inp, tables, ps.img_shape, ps.sinogram_shape, stream,
                      stream_is_current=ps.queue.stream is None and out.device == ps.queue.device)
    return _record(stream)


def forwardprojection(img, projectionsetting, sino=None, wait_for=[]):
    """Forward projection of ``img`` (gratopy.py:126-187).

    Writes into ``sino`` when given, otherwise allocates a sinogram of the
    image's order (and third dimension) on the image's device.
    """
    if sino is None:
        _require_device_array(img, "img")
        z = (int(img.shape[2]),) if img.dim() > 2 else ()
        sino = dev.empty(projectionsetting.sinogram_shape + z, img.dtype,
                         order="C" if dev.is_c_contiguous(img) else "F", device=img.device)
    projectionsetting.forwardprojection(sino, img, projectionsetting, wait_for=wait_for)
    return sino


def backprojection(sino, projectionsetting, img=None, wait_for=[]):
    """Backprojection of ``sino`` (gratopy.py:190-255)."""
    if img is None:
        _require_device_array(sino, "sino")
        z = (int(sino.shape[2]),) if sino.dim() > 2 else ()
        img = dev.empty(projectionsetting.img_shape + z, sino.dtype,
                        order="C" if dev.is_c_contiguous(sino) else "F", device=sino.device)
    projectionsetting.backprojection(img, sino, projectionsetting, wait_for=wait_for)
    return img


def radon(sino, img, projectionsetting, wait_for=[]):
    """Parallel-beam forward projection into ``sino`` (gratopy.py:258-316)."""
    return _project(sino, img, projectionsetting, False, wait_for)


def radon_ad(img, sino, projectionsetting, wait_for=[]):
    """Parallel-beam backprojection into ``img`` (gratopy.py:319-377)."""
    return _project(img, sino, projectionsetting, True, wait_for)


def fanbeam(sino, img, projectionsetting, wait_for=[]):
    """Fanbeam forward projection into ``sino`` (gratopy.py:581-640)."""
    return _project(sino, img, projectionsetting, False, wait_for)


def fanbeam_ad(img, sino, projectionsetting, wait_for=[]):
    """Fanbeam backprojection into ``img`` (gratopy.py:643-703)."""
    return _project(img, sino, projectionsetting, True, wait_for)


# ---------------------------------------------------------------------------
# geometry tables (host side, numpy float64 then cast per dtype)
# ---------------------------------------------------------------------------
_TABLE_DTYPES = (np.dtype("float64"), np.dtype("float32"))


def radon_struct(queue, img_shape, angles, angle_weights, n_detectors=None, detector_width=2.0,
                 image_width=2.0, midpoint_shift=(0.0, 0.0), detector_shift=0.0):
    """Tables of the parallel-beam transform (gratopy.py:380-578).

    ``ofs_dict[dtype]`` is an (8, Na) Fortran array with rows
    [X, Y, offset, 1/X (or 1/Y when flipped), angle weight, flipped, 0, 0];
    ``geo_dict[dtype]`` is [delta_x, delta_s, Nx, Ny, Ns, Na].  The
    expressions are evaluated in the same order as the reference so that the
    float64 tables agree bit for bit.
    """
    Nx, Ny = img_shape
    nd = int(np.ceil(np.hypot(Nx, Ny))) if n_detectors is None else n_detectors
    rho = detector_width / float(image_width) * max(img_shape) / nd  # delta_s / delta_x
    delta_x = image_width / float(max(img_shape))
    delta_s = float(detector_width) / nd
    centre = np.array([Nx - 1, Ny - 1]) / 2.0 - np.array(midpoint_shift) / delta_x
    centre_det = (nd - 1.0) / 2.0

    # unit normal of the rays, scaled to detector pixels per image pixel
    X = np.cos(angles - np.pi * 0.5) / rho
    Y = np.sin(angles - np.pi * 0.5) / rho
    steep = np.where(abs(X) < abs(Y))
    with np.errstate(divide="ignore"):
        inv = 1.0 / X
    inv[steep] = 1.0 / Y[steep]
    flipped = np.zeros(inv.shape)
    flipped[steep] = 1.0
    offset = centre_det - X * centre[0] - Y * centre[1] - detector_shift / delta_s

    n_angles = len(angles)
    ofs_dict, geo_dict, weight_dict = {}, {}, {}
    for dtype in _TABLE_DTYPES:
        ofs = np.zeros((8, n_angles), dtype=dtype, order="F")
        for row, values in enumerate((X, Y, offset, inv, angle_weights, flipped)):
            ofs[row, :] = values
        ofs_dict[dtype] = ofs
        weight_dict[dtype] = np.array(angle_weights, dtype=dtype)
        geo_dict[dtype] = np.array([delta_x, delta_s, Nx, Ny, nd, n_angles], dtype=dtype, order="F")
    return {
        "ofs_dict": ofs_dict,
        "img_shape": img_shape,
        "sinogram_shape": (nd, n_angles),
        "geo_dict": geo_dict,
        "angle_diff_dict": weight_dict,
    }


def fanbeam_struct(queue, img_shape, angles, detector_width, source_detector_dist,
                   source_origin_dist, angle_weights, n_detectors=None, detector_shift=0.0,
                   image_width=None, midpoint_shift=(0.0, 0.0), reverse_detector=False):
    """Tables of the fanbeam transform (gratopy.py:706-1003).

    ``ofs_dict[dtype]``: (8, Na) rows [xd, yd, qx, qy, dx0, dy0, angle weight, 0]
    (detector step, origin->source, origin->detector foot point) in pixel
    units; ``sdpd_dict[dtype]``: (1, Ns) source-to-detector-pixel distances;
    ``geo_dict[dtype]``: [R, RE, delta_xi, mid_x, mid_y, mid_det, Nx, Ny, Ns,
    Na, delta_x].  Same expression order as the reference.
    """
    detector_width = float(detector_width)
    R = float(source_detector_dist)
    RE = float(source_origin_dist)
    shift = np.array(midpoint_shift)
    pixels = max(img_shape[0], img_shape[1])

    nd = int(np.ceil(np.hypot(img_shape[0], img_shape[1]))) if n_detectors is None else n_detectors
    assert isinstance(nd, numbers.Integral), "Number of detectors must be integer"
    nd = int(nd)

    centre_det = (nd - 1.0) / 2.0
    centre_det = centre_det - detector_shift * nd / detector_width

    assert R > RE, "The origin is not between detector and source"

    if image_width is None:
        # largest square image that stays inside the fan from every angle
        slope = (0.5 * detector_width - abs(detector_shift)) / R
        image_width = 2 * slope * RE / np.sqrt(1 + slope**2)
        assert image_width > 0, (
            "The automatism for choosing the image_width failed as the image can never be fully "
            "contained in the fans from all directions (most likely due to detector_shift being "
            "larger than half the detector_width)! Please set the image_width parameter by hand!"
        )

    assert (
        image_width * 0.5 * np.sqrt(1 + (min(img_shape) / max(img_shape)) ** 2) + np.linalg.norm(shift)
        < RE
    ), "The source is not outside the image domain"

    mid_x = (img_shape[0] - 1) * 0.5 - (shift[0] * pixels / float(image_width))
    mid_y = (img_shape[1] - 1) * 0.5 - (shift[1] * pixels / float(image_width))

    # lengths in units of image pixels
    R *= pixels / float(image_width)
    RE *= pixels / float(image_width)
    detector_width *= pixels / float(image_width)

    cos_t = np.cos(angles)
    sin_t = np.sin(angles)
    orient = -1 if reverse_detector else 1
    XD = sin_t * detector_width / nd * orient
    YD = -cos_t * detector_width / nd * orient
    Qx = -cos_t * RE
    Qy = -sin_t * RE
    Dx0 = cos_t * (R - RE)
    Dy0 = sin_t * (R - RE)

    sinogram_shape = (nd, len(angles))
    ofs_dict, sdpd_dict, geo_dict, weight_dict = {}, {}, {}, {}
    for dtype in _TABLE_DTYPES:
        ofs = np.zeros((8, len(angles)), dtype=dtype, order="F")
        for row, values in enumerate((XD, YD, Qx, Qy, Dx0, Dy0, angle_weights)):
            ofs[row, :] = values
        ofs_dict[dtype] = ofs

        xi = (np.arange(0, nd) - centre_det) * detector_width / nd
        dist = np.array(np.sqrt(xi**2 + R**2), dtype=dtype, order="F")
        sdpd = np.zeros((1, len(dist)), dtype=dtype, order="F")
        sdpd[0, :] = dist[:]
        sdpd_dict[dtype] = sdpd

        geo_dict[dtype] = np.array(
            [R, RE, detector_width / nd, mid_x, mid_y, centre_det, img_shape[0], img_shape[1],
             sinogram_shape[0], sinogram_shape[1], image_width / float(max(img_shape))],
            dtype=dtype, order="F",
        )
        weight_dict[dtype] = np.array(angle_weights, dtype=dtype)

    return {
        "img_shape": img_shape,
        "sinogram_shape": sinogram_shape,
        "ofs_dict": ofs_dict,
        "sdpd_dict": sdpd_dict,
        "image_width": image_width,
        "geo_dict": geo_dict,
        "angle_diff_dict": weight_dict,
    }


def _merge_duplicates(sorted_angles, weights):
    """Angles equal to within 1e-6 share the mean of their weights
    (gratopy.py:1176-1189)."""
    n = len(sorted_angles)
    i = 0
    while i < n - 1:
        j = i + 1
        total = weights[i]
        while abs(sorted_angles[i] - sorted_angles[j]) < 1e-6:
            total += weights[j]
            j += 1
            if j > n - 1:
                break
        val = total / (j - i)
        for k in range(i, j):
            weights[k] = val
        i = j
    return weights


def read_angles(angles, angle_weights, projectionsetting):
    """Interpret the ``angles`` argument and derive angle weights
    (gratopy.py:1109-1296): an integer count (negative: reversed), an explicit
    list/array (full-angle natural weights), or limited-angle sections
    ``(count or angles, lower, upper)``."""
    ps = projectionsetting
    period = 2 * np.pi if ps.is_fan else np.pi
    weights_auto = None
    if np.isscalar(angles):
        na = abs(angles)
        angles_arr = np.linspace(0, period, abs(angles) + 1)[:-1]
        weights_auto = np.ones(len(angles_arr)) * (period / len(angles_arr))
        if angles < 0:
            angles_arr = np.flip(angles_arr)
            weights_auto = np.flip(weights_auto)
        angles = angles_arr
    elif isinstance(angles, (list, np.ndarray)) and np.isscalar(angles[0]):
        na = len(angles)
        angles = np.array(angles)
        order = np.argsort(angles % period)
        srt = angles[order] % period
        ext = np.array(np.hstack([-period + srt[-1], srt, srt[0] + period]))
        w = 0.5 * (abs(ext[2 : na + 2] - ext[0:na]))
        na = len(srt)
        w = _merge_duplicates(srt, w)
        weights_auto = np.zeros(w.shape)
        weights_auto[order] = w
    elif isinstance(angles, tuple) or isinstance(angles[0], tuple):
        sections = [angles] if isinstance(angles, tuple) else angles
        for sec in sections:
            assert isinstance(sec, tuple), (
                "When giving angles via tuples for limited angle setting all subsets must be "
                "given in tuple form!"
            )
            assert len(sec) == 3, (
                "When tuples are given for the limited angle setting, also angular bounds need to "
                "be given, i.e., tuple consists of number of angles or angles themselves as first "
                "entry and lower bound of  angular range as second, upper bound as third entry!"
            )
        all_angles, all_weights = [], []
        for first, lower, upper in sections:
            if isinstance(first, numbers.Integral):
                count = abs(first)
                half = (upper - lower) / count * 0.5
                current = np.linspace(lower + half, upper - half, count)
                if first < 0:
                    current = np.flip(current)
            elif isinstance(first, (list, np.ndarray)):
                current = np.array(first)
            else:
                raise TypeError(f"unsupported angular section {first!r}")
            all_angles += list(current)
            if angle_weights is None:
                srt = np.sort(current)
                order = np.argsort(current)
                ext = np.array(np.hstack([2 * lower - srt[0], srt, 2 * upper - srt[-1]]))
                w = 0.5 * (abs(ext[2 : len(ext)] - ext[0 : len(ext) - 2]))
                w = _merge_duplicates(srt, w)
                placed = np.zeros(w.shape)
                placed[order] = w
                all_weights += list(placed)
        angles = np.array(all_angles)
        na = len(angles)
        weights_auto = all_weights
    else:
        raise TypeError(f"unsupported angles argument {angles!r}")

    if angle_weights is None:
        angle_weights = np.array(weights_auto)
    elif np.isscalar(angle_weights):
        angle_weights = np.ones(na) * angle_weights
    elif isinstance(angle_weights, (list, np.ndarray)):
        assert na == len(angle_weights), (
            "The angle_weights given by the user do not have the same length as the number of "
            "angles considered"
        )
        angle_weights = np.array(angle_weights)
    return angles, angle_weights


class ProjectionSettings:
    """Geometry of a parallel-beam or fanbeam projection and its device
    tables (gratopy.py:1299-1756).  Same constructor and attributes as the
    reference; ``queue`` selects the HIP device and stream."""

    def __init__(self, queue, geometry, img_shape, angles, n_detectors=None, angle_weights=None,
                 detector_width=2.0, image_width=None, R=None, RE=None, detector_shift=0.0,
                 midpoint_shift=(0.0, 0.0), reverse_detector=False):
        self.geometry = geometry
        self.queue = dev.Queue(queue)
        self.cl_context = self.queue.device

        if np.isscalar(img_shape):
            img_shape = (img_shape, img_shape)
        if len(img_shape) > 2:
            img_shape = img_shape[0:2]
        self.img_shape = tuple(int(n) for n in img_shape)

        angles, weights = read_angles(angles, angle_weights, self)
        self.n_angles = len(angles)
        self.angle_weights = weights
        self.angles = angles

        if n_detectors is None:
            self.n_detectors = int(np.ceil(np.hypot(self.img_shape[0], self.img_shape[1])))
        else:
            self.n_detectors = n_detectors
        self.sinogram_shape = (self.n_detectors, self.n_angles)

        self.image_width = image_width
        self.detector_shift = detector_shift
        self.midpoint_shift = midpoint_shift
        self.detector_width = detector_width
        self.R = R
        self.RE = RE
        self.reverse_detector = reverse_detector
        self.buf_upload = {}

        if self.reverse_detector and self.is_parallel:
            print(
                "WARNING, the reverse_detector argument has no impact on the parallel beam setting."
                " To reverse the angles, the angles parameter can be translated by np.pi"
            )

        if self.is_fan:
            assert not (R is None or RE is None), (
                "For the Fanbeam geometry you need to set R (the normal distance from source to "
                "detector) and RE (distance from source to coordinate origin which is the rotation "
                "center)"
            )
            self.struct = fanbeam_struct(
                queue=self.queue, img_shape=self.img_shape, angles=self.angles,
                angle_weights=self.angle_weights, detector_width=self.detector_width,
                source_detector_dist=self.R, source_origin_dist=self.RE,
                n_detectors=self.n_detectors, detector_shift=self.detector_shift,
                image_width=image_width, midpoint_shift=self.midpoint_shift,
                reverse_detector=self.reverse_detector,
            )
            self.ofs_buf = dict(self.struct["ofs_dict"])
            self.sdpd_buf = dict(self.struct["sdpd_dict"])
            self.image_width = self.struct["image_width"]
            self.geometry_information = dict(self.struct["geo_dict"])
            self.angle_weights_buf = dict(self.struct["angle_diff_dict"])
            self.angle_weights = self.angle_weights_buf[np.dtype("float")].copy()
            self.delta_x = float(self.image_width) / max(self.img_shape)
            self.delta_s = float(detector_width) / self.n_detectors
            self.delta_ratio = self.delta_s / self.delta_x

        if self.is_parallel:
            self.image_width = 2.0 if image_width is None else float(self.image_width)
            self.struct = radon_struct(
                queue=self.queue, img_shape=self.img_shape, angles=self.angles,
                angle_weights=self.angle_weights, n_detectors=self.n_detectors,
                detector_width=self.detector_width, image_width=self.image_width,
                midpoint_shift=self.midpoint_shift, detector_shift=self.detector_shift,
            )
            self.ofs_buf = dict(self.struct["ofs_dict"])
            self.delta_x = self.image_width / max(self.img_shape)
            self.delta_s = self.detector_width / self.n_detectors
            self.delta_ratio = self.delta_s / self.delta_x
            self.geometry_information = dict(self.struct["geo_dict"])
            self.angle_weights_buf = dict(self.struct["angle_diff_dict"])
            self.angle_weights = self.angle_weights_buf[np.dtype("float")].copy()

    @property
    def is_parallel(self) -> bool:
        return self.geometry == GeometryType.RADON

    @property
    def is_fan(self) -> bool:
        return self.geometry == GeometryType.FANBEAM

    @property
    def geometry(self) -> GeometryType:
        return self._geometry

    @geometry.setter
    def geometry(self, value) -> None:
        self._geometry = GeometryType(value)

    def forwardprojection(self, sino, img, projectionsetting, wait_for=[]):
        if self.geometry == GeometryType.RADON:
            radon(sino, img, projectionsetting, wait_for=wait_for)
        elif self.geometry == GeometryType.FANBEAM:
            fanbeam(sino, img, projectionsetting, wait_for=wait_for)

    def backprojection(self, img, sino, projectionsetting, wait_for=[]):
        if self.geometry == GeometryType.RADON:
            radon_ad(img, sino, projectionsetting, wait_for=wait_for)
        elif self.geometry == GeometryType.FANBEAM:
            fanbeam_ad(img, sino, projectionsetting, wait_for=wait_for)

    def ensure_dtype(self, dtype):
        """Upload the tables of this dtype to the device once
        (gratopy.py:1687-1695, 1063-1106)."""
        dtype = np.dtype(dtype)
        if dtype in self.buf_upload:
            return
        device = self.queue.device
        self.ofs_buf[dtype] = dev.torch.from_numpy(np.ravel(self.ofs_buf[dtype], order="F").copy()).to(device)
        self.angle_weights_buf[dtype] = dev.torch.from_numpy(
            np.ravel(self.angle_weights_buf[dtype]).copy()).to(device)
        self.geometry_information[dtype] = dev.torch.from_numpy(
            np.ravel(self.geometry_information[dtype]).copy()).to(device)
        if self.is_fan:
            self.sdpd_buf[dtype] = dev.torch.from_numpy(np.ravel(self.sdpd_buf[dtype], order="F").copy()).to(device)
        self.buf_upload[dtype] = 1

    def set_angle_weights(self, angle_weights):
        """Replace the angle weights (gratopy.py:1697-1756)."""
        if self.is_parallel:
            self.struct = radon_struct(
                self.queue, self.img_shape, self.angles, angle_weights=angle_weights,
                n_detectors=self.n_detectors, detector_width=self.detector_width,
                image_width=self.image_width, midpoint_shift=self.midpoint_shift,
                detector_shift=self.detector_shift,
            )
        if self.is_fan:
            self.struct = fanbeam_struct(
                self.queue, self.img_shape, self.angles, angle_weights=angle_weights,
                detector_width=self.detector_width, source_detector_dist=self.R,
                source_origin_dist=self.RE, n_detectors=self.n_detectors,
                detector_shift=self.detector_shift, image_width=self.image_width,
                midpoint_shift=self.midpoint_shift, reverse_detector=self.reverse_detector,
            )
        self.ofs_buf = dict(self.struct["ofs_dict"])
        self.angle_weights_buf = dict(self.struct["angle_diff_dict"])
        self.angle_weights = self.angle_weights_buf[np.dtype("float")].copy()
        uploaded = list(self.buf_upload)
        for dtype in uploaded:
            device = self.queue.device
            self.ofs_buf[dtype] = dev.torch.from_numpy(np.ravel(self.ofs_buf[dtype], order="F").copy()).to(device)
            self.angle_weights_buf[dtype] = dev.torch.from_numpy(
                np.ravel(self.angle_weights_buf[dtype]).copy()).to(device)

    def create_sparse_matrix(self, dtype=np.dtype("float32"), order: Literal["C", "F"] = "F"):
        """System matrix of the forward projection as scipy COO
        (gratopy.py:2014-2155): column pos(x, y) holds the sinogram of a unit
        impulse at pixel (x, y), rows indexed by pos(s, a) in ``order``."""
        import scipy.sparse

        torch = dev.torch
        dtype = np.dtype(dtype)
        if order not in ("F", "C"):
            raise ValueError("Order (contiguity) not recognized, suitable choices are 'F' or 'C'")
        self.ensure_dtype(dtype)
        Nx, Ny = self.img_shape
        Ns, Na = self.sinogram_shape
        stream = _stream_of(self)
        geom = _hip.GEOM_FAN if self.is_fan else _hip.GEOM_RADON
        ofs = self.ofs_buf[dtype]
        geo = self.geometry_information[dtype]
        sdpd_ptr = self.sdpd_buf[dtype].data_ptr() if self.is_fan else None
        dcode = _hip.dtype_code(dtype)
        chunk = max(1, min(Nx * Ny, (64 << 20) // (Ns * Na * dtype.itemsize)))
        rows, cols, vals = [], [], []
        pixels = [(x, y) for x in range(Nx) for y in range(Ny)]
        with torch.cuda.stream(stream):
            buf = torch.zeros((chunk, Na, Ns), dtype=dev.to_torch_dtype(dtype), device=self.queue.device)
            for start in range(0, len(pixels), chunk):
                part = pixels[start : start + chunk]
                for k, (x, y) in enumerate(part):
                    _hip.call("gt_single_line", stream.cuda_stream, geom, dcode,
                              buf[k].data_ptr(), x, y, ofs.data_ptr(), sdpd_ptr, geo.data_ptr(),
                              Nx, Ny, Ns, Na, _hip.ORDER_F)
                sub = buf[: len(part)]
                k_idx, a_idx, s_idx = torch.nonzero(sub > 0, as_tuple=True)
                v = sub[k_idx, a_idx, s_idx].cpu().numpy()
                k_idx, a_idx, s_idx = (t.cpu().numpy() for t in (k_idx, a_idx, s_idx))
                px = np.array([p[0] for p in part])[k_idx]
                py = np.array([p[1] for p in part])[k_idx]
                if order == "F":
                    cols.append(px + Nx * py)
                    rows.append(s_idx + Ns * a_idx)
                else:
                    cols.append(px * Ny + py)
                    rows.append(s_idx * Na + a_idx)
                vals.append(v)
        print("\rSparse matrix creation complete")
        rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
        vals = np.concatenate(vals) if vals else np.zeros(0, dtype)
        return scipy.sparse.coo_matrix((vals, (rows, cols)), shape=(Ns * Na, Nx * Ny))

    def show_geometry(self, angle, figure=None, axes=None, show=True):
        """Plotting of the geometry is not part of this build; see
        :mod:`gratopy_amd` documentation."""
        raise NotImplementedError("show_geometry (matplotlib plotting) is not provided by gratopy_amd")


# ---------------------------------------------------------------------------
# streaming helpers
# ---------------------------------------------------------------------------
def weight_sinogram(sino, projectionsetting, sino_out=None, divide=False, wait_for=[]):
    """Multiply (or divide) every projection by its angle weight
    (gratopy.py:2188-2266)."""
    _require_device_array(sino, "sino")
    dtype = dev.to_np_dtype(sino.dtype)
    order = "C" if dev.is_c_contiguous(sino) else "F"
    if sino_out is None:
        sino_out = dev.empty(tuple(sino.shape), dtype, order=order, device=sino.device)
    projectionsetting.ensure_dtype(dtype)
    stream = _stream_of(projectionsetting)
    _wait(stream, wait_for)
    Ns, Na = int(sino.shape[0]), int(sino.shape[1])
    Nz = _dims(sino)
    if dev.order_of(sino_out) != dev.order_of(sino) and not (Nz == 1 and min(Ns, Na) == 1):
        raise ValueError("sino_out must have the memory order of sino")
    _hip.call("gt_weight_sino", stream.cuda_stream, _hip.dtype_code(dtype), sino_out.data_ptr(),
              sino.data_ptr(), projectionsetting.angle_weights_buf[dtype].data_ptr(), Ns, Na, Nz,
              dev.order_code(order), int(bool(divide)))
    return sino_out


def equ_mul_add(rhs, a, x, projectionsetting, wait_for=[]):
    """rhs += a*x (gratopy.py:2269-2295)."""
    stream = _stream_of(projectionsetting)
    _wait(stream, wait_for)
    _hip.call("gt_axpy", stream.cuda_stream, _hip.dtype_code(x.dtype), rhs.data_ptr(), float(a),
              x.data_ptr(), x.numel())
    return rhs


def mul_add_add(rhs, a, x, y, projectionsetting, wait_for=[]):
    """rhs = a*x + y (gratopy.py:2298-2326)."""
    stream = _stream_of(projectionsetting)
    _wait(stream, wait_for)
    _hip.call("gt_axpby", stream.cuda_stream, _hip.dtype_code(x.dtype), rhs.data_ptr(), float(a),
              x.data_ptr(), y.data_ptr(), x.numel())
    return rhs


class _Reducer:
    """Device scalars and scratch for the solvers' reductions."""

    def __init__(self, device, n_slots=8):
        torch = dev.torch
        self.values = torch.zeros(n_slots, dtype=torch.float64, device=device)
        self.ws_bytes = _hip.load_library().gt_reduce_workspace_bytes(0)
        self.ws = dev.workspace(self.ws_bytes, device)

    def slot(self, k):
        return self.values.data_ptr() + 8 * k

    def dot(self, stream, k, x, y):
        _hip.call("gt_dot", stream.cuda_stream, _hip.dtype_code(x.dtype), self.slot(k),
                  x.data_ptr(), y.data_ptr(), x.numel(), self.ws.data_ptr(), self.ws_bytes)

    def sum(self, stream, k, x):
        _hip.call("gt_sum", stream.cuda_stream, _hip.dtype_code(x.dtype), self.slot(k),
                  x.data_ptr(), x.numel(), self.ws.data_ptr(), self.ws_bytes)


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
def normest(projectionsetting, number_iterations=50, dtype="float32", allocator=None):
    """Operator norm by power iteration on R^# R (gratopy.py:2329-2377).

    Starts from ``np.random.randn(*img_shape)`` and rescales by the sum of
    the iterate, like the reference; the sums stay on the device, so the loop
    synchronises once at the end instead of once per iteration.
    """
    ps = projectionsetting
    torch = dev.torch
    stream = _stream_of(ps)
    start = np.require(np.random.randn(*ps.img_shape), dtype, "F")
    with torch.cuda.stream(stream):
        img = dev.to_device(start, ps.queue.device, order="F")
        red = _Reducer(ps.queue.device)
    sino = forwardprojection(img, ps)
    dcode = _hip.dtype_code(img.dtype)
    for _ in range(number_iterations):
        red.sum(stream, 0, img)
        _hip.call("gt_scale_dev", stream.cuda_stream, dcode, img.data_ptr(), red.slot(0), 1,
                  img.numel())
        forwardprojection(img, ps, sino=sino)
        backprojection(sino, ps, img=img)
    stream.synchronize()
    normsqr = float(red.values[0].item())
    return np.sqrt(normsqr)


def _iterate(bodies, number_iterations, projectionsetting, graph, label):
    """Run ``bodies[i % len(bodies)](i)`` for i < number_iterations.

    Eager mode enqueues every launch from Python, like the reference.  With
    ``graph=True`` a block of whole periods (all bodies in order, up to about
    16 iterations) is captured once into a HIP graph (on the settings'
    stream, or on torch's capture side stream when the settings use the
    current stream) and replayed on the stream the eager launches would use;
    leftover iterations run eagerly.  The captured launches are the eager
    ones, so results are identical.  Returns the number of iterations run.
    """
    torch = dev.torch
    period = len(bodies)
    per_graph = period * max(1, min(number_iterations // period, 16 // period)) if graph else 0
    n_graph = (number_iterations // per_graph) if per_graph else 0
    if n_graph:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=projectionsetting.queue.stream):
            for k in range(per_graph):
                bodies[k % period](k)
        with torch.cuda.stream(_stream_of(projectionsetting)):
            for r in range(n_graph):
                g.replay()
                sys.stdout.write("\rProgress at {:3.0%}".format(float(r * per_graph) / number_iterations))
        del g
    for i in range(n_graph * per_graph, number_iterations):
        sys.stdout.write("\rProgress at {:3.0%}".format(float(i) / number_iterations))
        bodies[i % period](i)
    print(f"\r{label} complete")
    return number_iterations


def landweber(sino, projectionsetting, number_iterations=100, w=1.0, *, norm_estimate=None,
              graph=False):
    """Landweber (SIRT) iteration u <- u - w/||R||^2 R^#(R u - f), u0 = 0
    (gratopy.py:2380-2446).  Keywords not in the reference:
    ``norm_estimate`` skips :func:`normest` when the operator norm is known;
    ``graph=True`` replays the iteration as a captured HIP graph."""
    ps = projectionsetting
    _require_device_array(sino, "sino")
    order = "C" if dev.is_c_contiguous(sino) else "F"
    if norm_estimate is None:
        norm_estimate = normest(ps, allocator=None)
    dtype = dev.to_np_dtype(sino.dtype)
    w = dtype.type(w / norm_estimate**2)
    stream = _stream_of(ps)
    ps.ensure_dtype(dtype)
    with dev.torch.cuda.stream(stream):
        sinonew = sino.clone(memory_format=dev.torch.preserve_format)
        z = (int(sino.shape[2]),) if sino.dim() > 2 else ()
        U = dev.empty(ps.img_shape + z, dtype, order=order, device=sino.device)
        Unew = dev.zeros(ps.img_shape + z, dtype, order=order, device=sino.device)
    dcode = _hip.dtype_code(dtype)

    def body(_):
        s = _stream_of(ps)
        forwardprojection(Unew, ps, sino=sinonew)
        _hip.call("gt_sub", s.cuda_stream, dcode, sinonew.data_ptr(), sinonew.data_ptr(),
                  sino.data_ptr(), sino.numel())
        backprojection(sinonew, ps, img=U)
        _hip.call("gt_axpy", s.cuda_stream, dcode, Unew.data_ptr(), float(-w), U.data_ptr(),
                  U.numel())

    _iterate([body], number_iterations, ps, graph, "Landweber reconstruction")
    return Unew


def conjugate_gradients(sino, projectionsetting, number_iterations=20, epsilon=0.0, x0=None, *,
                        graph=False):
    """Conjugate gradients on the normal equations in the weighted spaces
    (gratopy.py:2449-2551).  alpha and beta are formed on the device; the
    host reads the residual only when ``epsilon > 0`` (early stopping).
    ``graph=True`` (keyword, not in the reference; ignored when
    ``epsilon > 0``) replays the iterations as a captured HIP graph."""
    ps = projectionsetting
    _require_device_array(sino, "sino")
    torch = dev.torch
    dims = ps.img_shape + ((int(sino.shape[2]),) if sino.dim() > 2 else ())
    order = "C" if dev.is_c_contiguous(sino) else "F"
    dtype = dev.to_np_dtype(sino.dtype)
    stream = _stream_of(ps)
    ps.ensure_dtype(dtype)
    with torch.cuda.stream(stream):
        if x0 is None:
            x0 = dev.zeros(dims, dtype, order=order, device=sino.device)
        assert dev.is_c_contiguous(x0) == dev.is_c_contiguous(sino), (
            "The data sino and initial guess x0 must have the same contiguity!"
        )
        x = x0.clone(memory_format=torch.preserve_format)
        red = _Reducer(sino.device)
    d = forwardprojection(x, ps)
    dcode = _hip.dtype_code(dtype)
    _hip.call("gt_sub", stream.cuda_stream, dcode, d.data_ptr(), sino.data_ptr(), d.data_ptr(),
              d.numel())
    p = backprojection(d, ps)
    with torch.cuda.stream(stream):
        q = torch.empty_like(d)
        q_rescaled = torch.empty_like(d)
    snew = backprojection(d, ps)
    with torch.cuda.stream(stream):
        sold = snew.clone(memory_format=torch.preserve_format)
    n_img, n_sino = x.numel(), d.numel()
    scale = ps.delta_x**2 / ps.delta_s
    red.dot(stream, 0, sold, sold)
    if epsilon > 0:
        red.dot(stream, 6, sino, sino)
    buffers = [sold, snew]

    def body(k):
        # buffers[k % 2] holds s_old on entry; the new gradient goes to the other one
        s = _stream_of(ps)
        s_old, s_new = buffers[k % 2], buffers[(k + 1) % 2]
        forwardprojection(p, ps, sino=q)
        weight_sinogram(q, ps, q_rescaled)
        red.dot(s, 1, q_rescaled, q)
        _hip.call("gt_cg_scalars", s.cuda_stream, red.slot(0), float(scale))
        _hip.call("gt_axpy_dev", s.cuda_stream, dcode, x.data_ptr(), red.slot(4), 1.0,
                  p.data_ptr(), n_img)
        _hip.call("gt_axpy_dev", s.cuda_stream, dcode, d.data_ptr(), red.slot(4), -1.0,
                  q.data_ptr(), n_sino)
        backprojection(d, ps, img=s_new)
        red.dot(s, 2, s_new, s_new)
        _hip.call("gt_cg_scalars", s.cuda_stream, red.slot(0), 0.0)
        _hip.call("gt_axpby_dev", s.cuda_stream, dcode, p.data_ptr(), red.slot(5), 1.0,
                  p.data_ptr(), s_new.data_ptr(), n_img)

    if epsilon > 0:
        for k in range(number_iterations):
            sys.stdout.write("\rProgress at {:3.0%}".format(float(k) / number_iterations))
            body(k)
            red.dot(stream, 3, d, d)
            stream.synchronize()
            vals = red.values.cpu().numpy()
            residue = np.sqrt(vals[3] / vals[6])
            if residue < epsilon:
                print("\rProgress aborted prematurely as desired" + "precision is reached")
                break
        print("\rCG reconstruction complete")
    else:
        _iterate([body, body], number_iterations, ps, graph, "CG reconstruction")
    return x


def total_variation(sino, projectionsetting, mu, number_iterations=1000, slice_thickness=1.0,
                    stepsize_weighting=10.0, *, norm_estimate=None, graph=False):
    """Chambolle-Pock primal-dual iteration for
    min_u mu/2 ||R u - f||^2 + TV(u), u >= 0 (gratopy.py:2554-2762).

    Per iteration: fused dual update (v += sigma grad u_bar, projected onto
    the unit ball), forward projection, lambda update, backprojection, fused
    primal + extragradient update.  A 3-D sinogram is reconstructed slice by
    slice coupled through ``slice_thickness`` (the reference raises
    TypeError on 3-D input at gratopy.py:2626-2628); ``slice_thickness=0``
    decouples the slices.  Keywords not in the reference: ``norm_estimate``
    skips :func:`normest`; ``graph=True`` replays the iterations as a
    captured HIP graph.
    """
    ps = projectionsetting
    _require_device_array(sino, "sino")
    torch = dev.torch
    dtype = dev.to_np_dtype(sino.dtype)
    order = "C" if dev.is_c_contiguous(sino) else "F"
    img_shape = ps.img_shape
    if sino.dim() == 2:
        slice_thickness = 0
    else:
        img_shape = img_shape + (int(sino.shape[2]),)
    Nx, Ny = ps.img_shape
    Nz = img_shape[2] if len(img_shape) > 2 else 1
    stream = _stream_of(ps)
    with torch.cuda.stream(stream):
        U = dev.zeros(img_shape, dtype, order=order, device=sino.device)
        U_ = dev.zeros(img_shape, dtype, order=order, device=sino.device)
        V = torch.zeros(4 * U.numel(), dtype=dev.to_torch_dtype(dtype), device=sino.device)
        Lamb = dev.zeros(tuple(sino.shape), dtype, order=order, device=sino.device)
        KU = dev.zeros(tuple(sino.shape), dtype, order=order, device=sino.device)
        KSTAR = dev.zeros(img_shape, dtype, order=order, device=sino.device)

    if norm_estimate is None:
        norm_estimate = normest(ps)
    Lsqr = 13
    sigma = 1.0 / np.sqrt(Lsqr) * stepsize_weighting
    tau = 1.0 / np.sqrt(Lsqr) / stepsize_weighting
    assert mu > 0, "Regularization parameter mu must be positive"
    mu *= norm_estimate**2 * ps.delta_x
    mu = mu / (sigma + mu)

    dcode = _hip.dtype_code(dtype)
    ocode = dev.order_code(order)
    Ns, Na = ps.sinogram_shape
    ps.ensure_dtype(dtype)
    pair = [U, U_]

    def body(i):
        # pair[i % 2] is the primal iterate u, pair[(i+1) % 2] the extrapolation u_bar
        s = _stream_of(ps)
        u, u_bar = pair[i % 2], pair[(i + 1) % 2]
        _hip.call("gt_tv_dual", s.cuda_stream, dcode, V.data_ptr(), u_bar.data_ptr(),
                  float(sigma), float(slice_thickness), Nx, Ny, Nz, ocode)
        forwardprojection(u_bar, ps, sino=KU)
        _hip.call("gt_tv_update_lambda", s.cuda_stream, dcode, Lamb.data_ptr(), KU.data_ptr(),
                  sino.data_ptr(), float(sigma / norm_estimate), float(mu), Ns, Na, Nz, ocode)
        backprojection(Lamb, ps, img=KSTAR)
        # u_bar <- max(0, u + tau(div v - K*lambda/||R||)) (the new u); u <- 2 u_new - u
        _hip.call("gt_tv_primal", s.cuda_stream, dcode, u_bar.data_ptr(), u.data_ptr(),
                  V.data_ptr(), KSTAR.data_ptr(), float(tau), float(1.0 / norm_estimate),
                  float(slice_thickness), Nx, Ny, Nz, ocode)

    n = _iterate([body, body], number_iterations, ps, graph, "TV reconstruction")
    # after iteration i the new u sits in pair[(i+1) % 2]
    return pair[n % 2]
