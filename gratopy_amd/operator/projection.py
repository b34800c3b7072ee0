"""Projection operators of the operator interface.

:class:`Radon` mirrors gratopy/operator/projection.py:58-396: built from an
image domain, angles and detectors, ``R.T`` is the adjoint, extent
placeholders are resolved at construction, and the transform runs on the
gfx950 kernels of libgratopy_hip.so with the tables of ``radon_struct``.

:class:`Fanbeam` is a working fan-beam operator with the same interface; the
reference only declares a stub for it (projection.py:399-408).
"""

from __future__ import annotations

from copy import copy
from types import SimpleNamespace
from typing import Any

import numpy as np

from gratopy_amd import _device as dev
from gratopy_amd.gratopy import fanbeam_struct, launch_projection, radon_struct
from gratopy_amd.operator.hip import HipKernelSpec, _HipOperator, builtin_kernel_spec
from gratopy_amd.utilities import (
    Angles,
    Detectors,
    ExtentPlaceholder,
    ImageDomain,
    full_detector_given_image_halfcircle,
    full_image_given_detector_halfcircle,
    valid_detector_given_image_halfcircle,
    valid_image_given_detector_halfcircle,
)


class _ProjectionOperator(_HipOperator):
    """Common part of Radon and Fanbeam: state, adjoint flip, device tables."""

    _geometry_is_fan = False

    def _init_shapes(self):
        self.projection_settings = None
        self._host_struct = None
        self._device_struct: dict[tuple[Any, np.dtype], dict[str, Any]] = {}
        image_shape = tuple(self.image_domain.size)
        sinogram_shape = (self.detectors.number, len(self.angles))
        if self.adjoint:
            self.input_shape, self.output_shape = sinogram_shape, image_shape
        else:
            self.input_shape, self.output_shape = image_shape, sinogram_shape

    def __getstate__(self) -> dict[str, Any]:
        state = super().__getstate__()
        state["projection_settings"] = None
        state["_host_struct"] = None
        state["_device_struct"] = {}
        return state

    @property
    def image_domain(self) -> ImageDomain:
        return self.state["image_domain"]

    @property
    def angles(self) -> Angles:
        return self.state["angles"]

    @property
    def detectors(self) -> Detectors:
        return self.state["detectors"]

    @property
    def adjoint(self) -> bool:
        return self.state["adjoint"]

    @property
    def T(self):
        op = copy(self)
        op.state = copy(self.state)
        op.state["adjoint"] = not self.state["adjoint"]
        op.input_shape, op.output_shape = op.output_shape, op.input_shape
        return op

    def _repr_name_(self) -> str:
        return f"{self.name}.T" if self.adjoint else self.name

    def _ensure_device_struct(self, queue: dev.Queue, dtype) -> dict[str, Any]:
        self._ensure_host_struct(queue)
        dtype = np.dtype(dtype)
        key = (str(queue.device), dtype)
        if key not in self._device_struct:
            hs = self._host_struct
            tables = {
                "ofs": dev.torch.from_numpy(np.ravel(hs["ofs_dict"][dtype], order="F").copy()).to(queue.device),
                "geo": dev.torch.from_numpy(np.ravel(hs["geo_dict"][dtype]).copy()).to(queue.device),
                "angle_weights": dev.torch.from_numpy(
                    np.ravel(hs["angle_diff_dict"][dtype]).copy()).to(queue.device),
            }
            if "sdpd_dict" in hs:
                tables["sdpd"] = dev.torch.from_numpy(
                    np.ravel(hs["sdpd_dict"][dtype], order="F").copy()).to(queue.device)
            self._device_struct[key] = tables
        return self._device_struct[key]

    def _kernel_arguments(self, output, argument, queue: dev.Queue) -> tuple:
        """Geometry tables passed to a plugin projection kernel after (out, in):
        (ofs, geo) for Radon (projection.py:372-379), (ofs, sdpd, geo) for
        Fanbeam (gratopy.py:628-638)."""
        tables = self._ensure_device_struct(queue, dev.to_np_dtype(argument.dtype))
        if self._geometry_is_fan:
            return tables["ofs"], tables["sdpd"], tables["geo"]
        return tables["ofs"], tables["geo"]

    def _launch(self, output, argument, queue: dev.Queue) -> None:
        if not self.kernel_spec.is_builtin:
            super()._launch(output, argument, queue)
            return
        tables = self._ensure_device_struct(queue, dev.to_np_dtype(argument.dtype))
        sino_shape = (self.detectors.number, len(self.angles))
        launch_projection(self._geometry_is_fan, self.adjoint, output, argument, tables,
                          tuple(self.image_domain.size), sino_shape, queue.torch_stream())

    def apply_to(self, argument, output=None, queue=None, return_event=False, **kwargs):
        queue = self._infer_queue(argument=argument, output=output, queue=queue)
        self.projection_settings = SimpleNamespace(queue=queue)
        return super().apply_to(argument, output=output, queue=queue, return_event=return_event)


class Radon(_ProjectionOperator):
    """Parallel-beam Radon transform; ``R.T`` is its adjoint (backprojection).

    ``image_domain``: int, (Nx, Ny) or :class:`ImageDomain` (default extent 2.0).
    ``angles``: int (uniform half circle) or :class:`Angles`.
    ``detectors``: None (ceil(hypot(Nx, Ny)) pixels spanning the image extent),
    int, or :class:`Detectors`.  One of the two extents may be an
    :class:`ExtentPlaceholder`, resolved by :meth:`substitute_placeholder`.
    """

    def __init__(self, image_domain, angles, detectors=None, adjoint=False,
                 kernel_spec: HipKernelSpec | None = None):
        if not isinstance(image_domain, ImageDomain):
            image_domain = ImageDomain(size=image_domain, extent=2.0)
        if not isinstance(angles, Angles):
            angles = Angles.uniform(number=angles, half_circle=True)
        if not isinstance(detectors, Detectors):
            if detectors is None:
                detectors = int(np.ceil(np.hypot(*image_domain.size)))
            extent = image_domain.extent
            if isinstance(extent, ExtentPlaceholder):
                extent = ExtentPlaceholder.FULL
            detectors = Detectors(number=detectors, extent=extent)
        state = {"image_domain": image_domain, "angles": angles, "detectors": detectors,
                 "adjoint": adjoint}
        super().__init__(name="Radon", state=state, kernel_spec=kernel_spec)
        self.substitute_placeholder()
        self._init_shapes()

    def _default_kernel_spec(self) -> HipKernelSpec:
        return builtin_kernel_spec("radon")

    def substitute_placeholder(self) -> None:
        """Resolve an ExtentPlaceholder of the image or the detector into a
        number (projection.py:235-318)."""
        img_ext, det_ext = self.image_domain.extent, self.detectors.extent
        if isinstance(img_ext, ExtentPlaceholder) and isinstance(det_ext, ExtentPlaceholder):
            raise NotImplementedError(
                "Both the ImageDomain and Detectors use an ExtentPlaceholder. Please set at least "
                "one of the two extents to a specific value; the remaining placeholder will be "
                "resolved automatically."
            )
        Mx, My = self.image_domain.center
        Md = self.detectors.center
        Nx, Ny = self.image_domain.size
        if isinstance(img_ext, ExtentPlaceholder):
            Dd = float(det_ext)
            M = (Md, Mx, My)
            if img_ext == ExtentPlaceholder.FULL:
                result = full_image_given_detector_halfcircle(M, Dd, Nx / Ny)
                meaning = ("FULL means the largest image extent such that every ray through the "
                           "image also hits the detector")
            else:
                result = valid_image_given_detector_halfcircle(M, Dd, Nx / Ny)
                meaning = ("VALID means the smallest image extent such that every ray hitting the "
                           "detector also passes through the image")
            if result is None:
                raise ValueError(
                    f"Cannot resolve ExtentPlaceholder.{img_ext.name} for the image domain "
                    f"({meaning}): no such image extent exists with the given detector width "
                    f"Dd={Dd}, detector center Md={Md}, and image center ({Mx}, {My}). Consider "
                    "increasing the detector width or adjusting the center offsets."
                )
            self.image_domain.extent = max(result[0], result[1])
        if isinstance(det_ext, ExtentPlaceholder):
            extent = float(self.image_domain.extent)
            Dx = extent * Nx / max(Nx, Ny)
            Dy = extent * Ny / max(Nx, Ny)
            M = (Md, Mx, My)
            if det_ext == ExtentPlaceholder.FULL:
                result = full_detector_given_image_halfcircle(M, (Dx, Dy))
                meaning = ("FULL means the smallest detector width such that every ray through "
                           "the image also hits the detector")
            else:
                result = valid_detector_given_image_halfcircle(M, (Dx, Dy))
                meaning = ("VALID means the largest detector width such that every ray hitting "
                           "the detector also passes through the image")
            if result is None:
                raise ValueError(
                    f"Cannot resolve ExtentPlaceholder.{det_ext.name} for the detector ({meaning}): "
                    f"no such detector width exists with image dimensions ({Dx}, {Dy}), detector "
                    f"center Md={Md}, and image center ({Mx}, {My}). Consider increasing the image "
                    "extent or adjusting the center offsets."
                )
            self.detectors.extent = result

    def _ensure_host_struct(self, queue) -> None:
        if self._host_struct is not None:
            return
        self._host_struct = radon_struct(
            queue=queue, img_shape=tuple(self.image_domain.size), angles=self.angles.angles,
            angle_weights=self.angles.weights, n_detectors=self.detectors.number,
            detector_width=float(self.detectors.extent), image_width=float(self.image_domain.extent),
            midpoint_shift=self.image_domain.center, detector_shift=self.detectors.center,
        )


class Fanbeam(_ProjectionOperator):
    """Fan-beam transform; ``F.T`` is its adjoint.

    ``source_distances``: R (source-detector distance, then RE = R/2) or
    (R, RE).  ``image_domain`` as for :class:`Radon` but a plain int/tuple
    gets the automatic image extent of the reference's ProjectionSettings
    (the largest image inside the fan); ``angles``: int (uniform full
    circle) or :class:`Angles`; ``detectors``: None (ceil(hypot) pixels,
    width 2*RE... see below), int, or :class:`Detectors` (its ``reversed``
    flag flips the detector like ``reverse_detector``).  A default detector
    spans 2 physical units like the reference's ``detector_width`` default.
    """

    _geometry_is_fan = True

    def __init__(self, source_distances, image_domain, angles, detectors=None, adjoint=False,
                 kernel_spec: HipKernelSpec | None = None):
        if np.isscalar(source_distances):
            R, RE = float(source_distances), float(source_distances) / 2
        else:
            R, RE = (float(v) for v in source_distances)
        if not isinstance(image_domain, ImageDomain):
            image_domain = ImageDomain(size=image_domain, extent=None)
        if not isinstance(angles, Angles):
            angles = Angles.uniform(number=angles, half_circle=False)
        if not isinstance(detectors, Detectors):
            if detectors is None:
                detectors = int(np.ceil(np.hypot(*image_domain.size)))
            detectors = Detectors(number=detectors, extent=2.0)
        if isinstance(image_domain.extent, ExtentPlaceholder) or isinstance(detectors.extent, ExtentPlaceholder):
            raise NotImplementedError("extent placeholders are supported for the Radon operator only")
        state = {"image_domain": image_domain, "angles": angles, "detectors": detectors,
                 "adjoint": adjoint, "source_distances": (R, RE)}
        super().__init__(name="Fanbeam", state=state, kernel_spec=kernel_spec)
        self._init_shapes()

    def _default_kernel_spec(self) -> HipKernelSpec:
        return builtin_kernel_spec("fanbeam")

    def _ensure_host_struct(self, queue) -> None:
        if self._host_struct is not None:
            return
        R, RE = self.state["source_distances"]
        extent = self.image_domain.extent
        self._host_struct = fanbeam_struct(
            queue=queue, img_shape=tuple(self.image_domain.size), angles=self.angles.angles,
            detector_width=float(self.detectors.extent), source_detector_dist=R,
            source_origin_dist=RE, angle_weights=self.angles.weights,
            n_detectors=self.detectors.number, detector_shift=self.detectors.center,
            image_width=None if extent is None else float(extent),
            midpoint_shift=self.image_domain.center, reverse_detector=self.detectors.reversed,
        )
