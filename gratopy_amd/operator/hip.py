"""HIP execution pipeline of the operator interface.

Counterpart of gratopy/operator/opencl.py: :class:`HipKernelSpec` names the
kernel bundle of an operator (content-addressed, like ``OpenCLKernelSpec``,
opencl.py:55-155) and :class:`_HipOperator` implements the shared
``apply_to`` pipeline (opencl.py:398-441): infer the device queue, move
host arrays to the device, validate shapes, allocate the output in the
argument's order, enqueue the kernels, then apply the operator's scalar.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from gratopy_amd import _device as dev
from gratopy_amd import _hip
from gratopy_amd.operator.base import Operator


@dataclass(frozen=True)
class HipKernelSpec:
    """Kernel bundle of an operator: source or library paths plus the base
    name that selects ``{base}[_ad]_{float|double}_{out_order}{in_order}``.

    ``signature`` hashes the base name, the build options and the path and
    content of every file, so a changed file yields a new cache key
    (opencl.py:142-155).
    """

    paths: tuple[str, ...]
    base_name: str
    build_options: tuple[str, ...] = ()

    def __post_init__(self) -> None:
        object.__setattr__(self, "paths", tuple(str(Path(p).expanduser().resolve()) for p in self.paths))
        object.__setattr__(self, "build_options", tuple(self.build_options))

    @classmethod
    def from_path(cls, path, base_name: str, build_options=()) -> "HipKernelSpec":
        return cls(paths=(str(path),), base_name=base_name, build_options=build_options)

    @classmethod
    def from_paths(cls, paths, base_name: str, build_options=()) -> "HipKernelSpec":
        return cls(paths=tuple(str(p) for p in paths), base_name=base_name, build_options=build_options)

    def read_sources(self) -> tuple[bytes, ...]:
        return tuple(Path(p).read_bytes() for p in self.paths)

    @property
    def signature(self) -> str:
        h = hashlib.sha256()
        h.update(self.base_name.encode())
        for opt in self.build_options:
            h.update(b"\0")
            h.update(opt.encode())
        for path, content in zip(self.paths, self.read_sources()):
            h.update(b"\0")
            h.update(path.encode())
            h.update(b"\0")
            h.update(content)
        return h.hexdigest()

    def kernel_name(self, dtype, output_order: str, input_order: str, adjoint: bool = False) -> str:
        precision = "float" if np.dtype(dtype) == np.dtype("float32") else "double"
        ad = "_ad" if adjoint else ""
        return f"{self.base_name}{ad}_{precision}_{output_order.lower()}{input_order.lower()}"


def builtin_kernel_spec(base_name: str) -> HipKernelSpec:
    """Spec of a transform compiled into libgratopy_hip.so."""
    return HipKernelSpec.from_path(_hip.library_path(), base_name=base_name)


class _HipOperator(Operator):
    """Shared execution pipeline of HIP-backed operators.  Subclasses
    implement :meth:`_launch` (enqueue the kernels writing ``output``)."""

    def __init__(self, *, name=None, kernel_spec: HipKernelSpec | None = None, **operator_kwargs):
        super().__init__(name=name, **operator_kwargs)
        self.kernel_spec = kernel_spec or self._default_kernel_spec()
        self._last_queue = None

    def _default_kernel_spec(self) -> HipKernelSpec:
        raise NotImplementedError("Concrete HIP operators must define a kernel spec")

    def __getstate__(self) -> dict[str, Any]:
        state = self.__dict__.copy()
        state["_last_queue"] = None
        return state

    def _infer_queue(self, argument=None, output=None, queue=None) -> dev.Queue:
        if queue is not None:
            self._last_queue = dev.Queue(queue)
            return self._last_queue
        if dev.is_device_array(argument) and argument.device.type == "cuda":
            self._last_queue = dev.Queue(argument.device)
            return self._last_queue
        if dev.is_device_array(output) and output.device.type == "cuda":
            self._last_queue = dev.Queue(output.device)
            return self._last_queue
        if self._last_queue is not None:
            return self._last_queue
        raise ValueError(
            "No device queue available. Either pass an explicit queue, provide a device array "
            "as input, or provide an output array with an associated device."
        )

    @staticmethod
    def _coerce_argument(argument, queue: dev.Queue):
        """Host array-likes are copied to the queue's device, keeping their order."""
        if dev.is_device_array(argument):
            if argument.device.type != "cuda":
                argument = argument.numpy()
            else:
                return argument
        array = np.asarray(argument)
        if not array.flags["C_CONTIGUOUS"] and not array.flags["F_CONTIGUOUS"]:
            array = np.ascontiguousarray(array)
        return dev.to_device(array, queue.device)

    @staticmethod
    def _default_order(reference) -> str:
        return "C" if dev.is_c_contiguous(reference) else "F"

    @staticmethod
    def _allocate_output(queue: dev.Queue, shape, dtype, order="F"):
        return dev.empty(shape, dtype, order=order, device=queue.device)

    def _expected_output_shape(self, argument_shape) -> tuple[int, ...]:
        if self.input_shape is None or self.output_shape is None:
            raise NotImplementedError(
                "Concrete HIP operators must define _expected_output_shape() or set "
                "input_shape/output_shape appropriately."
            )
        return tuple(self.output_shape) + tuple(argument_shape[len(self.input_shape):])

    def _validate_argument(self, argument) -> None:
        if self.input_shape is None:
            return
        if tuple(argument.shape[0 : len(self.input_shape)]) != tuple(self.input_shape):
            raise ValueError(
                f"Input shape mismatch: expected {self.input_shape}, got {tuple(argument.shape)}"
            )

    def _validate_output(self, output, argument):
        expected = self._expected_output_shape(tuple(argument.shape))
        if output.dtype != argument.dtype:
            raise ValueError(f"Output dtype mismatch: expected {argument.dtype}, got {output.dtype}")
        if tuple(output.shape) != tuple(expected):
            raise ValueError(f"Output shape mismatch: expected {expected}, got {tuple(output.shape)}")
        return output

    def _launch(self, output, argument, queue: dev.Queue) -> None:
        raise NotImplementedError

    def apply_to(self, argument, output=None, queue=None, return_event=False, **kwargs):
        queue = self._infer_queue(argument=argument, output=output, queue=queue)
        stream = queue.torch_stream()
        with dev.torch.cuda.stream(stream):
            argument = self._coerce_argument(argument, queue)
            self._validate_argument(argument)
            if output is None:
                output = self._allocate_output(queue, self._expected_output_shape(tuple(argument.shape)),
                                               argument.dtype, order=self._default_order(argument))
            output = self._validate_output(output, argument)
            self._launch(output, argument, queue)
            if self.scalar != 1:
                _hip.call("gt_scale", stream.cuda_stream, _hip.dtype_code(output.dtype),
                          output.data_ptr(), float(self.scalar), output.numel())
            event = dev.torch.cuda.Event()
            event.record(stream)
        if return_event:
            return output, [event]
        return output
