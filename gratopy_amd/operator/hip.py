"""HIP execution pipeline of the operator interface.

Counterpart of gratopy/operator/opencl.py: :class:`HipKernelSpec` names the
kernel bundle of an operator (content-addressed, like ``OpenCLKernelSpec``,
opencl.py:55-155) and :class:`_HipOperator` implements the shared
``apply_to`` pipeline (opencl.py:398-441): infer the device queue, move
host arrays to the device, validate shapes, allocate the output in the
argument's order, enqueue the kernels, then apply the operator's scalar.

A spec that names HIP C++ sources (rather than the built-in library) is a
plugin: its sources are compiled at first use with hiprtc for the device's
architecture, loaded as a module per device, and cached by (device,
signature) like the reference's program cache (opencl.py:166, 262-281).
Plugin kernels follow the reference's naming ABI
``{base}[_ad]_{float|double}_{out_order}{in_order}`` and take
``(out, in, extra..., gt_range)``; ``gratopy_amd/csrc/gratopy_plugin.h``
provides ``gt_range`` and a macro that defines the eight variants.
"""

from __future__ import annotations

import ctypes
import hashlib
import struct
import threading
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

    @property
    def is_builtin(self) -> bool:
        """True for the transforms compiled into libgratopy_hip.so."""
        return self.paths == (str(Path(_hip.library_path()).resolve()),)

    def build_code(self) -> str:
        """The plugin translation unit: the sources in order, each behind a
        ``#line`` marker so compiler messages name the original file (the
        reference concatenates its expanded sources, opencl.py:252-260)."""
        parts = []
        for path, content in zip(self.paths, self.read_sources()):
            parts.append(f'#line 1 "{path}"\n')
            parts.append(content.decode())
            parts.append("\n")
        return "".join(parts)

    def compile_options(self) -> tuple[str, ...]:
        """Include paths (the plugin header's directory, then each source's
        directory) followed by the spec's own build options."""
        dirs = [PLUGIN_INCLUDE_DIR]
        for p in self.paths:
            d = str(Path(p).parent)
            if d not in dirs:
                dirs.append(d)
        return tuple(f"-I{d}" for d in dirs) + self.build_options


def builtin_kernel_spec(base_name: str) -> HipKernelSpec:
    """Spec of a transform compiled into libgratopy_hip.so."""
    return HipKernelSpec.from_path(_hip.library_path(), base_name=base_name)


# ---- plugins: run-time compilation, module cache, launch ---------------------

PLUGIN_INCLUDE_DIR = str(Path(__file__).resolve().parent.parent / "csrc")


class GtRange(ctypes.Structure):
    """The ``gt_range`` kernel argument: the output shape (global size)."""

    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int), ("n2", ctypes.c_int)]


def compile_plugin(code: str, name: str, arch: str, options=()) -> bytes:
    """hiprtc-compile a plugin translation unit into a code object for
    ``arch`` (e.g. ``"gfx950"``).  Raises HipLibraryError with the compiler log."""
    lib = _hip.load_library()
    opts = [o.encode() for o in options]
    arr = (ctypes.c_char_p * max(1, len(opts)))(*opts)
    size = ctypes.c_size_t(0)
    src = code.encode()
    _hip.check(lib.gt_plugin_compile(src, name.encode(), arch.encode(), arr, len(opts), None, 0,
                                     ctypes.byref(size)), "gt_plugin_compile")
    buf = ctypes.create_string_buffer(size.value)
    _hip.check(lib.gt_plugin_compile(src, name.encode(), arch.encode(), arr, len(opts), buf,
                                     size.value, ctypes.byref(size)), "gt_plugin_compile")
    return buf.raw[: size.value]


def code_object_kernels(image: bytes) -> list[str]:
    """Kernel names of an AMDGPU code object (ELF64): the symbols whose kernel
    descriptor ``<name>.kd`` is defined, in symbol-table order."""
    if image[:4] != b"\x7fELF" or image[4] != 2 or image[5] != 1:
        raise ValueError("not a little-endian ELF64 code object")
    shoff, = struct.unpack_from("<Q", image, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", image, 0x3A)
    sections = []
    for i in range(shnum):
        off = shoff + i * shentsize
        _name, stype, _flags, _addr, offset, size, link, _info, _align, entsize = struct.unpack_from(
            "<IIQQQQIIQQ", image, off)
        sections.append((stype, offset, size, link, entsize))
    names = []
    for stype, offset, size, link, entsize in sections:
        if stype not in (2, 11) or entsize == 0:  # SHT_SYMTAB, SHT_DYNSYM
            continue
        str_off = sections[link][1]
        for k in range(size // entsize):
            st_name, _info, _other, shndx, _value, _size = struct.unpack_from(
                "<IBBHQQ", image, offset + k * entsize)
            if shndx == 0:
                continue
            end = image.index(b"\0", str_off + st_name)
            sym = image[str_off + st_name:end].decode()
            if sym.endswith(".kd") and sym[:-3] not in names:
                names.append(sym[:-3])
    return names


class PluginProgram:
    """A compiled plugin loaded on one device: ``kernels`` maps each kernel
    name of the code object to its function handle (the reference's
    ``{kernel.function_name: kernel}`` dictionary, opencl.py:277-280)."""

    def __init__(self, spec: HipKernelSpec, device):
        self.signature = spec.signature
        self.device = device
        lib = _hip.load_library()
        with dev.torch.cuda.device(device):
            buf = ctypes.create_string_buffer(256)
            _hip.check(lib.gt_plugin_device_arch(buf, len(buf)), "gt_plugin_device_arch")
            self.arch = buf.value.decode()
            self.image = compile_plugin(spec.build_code(), f"{spec.base_name}.hip", self.arch,
                                        spec.compile_options())
            module = ctypes.c_void_p()
            _hip.check(lib.gt_plugin_module_load(self.image, ctypes.byref(module)),
                       "gt_plugin_module_load")
            self.module = module.value
            self.kernels: dict[str, int] = {}
            for name in code_object_kernels(self.image):
                fn = ctypes.c_void_p()
                _hip.check(lib.gt_plugin_function(self.module, name.encode(), ctypes.byref(fn)),
                           "gt_plugin_function")
                self.kernels[name] = fn.value


def _ctypes_argument(value):
    if isinstance(value, ctypes.Structure):
        return value
    if dev.is_device_array(value):
        if value.device.type != "cuda":
            raise TypeError("plugin kernel arguments must be device arrays")
        return ctypes.c_void_p(value.data_ptr())
    if isinstance(value, np.generic):
        return np.ctypeslib.as_ctypes_type(value.dtype)(value.item())
    raise TypeError(
        f"plugin kernel argument of type {type(value).__name__}: pass device arrays or numpy "
        "scalars (np.int32, np.float32, ...) so that the argument's C type is explicit")


def launch_plugin_kernel(function: int, stream_handle: int, global_shape, *arguments) -> None:
    """Launch ``function(arguments..., gt_range(global_shape))`` over the
    global shape (at most 3 axes) on the given stream: blocks of 64x4 threads
    (256 along axis 0 for one-axis shapes), as many as cover the shape."""
    shape = tuple(int(n) for n in global_shape)
    if len(shape) > 3:
        raise ValueError(f"plugin kernels cover at most 3 axes, got global shape {shape}")
    shape = shape + (1,) * (3 - len(shape))
    if min(shape) == 0:
        return
    block = (64, 4, 1) if shape[1] > 1 else (256, 1, 1)
    grid = tuple(-(-n // b) for n, b in zip(shape, block))
    values = [_ctypes_argument(a) for a in arguments] + [GtRange(*shape)]
    ptrs = (ctypes.c_void_p * len(values))(*[ctypes.cast(ctypes.pointer(v), ctypes.c_void_p)
                                             for v in values])
    _hip.check(_hip.load_library().gt_plugin_launch(
        stream_handle, function, (ctypes.c_uint * 3)(*grid), (ctypes.c_uint * 3)(*block), ptrs),
        "gt_plugin_launch")


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

    # ---- plugin kernels (opencl.py:262-305, 336-396) ----

    def _get_program(self, queue: dev.Queue) -> dict[str, int]:
        """Compiled kernels of the spec on the queue's device, compiled on
        first use and cached by (device, source signature)."""
        device = dev.torch.device(queue.device)
        key = (str(device), self.kernel_spec.signature)
        with _PROGRAM_LOCK:
            program = _PROGRAM_CACHE.get(key)
            if program is None:
                program = PluginProgram(self.kernel_spec, device)
                _PROGRAM_CACHE[key] = program
        return program.kernels

    def _get_kernel(self, output, argument, queue: dev.Queue) -> int:
        """The plugin kernel for the argument's dtype, the output and input
        orders and the operator's direction; KeyError when the plugin does
        not define it (opencl.py:289-305, 351-365)."""
        name = self.kernel_spec.kernel_name(
            dev.to_np_dtype(argument.dtype), self._default_order(output),
            self._default_order(argument), adjoint=getattr(self, "adjoint", False))
        return self._get_program(queue)[name]

    def _kernel_arguments(self, output, argument, queue: dev.Queue) -> tuple:
        """Arguments after (out, in): device arrays or numpy scalars."""
        return ()

    def _global_shape(self, output, argument) -> tuple[int, ...]:
        return tuple(output.shape)

    def _launch(self, output, argument, queue: dev.Queue) -> None:
        """Run the spec's plugin kernel on ``(output, argument, extras...)``."""
        if self.kernel_spec.is_builtin:
            raise NotImplementedError(f"{type(self).__name__} has no built-in transform")
        kernel = self._get_kernel(output, argument, queue)
        launch_plugin_kernel(kernel, queue.handle(), self._global_shape(output, argument),
                             output, argument, *self._kernel_arguments(output, argument, queue))

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


# Compiled plugins by (device, signature).  Invalidation only forgets them: the
# modules stay loaded until the process ends, since HIP graphs captured earlier
# may still launch their kernels.
_PROGRAM_CACHE: dict[tuple[str, str], PluginProgram] = {}
_RETIRED_PROGRAMS: list[PluginProgram] = []
_PROGRAM_LOCK = threading.Lock()


def invalidate_kernel_cache() -> None:
    """Forget every compiled plugin, so the next use compiles its sources
    again (opencl.py:457-459)."""
    with _PROGRAM_LOCK:
        _RETIRED_PROGRAMS.extend(_PROGRAM_CACHE.values())
        _PROGRAM_CACHE.clear()
