"""Device-array plumbing: torch tensors stand in for pyopencl arrays.

The reference's ``queue`` (a ``pyopencl.CommandQueue``) becomes a
:class:`Queue`: a torch CUDA(HIP) device plus an optional stream.  All
kernels of one :class:`~gratopy_amd.ProjectionSettings` run in order on its
stream, which is what the reference's in-order queue and event lists
(``wait_for``) guarantee.
"""

from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

_NP_TO_TORCH = {}
if torch is not None:
    _NP_TO_TORCH = {np.dtype("float32"): torch.float32, np.dtype("float64"): torch.float64}
_TORCH_TO_NP = {v: k for k, v in _NP_TO_TORCH.items()}


def to_torch_dtype(dtype):
    if torch is not None and isinstance(dtype, torch.dtype):
        return dtype
    dt = np.dtype(dtype)
    if dt not in _NP_TO_TORCH:
        raise TypeError(f"unsupported dtype {dt}; gratopy arrays are float32 or float64")
    return _NP_TO_TORCH[dt]


def to_np_dtype(dtype) -> np.dtype:
    if torch is not None and isinstance(dtype, torch.dtype):
        if dtype not in _TORCH_TO_NP:
            raise TypeError(f"unsupported dtype {dtype}; gratopy arrays are float32 or float64")
        return _TORCH_TO_NP[dtype]
    return np.dtype(dtype)


class Queue:
    """Device and stream on which a ProjectionSettings enqueues its work.

    Accepts ``None`` (``cuda:0``, torch's current stream at launch time), a
    device index, a device string / ``torch.device``, a ``torch.cuda.Stream``,
    or another :class:`Queue`.
    """

    def __init__(self, spec=None):
        if isinstance(spec, Queue):
            self.device, self.stream = spec.device, spec.stream
            return
        self.stream = None
        if torch is not None and isinstance(spec, torch.cuda.Stream):
            self.stream = spec
            self.device = spec.device
        elif spec is None:
            self.device = torch.device("cuda", 0) if torch is not None else None
        elif isinstance(spec, int):
            self.device = torch.device("cuda", spec)
        else:
            dev = torch.device(spec)
            if dev.type != "cuda":
                raise ValueError(f"gratopy_amd runs on HIP devices, got {dev}")
            self.device = dev if dev.index is not None else torch.device("cuda", 0)

    def torch_stream(self):
        """The stream kernels go to: the explicit one, else torch's current one."""
        if self.stream is not None:
            return self.stream
        return torch.cuda.current_stream(self.device)

    def handle(self) -> int:
        return self.torch_stream().cuda_stream

    def __repr__(self):
        return f"Queue(device={self.device}, stream={self.stream})"

    def __eq__(self, other):
        return isinstance(other, Queue) and self.device == other.device and self.stream == other.stream

    def __hash__(self):
        return hash((str(self.device), id(self.stream)))


def is_device_array(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


def is_c_contiguous(t) -> bool:
    return t.is_contiguous()


def is_f_contiguous(t) -> bool:
    return t.permute(*reversed(range(t.dim()))).is_contiguous()


def order_of(t) -> str:
    """'C' or 'F', with C preferred when both hold (the reference's
    ``{0: 'F', 1: 'C'}[flags.c_contiguous]``)."""
    if is_c_contiguous(t):
        return "C"
    if is_f_contiguous(t):
        return "F"
    raise ValueError(
        "gratopy arrays must be C- or F-contiguous; got strides "
        f"{tuple(t.stride())} for shape {tuple(t.shape)}"
    )


def order_code(order: str) -> int:
    return 0 if order.upper() == "F" else 1


def empty(shape, dtype, order="F", device=None):
    """Uninitialised device array of the given memory order."""
    shape = tuple(int(n) for n in shape)
    tdt = to_torch_dtype(dtype)
    if order.upper() == "C":
        return torch.empty(shape, dtype=tdt, device=device)
    return torch.empty(shape[::-1], dtype=tdt, device=device).permute(*reversed(range(len(shape))))


def zeros(shape, dtype, order="F", device=None):
    out = empty(shape, dtype, order, device)
    out.zero_()
    return out


def to_device(array, device, order=None):
    """Copy a numpy array to the device, keeping its memory order."""
    a = np.asarray(array)
    if order is None:
        order = "C" if a.flags.c_contiguous else ("F" if a.flags.f_contiguous else "C")
    a = np.require(a, requirements=order.upper())
    if order.upper() == "C":
        return torch.from_numpy(a).to(device)
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to(device)
    return t.permute(*reversed(range(t.dim())))


def to_host(t) -> np.ndarray:
    """Copy a device array back to numpy, keeping its memory order."""
    order = order_of(t)
    if order == "C":
        return t.detach().cpu().numpy()
    return np.asfortranarray(t.detach().permute(*reversed(range(t.dim()))).contiguous().cpu().numpy().T)


def ptr(t) -> int:
    return t.data_ptr()


def workspace(nbytes: int, device):
    """Scratch device memory from torch's caching allocator (current stream)."""
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
