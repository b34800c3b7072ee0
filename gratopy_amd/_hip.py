"""ctypes binding of ``libgratopy_hip.so``, the C-ABI of the HIP kernels.

The library is built in-tree by ``__graft_entry__.build()`` (hipcc,
gfx950).  There is no fallback: if the library is absent or no GPU is
visible, every compute entry point raises.  Device arrays are torch tensors;
only raw pointers, sizes and the HIP stream handle cross the boundary.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

LIB_NAME = "libgratopy_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# enum values shared with include/gratopy_hip.h
DTYPE_F32 = 0
DTYPE_F64 = 1
ORDER_F = 0
ORDER_C = 1
GEOM_RADON = 0
GEOM_FAN = 1

# status codes of include/gratopy_hip.h
GT_OK = 0
GT_EINVAL = 1
GT_ENOSPC = 2
GT_EHIP = 3

_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
_d = ctypes.c_double
_f = ctypes.c_float

# name -> (restype, argtypes); kept in sync with include/gratopy_hip.h
PROTOTYPES: dict[str, tuple] = {
    "gt_version": (_i, []),
    "gt_tune_set": (_i, [_i, _i]),
    "gt_tune_get": (_i, [_i]),
    "gt_last_error": (_sz, [ctypes.c_char_p, _sz]),
    "gt_projection_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i, _i, _i, _i]),
    "gt_projection_shape": (_i, [_i, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_i)]),
    "gt_radon_fwd": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _sz]),
    "gt_radon_back": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _sz]),
    "gt_fan_fwd": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _sz]),
    "gt_fan_back": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _sz]),
    "gt_projection_noalloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i,
                                   _i, _i]),
    "gt_single_line": (_i, [_vp, _i, _i, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i]),
    "gt_weight_sino": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i]),
    "gt_axpy": (_i, [_vp, _i, _vp, _d, _vp, _sz]),
    "gt_axpby": (_i, [_vp, _i, _vp, _d, _vp, _vp, _sz]),
    "gt_axpy_dev": (_i, [_vp, _i, _vp, _vp, _d, _vp, _sz]),
    "gt_axpby_dev": (_i, [_vp, _i, _vp, _vp, _d, _vp, _vp, _sz]),
    "gt_scale_dev": (_i, [_vp, _i, _vp, _vp, _i, _sz]),
    "gt_scale": (_i, [_vp, _i, _vp, _d, _sz]),
    "gt_fill": (_i, [_vp, _i, _vp, _d, _sz]),
    "gt_sub": (_i, [_vp, _i, _vp, _vp, _vp, _sz]),
    "gt_reduce_workspace_bytes": (_sz, [_sz]),
    "gt_dot": (_i, [_vp, _i, _vp, _vp, _vp, _sz, _vp, _sz]),
    "gt_sum": (_i, [_vp, _i, _vp, _vp, _sz, _vp, _sz]),
    "gt_cg_scalars": (_i, [_vp, _vp, _d]),
    "gt_tv_update_v": (_i, [_vp, _i, _vp, _vp, _f, _f, _i, _i, _i, _i]),
    "gt_tv_update_normv": (_i, [_vp, _i, _vp, _vp, _i, _i, _i, _i]),
    "gt_tv_update_lambda": (_i, [_vp, _i, _vp, _vp, _vp, _f, _f, _i, _i, _i, _i]),
    "gt_tv_update_u": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _f, _f, _f, _i, _i, _i, _i]),
    "gt_tv_update_extra": (_i, [_vp, _i, _vp, _vp, _sz]),
    "gt_tv_dual": (_i, [_vp, _i, _vp, _vp, _f, _f, _i, _i, _i, _i]),
    "gt_tv_primal": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _f, _f, _f, _i, _i, _i, _i]),
    "gt_tv_adj_primal": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _sz]),
    "gt_tv_adj_dual": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _i, _i, _i, _i]),
    "gt_tv_adj_div": (_i, [_vp, _i, _vp, _vp, _d, _vp, _f, _f, _i, _i, _i, _i]),
    "gt_tv_adj_lambda": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _d, _f, _f, _i, _i, _i, _i]),
    "gt_batch_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _vp]),
    "gt_batch_plan_destroy": (_i, [_vp]),
    "gt_batch_plan_bytes": (_sz, [_vp]),
    "gt_batch_launch": (_i, [_vp, _vp, _i, _d]),
    "gt_plugin_compile": (_i, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                               ctypes.POINTER(ctypes.c_char_p), _i, _vp, _sz,
                               ctypes.POINTER(_sz)]),
    "gt_plugin_device_arch": (_i, [ctypes.c_char_p, _sz]),
    "gt_plugin_module_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "gt_plugin_module_unload": (_i, [_vp]),
    "gt_plugin_function": (_i, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "gt_plugin_launch": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_uint),
                              ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(_vp)]),
}

# operations of gt_batch_launch
BATCH_FORWARD = 0
BATCH_BACKWARD = 1
BATCH_LANDWEBER = 2
BATCH_TV_EVEN = 3
BATCH_TV_ODD = 4


class BatchItem(ctypes.Structure):
    """struct gt_batch_item of include/gratopy_hip.h."""

    _fields_ = [
        ("geometry", _i),
        ("Nx", _i),
        ("Ny", _i),
        ("Ns", _i),
        ("Na", _i),
        ("ofs", _vp),
        ("sdpd", _vp),
        ("geo", _vp),
        ("img", _vp),
        ("sino", _vp),
        ("data", _vp),
        ("img2", _vp),
        ("field", _vp),
        ("p", _d * 6),
    ]

_lib = None
_lock = threading.Lock()


class HipLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or reported a HIP error."""


def library_path() -> str:
    return LIB_PATH


def load_library() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HipLibraryError(
                f"{LIB_NAME} is not built (expected at {LIB_PATH}); "
                "run `python -c 'import __graft_entry__ as g; g.build()'` from the repository root"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def last_error() -> str:
    lib = load_library()
    buf = ctypes.create_string_buffer(lib.gt_last_error(None, 0) + 1)
    lib.gt_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    """Map a C-ABI status code to an exception (ValueError for argument
    errors, HipLibraryError for HIP runtime failures)."""
    if rc == GT_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc in (GT_EINVAL, GT_ENOSPC):
        raise ValueError(msg)
    raise HipLibraryError(msg)


def call(name: str, *args) -> None:
    lib = load_library()
    check(getattr(lib, name)(*args), name)


def dtype_code(dtype) -> int:
    dtype = np.dtype(dtype) if not _is_torch_dtype(dtype) else _torch_to_np(dtype)
    if dtype == np.float32:
        return DTYPE_F32
    if dtype == np.float64:
        return DTYPE_F64
    raise ValueError(f"unsupported dtype {dtype}; only float32 and float64 are supported")


def _is_torch_dtype(dtype) -> bool:
    try:
        import torch

        return isinstance(dtype, torch.dtype)
    except ImportError:  # pragma: no cover
        return False


def _torch_to_np(dtype):
    import torch

    return {torch.float32: np.dtype("float32"), torch.float64: np.dtype("float64")}.get(dtype, dtype)
