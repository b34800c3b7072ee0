"""CPU oracle of gratopy's projection hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker: the product
(``gratopy_amd``) never calls it.

* ``oracle/oracle.c`` (+ ``oracle_kernels.h``): plain-C restatement of the
  reference OpenCL kernels, built into ``oracle/liboracle.so`` by
  ``oracle/Makefile``.  Bound here with ctypes on numpy arrays.
* ``oracle/closed_form.py``: an independent dense closed-form evaluation of
  the projection operators from the geometry (numpy, float64) used to pin the
  C restatement on small grids.
* ``oracle/solvers.py``: the reference's solver loops (gratopy.py:2329-2762)
  restated on top of the C kernels.

Every function of the C library cites the reference lines it follows.  How
each one is pinned (control numbers of the reference's tests, the closed form,
adjointness and mass invariants) is listed in tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


FAST_PATHS = {v: os.path.join(HERE, f"liboracle_fast_{v}.so") for v in ("v3", "v4")}


def build(force: bool = False) -> str:
    """Compile liboracle.so (the checker) and the two CPU-baseline timing
    builds with the committed Makefile (gcc, OpenMP)."""
    if force or not all(os.path.exists(p) for p in (LIB_PATH, *FAST_PATHS.values())):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as fh:
            return any(line.startswith("flags") and " avx512f" in line for line in fh)
    except OSError:
        return False


class timing_build:
    """Context manager: the oracle functions run on the CPU-baseline timing
    build (FMA contraction allowed; AVX-512 where the host has it) instead of
    the checker.  Only bench.py's CPU baseline uses it."""

    def __enter__(self):
        global _lib
        variant = "v4" if _has_avx512() else "v3"
        path = FAST_PATHS[variant]
        if not os.path.exists(path):
            build()
        self._saved = _lib
        _lib = ctypes.CDLL(path)
        self.variant = variant
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def order_of(a: np.ndarray) -> int:
    """0 for Fortran order, 1 for C order (C wins for arrays that are both,
    as in the reference's ``{0: 'F', 1: 'C'}[flags.c_contiguous]``)."""
    if a.flags.c_contiguous:
        return 1
    if a.flags.f_contiguous:
        return 0
    raise ValueError("oracle arrays must be C or F contiguous")


def _dims3(shape):
    return (shape[0], shape[1], shape[2] if len(shape) > 2 else 1)


def _call(name, *args):
    fn = getattr(lib(), name)
    fn.restype = None
    fn(*args)


def _table(t, dtype):
    return np.require(t, dtype, "F")


def radon_fwd(img, ofs, geo, n_detectors, n_angles, sino_order="F"):
    """Reference kernel radon_* (gratopy/radon.cl:66-161) on the CPU."""
    dt = img.dtype
    Nx, Ny, Nz = _dims3(img.shape)
    shape = (n_detectors, n_angles) + ((Nz,) if img.ndim > 2 else ())
    sino = np.zeros(shape, dt, order=sino_order)
    _call(f"oracle_radon_fwd_{_sfx(dt)}", _ptr(sino), _ptr(img), _ptr(_table(ofs, dt)),
          _ptr(_table(geo, dt)), Nx, Ny, n_detectors, n_angles, Nz, order_of(sino), order_of(img))
    return sino


def radon_back(sino, ofs, geo, img_shape, img_order="F"):
    """Reference kernel radon_ad_* (gratopy/radon.cl:183-241) on the CPU."""
    dt = sino.dtype
    Ns, Na, Nz = _dims3(sino.shape)
    shape = tuple(img_shape) + ((Nz,) if sino.ndim > 2 else ())
    img = np.zeros(shape, dt, order=img_order)
    _call(f"oracle_radon_back_{_sfx(dt)}", _ptr(img), _ptr(sino), _ptr(_table(ofs, dt)),
          _ptr(_table(geo, dt)), img_shape[0], img_shape[1], Ns, Na, Nz, order_of(img),
          order_of(sino))
    return img


def fan_fwd(img, ofs, sdpd, geo, n_detectors, n_angles, sino_order="F"):
    """Reference kernel fanbeam_* (gratopy/fanbeam.cl:72-256) on the CPU."""
    dt = img.dtype
    Nx, Ny, Nz = _dims3(img.shape)
    shape = (n_detectors, n_angles) + ((Nz,) if img.ndim > 2 else ())
    sino = np.zeros(shape, dt, order=sino_order)
    _call(f"oracle_fan_fwd_{_sfx(dt)}", _ptr(sino), _ptr(img), _ptr(_table(ofs, dt)),
          _ptr(_table(sdpd, dt)), _ptr(_table(geo, dt)), Nx, Ny, n_detectors, n_angles, Nz,
          order_of(sino), order_of(img))
    return sino


def fan_back(sino, ofs, sdpd, geo, img_shape, img_order="F"):
    """Reference kernel fanbeam_ad_* (gratopy/fanbeam.cl:290-391) on the CPU."""
    dt = sino.dtype
    Ns, Na, Nz = _dims3(sino.shape)
    shape = tuple(img_shape) + ((Nz,) if sino.ndim > 2 else ())
    img = np.zeros(shape, dt, order=img_order)
    _call(f"oracle_fan_back_{_sfx(dt)}", _ptr(img), _ptr(sino), _ptr(_table(ofs, dt)),
          _ptr(_table(sdpd, dt)), _ptr(_table(geo, dt)), img_shape[0], img_shape[1], Ns, Na, Nz,
          order_of(img), order_of(sino))
    return img


def single_line(geometry_is_fan, x, y, ofs, sdpd, geo, img_shape, n_detectors, n_angles, dtype,
                sino_order="F"):
    """Reference kernels single_line_radon_* / single_line_fanbeam_*
    (gratopy/radon.cl:263-341, gratopy/fanbeam.cl:422-568)."""
    dt = np.dtype(dtype)
    sino = np.zeros((n_detectors, n_angles), dt, order=sino_order)
    if geometry_is_fan:
        _call(f"oracle_single_line_fan_{_sfx(dt)}", _ptr(sino), int(x), int(y), _ptr(_table(ofs, dt)),
              _ptr(_table(sdpd, dt)), _ptr(_table(geo, dt)), img_shape[0], img_shape[1],
              n_detectors, n_angles, order_of(sino))
    else:
        _call(f"oracle_single_line_radon_{_sfx(dt)}", _ptr(sino), int(x), int(y),
              _ptr(_table(ofs, dt)), _ptr(_table(geo, dt)), img_shape[0], img_shape[1],
              n_detectors, n_angles, order_of(sino))
    return sino


def weight_sino(sino, angle_weights, divide=False):
    """Reference kernels multiply_* / divide_* (gratopy/utilities.cl:38-67)."""
    dt = sino.dtype
    out = np.zeros_like(sino)
    Ns, Na, Nz = _dims3(sino.shape)
    w = np.require(angle_weights, dt, "C")
    _call(f"oracle_weight_sino_{_sfx(dt)}", _ptr(out), _ptr(sino), _ptr(w), Ns, Na, Nz,
          order_of(sino), int(divide))
    return out


def equ_mul_add(rhs, a, x):
    """Reference kernel equ_mul_add_* (gratopy/utilities.cl:69-73), in place."""
    dt = rhs.dtype
    ct = ctypes.c_float if dt == np.float32 else ctypes.c_double
    _call(f"oracle_equ_mul_add_{_sfx(dt)}", _ptr(rhs), ct(a), _ptr(x), ctypes.c_size_t(rhs.size))
    return rhs


def mul_add_add(rhs, a, x, y):
    """Reference kernel mul_add_add_* (gratopy/utilities.cl:75-80), in place."""
    dt = rhs.dtype
    ct = ctypes.c_float if dt == np.float32 else ctypes.c_double
    _call(f"oracle_mul_add_add_{_sfx(dt)}", _ptr(rhs), ct(a), _ptr(x), _ptr(y),
          ctypes.c_size_t(rhs.size))
    return rhs


def tv_update_v(v, u, sigma, z_distance):
    """Reference kernel update_v_* (gratopy/total_variation.cl:54-93), in place.
    ``v`` is the (4,)+img_shape dual array of the same order as ``u``."""
    Nx, Ny, Nz = _dims3(u.shape)
    _call(f"oracle_tv_update_v_{_sfx(u.dtype)}", _ptr(v), _ptr(u), ctypes.c_float(sigma),
          ctypes.c_float(z_distance), Nx, Ny, Nz, order_of(u))


def tv_update_normv(v, img_shape, order):
    """Reference kernel update_NormV_unchor_* (gratopy/total_variation.cl:150-167)."""
    Nx, Ny, Nz = _dims3(img_shape)
    _call(f"oracle_tv_update_normv_{_sfx(v.dtype)}", _ptr(v), Nx, Ny, Nz, 0 if order == "F" else 1)


def tv_update_lambda(lamb, ku, f, sigma, sigmap1inv):
    """Reference kernel update_lambda_L2_* (gratopy/total_variation.cl:95-106)."""
    _call(f"oracle_tv_update_lambda_{_sfx(lamb.dtype)}", _ptr(lamb), _ptr(ku), _ptr(f),
          ctypes.c_float(sigma), ctypes.c_float(sigmap1inv), ctypes.c_size_t(lamb.size))


def tv_update_u(u, u_, v, kstar, tau, norming, z_distance):
    """Reference kernel update_u_* (gratopy/total_variation.cl:108-148)."""
    Nx, Ny, Nz = _dims3(u.shape)
    _call(f"oracle_tv_update_u_{_sfx(u.dtype)}", _ptr(u), _ptr(u_), _ptr(v), _ptr(kstar),
          ctypes.c_float(tau), ctypes.c_float(norming), ctypes.c_float(z_distance), Nx, Ny, Nz,
          order_of(u))


def tv_update_extra(u_, u):
    """ElementwiseKernel update_extra_ of gratopy/gratopy.py:2708-2717: u = 2 u_ - u."""
    _call(f"oracle_tv_update_extra_{_sfx(u.dtype)}", _ptr(u_), _ptr(u), ctypes.c_size_t(u.size))
