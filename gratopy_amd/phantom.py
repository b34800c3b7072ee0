"""Ellipse phantoms (gratopy/phantom.py:33-215), rasterised on the host and
copied to the device.

``ct_shepp_logan`` / ``easy_phantom`` follow the reference's ellipse rule:
a field of view [-1, 1]^2 sampled by ``meshgrid(linspace(-1, 1, N),
-linspace(-1, 1, M))`` and a point belongs to ellipse (grey, a, b, xc, yc,
theta) when ((X-xc)cos + (Y-yc)sin)^2/a^2 + ((X-xc)sin - (Y-yc)cos)^2/b^2 <= 1.
``random_ellipses`` draws the synthetic phantoms of the benchmark workloads
from a seeded generator with the same rule.
"""

from __future__ import annotations

import numpy as np

from . import _device as dev


def ct_shepp_logan_params_2d():
    """Ellipses [grey, a, b, xc, yc, theta] of the original Shepp-Logan phantom
    (Shepp & Logan 1974; gratopy/phantom.py:186-202)."""
    E = np.zeros((10, 6))
    E[:, 0] = [2, -0.98, -0.02, -0.02, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01]
    E[:, 1] = [0.69, 0.6624, 0.11, 0.16, 0.21, 0.046, 0.046, 0.046, 0.023, 0.023]
    E[:, 2] = [0.92, 0.874, 0.31, 0.41, 0.25, 0.046, 0.046, 0.023, 0.023, 0.046]
    E[:, 3] = [0, 0, 0.22, -0.22, 0, 0, 0, -0.08, 0, 0.06]
    E[:, 4] = [0, -0.0184, 0, 0, 0.35, 0.1, -0.1, -0.605, -0.605, -0.605]
    E[:, 5] = np.deg2rad([0, 0, -18, 18, 0, 0, 0, 0, 0, 0])
    return E


def ct_modified_shepp_logan_params_2d():
    """Same ellipses with the higher-contrast grey values (Toft 1996, table B.1)."""
    E = ct_shepp_logan_params_2d()
    E[:, 0] = [1, -0.8, -0.2, -0.2, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1]
    return E


def rasterize_ellipses(M: int, N: int, E: np.ndarray) -> np.ndarray:
    """Host (M, N) float64 image: sum of the ellipses of ``E`` on [-1, 1]^2.

    Each ellipse is tested only inside its bounding box (widened by two
    grid steps); the per-pixel test is the same expression as on the full
    grid, so the result is identical.
    """
    xs = np.linspace(-1, 1, N)
    ys = -np.linspace(-1, 1, M)
    ph = np.zeros((M, N))
    ct, st = np.cos(E[:, 5]), np.sin(E[:, 5])
    step_x = 2.0 / max(N - 1, 1)
    step_y = 2.0 / max(M - 1, 1)
    for k in range(E.shape[0]):
        grey, a, b, xc, yc = E[k, 0], E[k, 1], E[k, 2], E[k, 3], E[k, 4]
        hx = np.sqrt((a * ct[k]) ** 2 + (b * st[k]) ** 2) + 2 * step_x
        hy = np.sqrt((a * st[k]) ** 2 + (b * ct[k]) ** 2) + 2 * step_y
        cols = np.nonzero(np.abs(xs - xc) <= hx)[0]
        rows = np.nonzero(np.abs(ys - yc) <= hy)[0]
        if len(cols) == 0 or len(rows) == 0:
            continue
        c0, c1 = cols[0], cols[-1] + 1
        r0, r1 = rows[0], rows[-1] + 1
        X, Y = np.meshgrid(xs[c0:c1], ys[r0:r1])
        inside = ((X - xc) * ct[k] + (Y - yc) * st[k]) ** 2 / a**2 + (
            (X - xc) * st[k] - (Y - yc) * ct[k]
        ) ** 2 / b**2 <= 1
        ph[r0:r1, c0:c1][inside] += grey
    return ph


def ct_shepp_logan(queue, N, modified=True, E=None, ret_E=False, dtype="double", allocator=None):
    """Shepp-Logan phantom of size (N, N) or (M, N) as a device array
    (gratopy/phantom.py:58-183)."""
    if np.isscalar(N):
        M, N = N, N
    else:
        if len(N) != 2:
            raise ValueError("Dimension must be scalar or 2D!")
        M, N = N[:]
    if E is None:
        E = ct_modified_shepp_logan_params_2d() if modified else ct_shepp_logan_params_2d()
    E = np.asarray(E)
    ph = rasterize_ellipses(M, N, E).astype(dtype)
    q = dev.Queue(queue)
    out = dev.to_device(ph, q.device, order="C")
    if ret_E:
        return (out, E)
    return out


def easy_phantom(N, queue=None, modified=True, E=None, ret_E=False, dtype=None, allocator=None):
    """Shepp-Logan phantom with a default queue and float64 default dtype
    (gratopy/phantom.py:33-55)."""
    if dtype is None:
        dtype = "float64"
    return ct_shepp_logan(queue, N, modified=modified, E=E, ret_E=ret_E, dtype=dtype,
                          allocator=allocator)


def random_ellipse_params(rng: np.random.Generator, n_ellipses: int) -> np.ndarray:
    """E matrix of ``n_ellipses`` random ellipses drawn in the order: centres
    U[-0.6, 0.6]^2, semi-axes U[0.05, 0.4]^2, rotation U[0, pi), grey U[0, 1]."""
    centres = rng.uniform(-0.6, 0.6, size=(n_ellipses, 2))
    axes = rng.uniform(0.05, 0.4, size=(n_ellipses, 2))
    theta = rng.uniform(0.0, np.pi, size=n_ellipses)
    grey = rng.uniform(0.0, 1.0, size=n_ellipses)
    E = np.zeros((n_ellipses, 6))
    E[:, 0] = grey
    E[:, 1:3] = axes
    E[:, 3:5] = centres
    E[:, 5] = theta
    return E


def random_ellipses(N: int, n_ellipses: int, seed: int = 0, dtype="float32", rng=None) -> np.ndarray:
    """Host (N, N) phantom: sum of ``n_ellipses`` random ellipses (seeded)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    return rasterize_ellipses(N, N, random_ellipse_params(rng, n_ellipses)).astype(dtype)


def add_noise(f: np.ndarray, percent: float, rng: np.random.Generator) -> np.ndarray:
    """f + (percent/100) * (||f||_2 / sqrt(n)) * N(0, 1), drawn from ``rng``."""
    scale = percent / 100.0 * np.linalg.norm(f) / np.sqrt(f.size)
    return (f + scale * rng.standard_normal(f.shape)).astype(f.dtype)
