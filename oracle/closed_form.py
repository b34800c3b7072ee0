"""Dense closed-form evaluation of the projection operators -- TEST ONLY.

Derived from the geometry (physical source, detector and pixel positions),
not from the reference's loop structure or its ofs tables, so that it pins
the strip bounds, axis swaps and incremental recurrences of the C oracle:

* parallel beam: a pixel centre p projects to the detector coordinate
  s(theta) = p_x sin(theta) - p_y cos(theta); with the continuous detector
  index u = (s - detector_shift)/delta_s + (Ns-1)/2 the interpolation
  weight of bin k is hat(u - k) = max(0, 1 - |u - k|);
* fan beam: source S = -RE (cos, sin), detector line at distance R from S
  with direction e = orient (sin, -cos); the ray S -> p hits the detector at
  xi = ((p-S).e) R / ((p-S).n), n = (cos, sin); t = (p-S).n / R; weight
  hat(u - k)/t with u = xi/delta_s + mid_det.

Forward: sino = scale * A img; backward: img = A^T (weights * sino), see the
per-geometry scale factors below (gratopy.py:553-569, 950-988; the
reference kernels radon.cl:159-160, fanbeam.cl:254-255, 384-387).
"""

from __future__ import annotations

import numpy as np


def _pixel_centres(img_shape, delta_x, midpoint_shift):
    Nx, Ny = img_shape
    px = (np.arange(Nx) - (Nx - 1) / 2.0) * delta_x + midpoint_shift[0]
    py = (np.arange(Ny) - (Ny - 1) / 2.0) * delta_x + midpoint_shift[1]
    return np.meshgrid(px, py, indexing="ij")  # (Nx, Ny) arrays


def _hat(u, n_bins):
    """(n_bins, ...) array hat(u - k) for k = 0..n_bins-1."""
    k = np.arange(n_bins).reshape((n_bins,) + (1,) * u.ndim)
    return np.maximum(0.0, 1.0 - np.abs(u[None] - k))


def radon_matrix(img_shape, angles, n_detectors, detector_width, image_width,
                 midpoint_shift=(0.0, 0.0), detector_shift=0.0):
    """A[k, a, x, y] of the parallel-beam transform (float64)."""
    delta_x = image_width / max(img_shape)
    delta_s = detector_width / n_detectors
    PX, PY = _pixel_centres(img_shape, delta_x, midpoint_shift)
    A = np.zeros((n_detectors, len(angles)) + tuple(img_shape))
    for a, th in enumerate(angles):
        s = PX * np.sin(th) - PY * np.cos(th)
        u = (s - detector_shift) / delta_s + (n_detectors - 1) / 2.0
        A[:, a] = _hat(u, n_detectors)
    return A


def radon_scale(img_shape, n_detectors, detector_width, image_width, float_quirk=True):
    """delta_x^2/delta_s; the reference kernel reads both as float even in
    double precision (radon.cl:82-83), reproduced when float_quirk is set."""
    delta_x = image_width / max(img_shape)
    delta_s = detector_width / n_detectors
    if float_quirk:
        dx = float(np.float32(delta_x))
        return dx * dx / float(np.float32(delta_s))
    return delta_x**2 / delta_s


def fan_matrix(img_shape, angles, n_detectors, detector_width, R, RE, image_width,
               midpoint_shift=(0.0, 0.0), detector_shift=0.0, reverse_detector=False):
    """(A[k, a, x, y], sdpd[k]) of the fan-beam transform in pixel units."""
    pixels = max(img_shape)
    unit = pixels / float(image_width)  # physical -> pixel units
    Rp, REp, dwp = R * unit, RE * unit, detector_width * unit
    delta_xi = dwp / n_detectors
    mid_det = (n_detectors - 1.0) / 2.0 - detector_shift * n_detectors / detector_width
    PX, PY = _pixel_centres(img_shape, 1.0, (midpoint_shift[0] * unit, midpoint_shift[1] * unit))
    orient = -1.0 if reverse_detector else 1.0
    A = np.zeros((n_detectors, len(angles)) + tuple(img_shape))
    for a, th in enumerate(angles):
        n = np.array([np.cos(th), np.sin(th)])
        e = orient * np.array([np.sin(th), -np.cos(th)])
        Sx, Sy = -REp * n
        dx, dy = PX - Sx, PY - Sy
        depth = dx * n[0] + dy * n[1]
        xi = (dx * e[0] + dy * e[1]) * Rp / depth
        t = depth / Rp
        u = xi / delta_xi + mid_det
        A[:, a] = _hat(u, n_detectors) / t[None]
    xi_k = (np.arange(n_detectors) - mid_det) * delta_xi
    sdpd = np.sqrt(xi_k**2 + Rp**2)
    return A, sdpd, Rp, delta_xi, float(image_width) / pixels


def radon_fwd(img, A, scale):
    return scale * np.einsum("kaxy,xy->ka", A, img)


def radon_back(sino, A, angle_weights):
    return np.einsum("kaxy,ka->xy", A, sino * angle_weights[None, :])


def fan_fwd(img, A, sdpd, Rp, delta_xi, delta_x):
    acc = np.einsum("kaxy,xy->ka", A, img)
    return acc * sdpd[:, None] * delta_x / (delta_xi * Rp)


def fan_back(sino, A, sdpd, Rp, angle_weights):
    weighted = sino * sdpd[:, None] * angle_weights[None, :] / Rp
    return np.einsum("kaxy,ka->xy", A, weighted)
