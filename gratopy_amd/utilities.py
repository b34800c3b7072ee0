"""Geometry description objects for the operator interface.

Mirrors the public surface of gratopy's ``gratopy/utilities.py``:
``GeometryType`` (utilities.py:44-50), ``ExtentPlaceholder`` (15-41),
``Angles`` (53-250), ``Detectors`` (255-305), ``ImageDomain`` (309-350) and
the half-circle extent solvers (358-580).  Everything here is host-side
numpy; nothing touches the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import TypeAlias

import numpy as np
import numpy.typing as npt

Numeric: TypeAlias = float | int | np.int_ | np.float32 | np.double


class ExtentPlaceholder(Enum):
    """Symbolic image or detector extent, resolved when an operator is built.

    FULL: every ray through the image domain also hits the detector
    (smallest such detector / largest such image).
    VALID: every ray hitting the detector also passes through the image
    (largest such detector / smallest such image).
    """

    FULL = "full"
    VALID = "valid"


class GeometryType(Enum):
    """Projection geometry: parallel beam (``RADON``) or fan beam (``FANBEAM``)."""

    RADON = "radon"
    FANBEAM = "fanbeam"


def _natural_weights(sorted_angles: np.ndarray, extended: np.ndarray) -> np.ndarray:
    """Half the distance between the two neighbours of every sorted angle,
    with runs of (numerically) equal angles sharing their summed weight.

    ``extended`` is ``sorted_angles`` with one wrap-around neighbour added at
    each end.  Same rule as gratopy.py:1176-1189 / utilities.py:226-247.
    """
    n = len(sorted_angles)
    weights = 0.5 * np.abs(extended[2 : n + 2] - extended[0:n])
    tol = 1e-6
    i = 0
    while i < n - 1:
        j = i + 1
        total = weights[i]
        while abs(sorted_angles[i] - sorted_angles[j]) < tol:
            total += weights[j]
            j += 1
            if j > n - 1:
                break
        weights[i:j] = total / (j - i)
        i = j
    return weights


class Angles:
    """Projection angles (radians) with their quadrature weights.

    ``half_circle`` selects pi- instead of 2pi-periodicity for the
    constructors that generate a complete sampling (utilities.py:53-250).
    """

    def __init__(self, angles: npt.ArrayLike, weights: npt.ArrayLike, half_circle: bool = False):
        angles = np.asarray(angles)
        weights = np.asarray(weights)
        if len(angles) != len(weights):
            raise ValueError("Angles and weights must have the same length.")
        self.angles = angles
        self.weights = weights
        self.half_circle = half_circle

    def __repr__(self):
        return f"Angles({self.angles}, weights={self.weights})"

    def __len__(self):
        return len(self.angles)

    def __eq__(self, other):
        if not isinstance(other, Angles):
            return NotImplemented
        return (
            self.half_circle == other.half_circle
            and np.array_equal(self.angles, other.angles)
            and np.array_equal(self.weights, other.weights)
        )

    @staticmethod
    def sparse(number: int, half_circle: bool = False) -> Angles:
        """Equispaced angles over [0, pi) or [0, 2pi) with unit weights."""
        period = np.pi if half_circle else 2 * np.pi
        grid = np.linspace(0, period, number, endpoint=False)
        return Angles(grid, np.ones_like(grid), half_circle=half_circle)

    @staticmethod
    def uniform(number: int, half_circle: bool = False) -> Angles:
        """Equispaced angles with the uniform weight period/number."""
        period = np.pi if half_circle else 2 * np.pi
        result = Angles.sparse(number, half_circle=half_circle)
        result.weights = result.weights * (period / number)
        return result

    @staticmethod
    def uniform_interval(start: float, end: float, number: int, half_circle: bool = False) -> Angles:
        """``number`` midpoints of a uniform partition of [start, end]."""
        half_step = abs(end - start) / (2 * number)
        grid = np.linspace(start + half_step, end - half_step, number)
        return Angles(grid, 2 * half_step * np.ones_like(grid), half_circle=half_circle)

    @staticmethod
    def intervals(number_list, start_list, end_list, half_circle: bool = False) -> Angles:
        """Concatenation of :meth:`uniform_interval` samplings."""
        if not len(number_list) == len(start_list) == len(end_list):
            raise ValueError("All input lists must have the same length.")
        all_angles: list[float] = []
        all_weights: list[float] = []
        for n, lo, hi in zip(number_list, start_list, end_list):
            part = Angles.uniform_interval(lo, hi, n, half_circle=half_circle)
            all_angles.extend(part.angles)
            all_weights.extend(part.weights)
        return Angles(np.array(all_angles), np.array(all_weights), half_circle=half_circle)

    @staticmethod
    def from_list(angles: npt.ArrayLike, half_circle: bool = False) -> Angles:
        """Explicit angles with natural (neighbour-distance) weights."""
        angles = np.asarray(angles)
        period = np.pi if half_circle else 2 * np.pi
        order = np.argsort(angles % period)
        srt = angles[order] % period
        ext = np.hstack([-period + srt[-1], srt, srt[0] + period])
        weights_sorted = _natural_weights(srt, ext)
        weights = np.empty_like(weights_sorted)
        weights[order] = weights_sorted
        return Angles(angles, weights, half_circle=half_circle)


@dataclass(init=False)
class Detectors:
    """Detector line: pixel count, physical extent, centre shift, orientation.

    A negative ``number`` means ``reversed=True`` with ``abs(number)`` pixels
    unless ``reversed`` is given explicitly (utilities.py:291-305).
    """

    number: int
    extent: float | ExtentPlaceholder = ExtentPlaceholder.FULL
    center: float = 0.0
    reversed: bool = False

    def __init__(self, number: int, extent=ExtentPlaceholder.FULL, center: float = 0.0, reversed: bool | None = None):
        self.number = abs(number)
        self.extent = extent
        self.center = center
        self.reversed = (number < 0) if reversed is None else reversed


@dataclass(init=False)
class ImageDomain:
    """Image grid size (Nx, Ny), physical extent of the longer side, centre shift."""

    size: tuple[int, int]
    extent: float | ExtentPlaceholder = 2.0
    center: tuple[float, float] = (0.0, 0.0)

    def __init__(self, size, extent=2.0, center=(0.0, 0.0)):
        if isinstance(size, (int, np.integer)):
            size = (int(size), int(size))
        self.size = tuple(size)
        self.extent = extent
        self.center = center


# ---------------------------------------------------------------------------
# Half-circle extent solvers (parallel beam, phi in [0, pi]).
# The sinogram coordinate of a point (x, y) is s(phi) = x sin(phi) - y cos(phi).
# ---------------------------------------------------------------------------


def _solve_quadratic(a: float, b: float, d: float) -> tuple[float, float]:
    """Roots (x1 <= x2) of a x^2 + b x + d = 0, or (nan, nan) when complex."""
    disc = b * b - 4 * a * d
    if disc < 0:
        return (float("nan"), float("nan"))
    root = np.sqrt(disc)
    return ((-b - root) / (2 * a), (-b + root) / (2 * a))


def _point_image_fits_detector_halfcircle(M, Dd) -> bool:
    """Whether a single point at (Mx, My) projects inside the detector for all
    phi in [0, pi]: its trajectory spans [-|My|, r] (Mx >= 0) or [-r, |My|]."""
    (Md, Mx, My) = M
    r = np.sqrt(Mx**2 + My**2)
    sgn = 1.0 if Mx >= 0 else -1.0
    return r <= sgn * Md + Dd / 2 and abs(My) <= Dd / 2 - sgn * Md


def full_detector_given_image_halfcircle(M, D) -> float:
    """Smallest detector width such that every ray through the image (size D,
    centre (Mx, My)) hits the detector centred at Md (utilities.py:408-447)."""
    (Md, Mx, My) = M
    (Dx, Dy) = D
    # rotate to the a cos + b sin convention of the max/min formulas
    (Mx, My) = (-My, Mx)
    (Dx, Dy) = (Dy, Dx)
    if My >= -Dy / 2:
        top = np.sqrt((abs(Mx) + Dx / 2) ** 2 + (My + Dy / 2) ** 2)
    else:
        top = abs(Mx) + Dx / 2
    if My <= Dy / 2:
        bottom = -np.sqrt((abs(Mx) + Dx / 2) ** 2 + (My - Dy / 2) ** 2)
    else:
        bottom = -abs(Mx) - Dx / 2
    return 2 * max(Md - bottom, top - Md)


def full_image_given_detector_halfcircle(M, Dd: float, c: float = 1.0):
    """Largest image (aspect Dx/Dy = c) such that every ray through it hits the
    detector of width Dd (utilities.py:450-513); None when impossible."""
    (Md, Mx, My) = M
    if not _point_image_fits_detector_halfcircle(M, Dd):
        return None
    quad = (1 + c**2) / (4 * c**2)
    (x1, x2) = _solve_quadratic(quad, abs(My) / c + Mx, -((Md + Dd / 2) ** 2) + Mx**2 + My**2)
    (y1, y2) = _solve_quadratic(quad, abs(My) / c - Mx, -((Md - Dd / 2) ** 2) + Mx**2 + My**2)

    Dx = None
    if (y2 <= x2) and (x1 <= y2):
        Dx = y2
    elif (x2 <= y2) and (y1 <= x2):
        Dx = x2
    if Dx is not None and Dx >= 2 * abs(Mx):
        return (Dx, Dx / c)

    def _linear(lo, hi, bound):
        if hi <= bound:
            return hi
        if lo <= bound:
            return bound
        return None

    if Mx < 0:
        Dx = _linear(y1, y2, (Md + Dd / 2 - abs(My)) * 2 * c)
    elif Mx > 0:
        Dx = _linear(x1, x2, (Dd / 2 - Md - abs(My)) * 2 * c)
    else:
        return None
    if Dx is None or Dx < 0:
        return None
    return (Dx, Dx / c)


def valid_image_given_detector_halfcircle(M, Dd: float, c: float = 1.0):
    """Smallest image (aspect c) such that every ray hitting the detector passes
    through it (utilities.py:516-546)."""
    (Md, Mx, My) = M
    Dx = 2 * max(Mx - Md + Dd / 2, -Mx + Md + Dd / 2)
    Dy = 2 * max(abs(Md) + Dd / 2 - My, abs(Md) + Dd / 2 + My)
    Dx = max(Dx, c * Dy)
    return (Dx, Dx / c)


def valid_detector_given_image_halfcircle(M, D):
    """Largest detector such that every ray hitting it passes through the image
    (utilities.py:549-580); None when no positive width exists."""
    (Md, Mx, My) = M
    (Dx, Dy) = D
    Dd = 2 * min(-Mx + Dx / 2 + Md, Mx - Md + Dx / 2, My - abs(Md) + Dy / 2, -My - abs(Md) + Dy / 2)
    if Dd <= 0:
        return None
    return Dd
