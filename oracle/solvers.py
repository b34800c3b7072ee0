"""The reference's solver loops on the CPU oracle kernels -- TEST ONLY.

Each function restates one loop of gratopy/gratopy.py with numpy arrays and
the C oracle kernels; ``geom`` is an :class:`OracleGeometry` wrapping the
tables of one geometry.  Reductions are numpy float64 dot products (the
reference's pyopencl reduction order is not reproducible, see
gratopy.py:2372, 2527-2543).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import oracle


@dataclass
class OracleGeometry:
    is_fan: bool
    img_shape: tuple
    sinogram_shape: tuple
    struct: dict
    delta_x: float
    delta_s: float

    @classmethod
    def from_settings(cls, ps):
        """Borrow the host tables of a ProjectionSettings (numpy, per dtype)."""
        return cls(ps.is_fan, tuple(ps.img_shape), tuple(ps.sinogram_shape), ps.struct,
                   ps.delta_x, ps.delta_s)

    def _tabs(self, dtype):
        dt = np.dtype(dtype)
        s = self.struct
        return s["ofs_dict"][dt], s.get("sdpd_dict", {}).get(dt), s["geo_dict"][dt]

    def fwd(self, img, order="F"):
        ofs, sdpd, geo = self._tabs(img.dtype)
        if self.is_fan:
            return oracle.fan_fwd(img, ofs, sdpd, geo, *self.sinogram_shape, sino_order=order)
        return oracle.radon_fwd(img, ofs, geo, *self.sinogram_shape, sino_order=order)

    def back(self, sino, order="F"):
        ofs, sdpd, geo = self._tabs(sino.dtype)
        if self.is_fan:
            return oracle.fan_back(sino, ofs, sdpd, geo, self.img_shape, img_order=order)
        return oracle.radon_back(sino, ofs, geo, self.img_shape, img_order=order)

    def angle_weights(self, dtype):
        return self.struct["angle_diff_dict"][np.dtype(dtype)]


def _order(a):
    return "C" if a.flags.c_contiguous else "F"


def normest(geom, number_iterations=50, dtype="float32", start=None):
    """gratopy.py:2329-2377 with an optional explicit start vector."""
    if start is None:
        start = np.random.randn(*geom.img_shape)
    img = np.require(start, dtype, "F")
    normsqr = 1.0
    for _ in range(number_iterations):
        normsqr = float(np.sum(img, dtype=np.float64))
        img = np.require(img / img.dtype.type(normsqr), dtype, "F")
        img = geom.back(geom.fwd(img))
    return np.sqrt(normsqr)


def landweber(sino, geom, number_iterations=100, w=1.0, norm_estimate=None):
    """gratopy.py:2380-2446."""
    if norm_estimate is None:
        norm_estimate = normest(geom)
    order = _order(sino)
    w = sino.dtype.type(w / norm_estimate**2)
    z = (sino.shape[2],) if sino.ndim > 2 else ()
    unew = np.zeros(geom.img_shape + z, sino.dtype, order=order)
    for _ in range(number_iterations):
        sinonew = geom.fwd(unew, order) - sino
        sinonew = np.require(sinonew, sino.dtype, order)
        u = geom.back(sinonew, order)
        oracle.equ_mul_add(unew, -w, np.require(u, sino.dtype, order))
    return unew


def conjugate_gradients(sino, geom, number_iterations=20, epsilon=0.0, x0=None):
    """gratopy.py:2449-2551 (alpha, beta in float64)."""
    x = None
    for _, x in _cg_iterates(sino, geom, number_iterations, epsilon, x0):
        pass
    return x


def conjugate_gradients_iterates(sino, geom, iterations):
    """{n: x after n CG iterations} for every n in `iterations` (one run)."""
    want = set(iterations)
    return {n: x.copy(order="K") for n, x in _cg_iterates(sino, geom, max(want)) if n in want}


def _cg_iterates(sino, geom, number_iterations, epsilon=0.0, x0=None):
    """Yields (n, x) after each iteration n = 1, 2, ... of gratopy.py:2449-2551;
    with 0 iterations, (0, x0)."""
    order = _order(sino)
    z = (sino.shape[2],) if sino.ndim > 2 else ()
    dt = sino.dtype
    x = np.zeros(geom.img_shape + z, dt, order=order) if x0 is None else np.array(x0, order=order)
    d = np.require(sino - geom.fwd(x, order), dt, order)
    p = geom.back(d, order)
    snew = geom.back(d, order)
    sold = snew.copy(order="K")
    w = geom.angle_weights(dt)
    vdot = lambda a, b: float(np.dot(a.ravel(order="K").astype(np.float64), b.ravel(order="K").astype(np.float64)))
    sino_norm = vdot(sino, sino)
    for it in range(number_iterations):
        q = geom.fwd(p, order)
        qw = oracle.weight_sino(q, w)
        alpha = dt.type(geom.delta_x**2 / geom.delta_s * (vdot(sold, sold) / vdot(qw, q)))
        oracle.equ_mul_add(x, alpha, p)
        oracle.equ_mul_add(d, -alpha, q)
        snew = np.require(geom.back(d, order), dt, order)
        beta = vdot(snew, snew) / vdot(sold, sold)
        sold, snew = snew, sold
        oracle.mul_add_add(p, beta, p.copy(order="K"), sold)
        yield it + 1, x
        residue = np.sqrt(vdot(d, d) / sino_norm)
        if residue < epsilon:
            break
    if number_iterations == 0:
        yield 0, x


def total_variation(sino, geom, mu, number_iterations=1000, slice_thickness=1.0,
                    stepsize_weighting=10.0, norm_estimate=None):
    """gratopy.py:2554-2762 (3-D input coupled through slice_thickness)."""
    dt = sino.dtype
    order = _order(sino)
    img_shape = geom.img_shape
    if sino.ndim == 2:
        slice_thickness = 0
    else:
        img_shape = img_shape + (sino.shape[2],)
    U = np.zeros(img_shape, dt, order=order)
    U_ = np.zeros(img_shape, dt, order=order)
    V = np.zeros(4 * U.size, dt)
    Lamb = np.zeros(sino.shape, dt, order=order)
    KSTAR = np.zeros(img_shape, dt, order=order)
    if norm_estimate is None:
        norm_estimate = normest(geom)
    sigma = 1.0 / np.sqrt(13) * stepsize_weighting
    tau = 1.0 / np.sqrt(13) / stepsize_weighting
    mu = mu * norm_estimate**2 * geom.delta_x
    mu = mu / (sigma + mu)
    for _ in range(number_iterations):
        oracle.tv_update_v(V, U_, sigma, slice_thickness)
        oracle.tv_update_normv(V, img_shape, order)
        KU = np.require(geom.fwd(U_, order), dt, order)
        oracle.tv_update_lambda(Lamb, KU, sino, sigma / norm_estimate, mu)
        KSTAR = np.require(geom.back(Lamb, order), dt, order)
        oracle.tv_update_u(U_, U, V, KSTAR, tau, 1.0 / norm_estimate, slice_thickness)
        oracle.tv_update_extra(U_, U)
        U, U_ = U_, U
    return U
