"""Shared test helpers: the reference's control-number check and test images.

Control numbers (tests/helpers.py:13-111 of the reference): 1000 fixed
pseudo-random entries of an array, weighted by fixed N(0,1) factors and
summed; the reference's tests pin their results to rtol 1e-3.  The index and
factor stream is regenerated from its recipe (default_rng(1), draws in the
order s, phi, z, factors, x, y, each of length 1000, integers in [0, 2000)),
which reproduces the reference's tests/rng.txt exactly.
"""

from __future__ import annotations

import functools

import numpy as np

from gratopy_amd.phantom import ct_modified_shepp_logan_params_2d, rasterize_ellipses


@functools.lru_cache(maxsize=1)
def _control_stream():
    rng = np.random.default_rng(1)
    m, M = 1000, 2000
    s = rng.integers(0, M, m)
    phi = rng.integers(0, M, m)
    z = rng.integers(0, M, m)
    factors = rng.normal(0, 1, m)
    x = rng.integers(0, M, m)
    y = rng.integers(0, M, m)
    return s, phi, z, factors, x, y


def control_number(data, dims, classified):
    """Sum of factors[i] * data[i-th control index] (float64 accumulation)."""
    Nx, Ny, Ns, Na, Nz = dims
    s, phi, z, factors, x, y = _control_stream()
    data = np.asarray(data)
    if data.ndim == 2:
        data = data.reshape(data.shape[0], data.shape[1], 1)
    if classified == "img":
        i1, i2 = x % Nx, y % Ny
    else:
        i1, i2 = s % Ns, phi % Na
    return float(np.sum(factors * data[i1, i2, z % Nz].astype(np.float64)))


def assert_control(data, dims, expected, classified, name, rtol=1e-3):
    got = control_number(data, dims, classified)
    assert abs(got - expected) < abs(expected) * rtol, (
        f"control number of {name}: expected {expected}, got {got}"
    )


def phantom(N, dtype="float64"):
    """The reference's gratopy.phantom(queue, N, dtype) on the host: modified
    Shepp-Logan, (M, N) C-ordered."""
    if np.isscalar(N):
        M, N = N, N
    else:
        M, N = N
    return rasterize_ellipses(M, N, ct_modified_shepp_logan_params_2d()).astype(dtype)


def create_phantoms(N, dtype="float64", order="F"):
    """The reference's two-slice test image (tests/helpers.py:114-134):
    Shepp-Logan scaled to max 255 and a two-bar image, stacked on axis 2."""
    dtype = np.dtype(dtype)
    A = phantom(N, dtype)
    A *= dtype.type(255) / A.max()
    B = np.empty(A.shape, dtype)
    B[:] = 255 - 120
    B[int(N / 3) : int(2 * N / 3)] = 0
    B[0 : int(N / 4)] = 0
    B[int(N - N / 4) : N] = 0
    return np.require(np.stack([A, B], axis=-1), dtype, order)


def rel_err(got, ref):
    """Normwise max|got - ref| / max|ref|."""
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.abs(np.asarray(got, dtype=np.float64) - ref).max() / max(np.abs(ref).max(), 1e-300))


def record_parity(config, quantity, **values):
    """Print one measured set of parity numbers as a line 'parity {json}'
    (run pytest with -s; profiles/parity_configs.txt collects these lines
    from a run's output)."""
    import json

    line = json.dumps({"config": config, "quantity": quantity, **values}, default=float)
    print("parity", line, flush=True)
