"""The reference test suite's control numbers and invariants that the other
GPU tests do not cover, through the drop-in API on the HIP path (rtol 1e-3,
the reference's tolerance; tests/helpers.py regenerates its control
stream):

* Radon limited angles: projection with incorrect angle weights and the
  backprojections with incorrect and repaired weights (tests/test_radon.py:427-623);
* Radon midpoint shift (tests/test_radon.py:799-884);
* fanbeam limited angles in two sections (tests/test_fanbeam.py:452-675);
* fanbeam non-square image (tests/test_fanbeam.py:1385-1475);
* the 24 orientation numbers, detector and midpoint shifts, parallel and
  fanbeam, C-ordered images (tests/test_fanbeam.py:769-943);
* the sparse system matrix's products (tests/test_radon.py:715-796,
  tests/test_fanbeam.py:1491-1579);
* mass preservation (tests/test_radon.py:273-337, tests/test_fanbeam.py:280-355).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import gratopy_amd as gratopy
from gratopy_amd import _device as dev
from tests.helpers import assert_control, create_phantoms, phantom

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")


def _ps(geometry, img_shape, angles, n_det, **kw):
    return gratopy.ProjectionSettings("cuda:0", geometry, img_shape, angles, n_det, **kw)


def _fwd(ps, img, order="F"):
    out = dev.to_host(gratopy.forwardprojection(dev.to_device(np.require(img, None, order), "cuda:0",
                                                              order=order), ps))
    torch.cuda.synchronize()
    return out


def _back(ps, sino, order="F"):
    out = dev.to_host(gratopy.backprojection(dev.to_device(np.require(sino, None, order), "cuda:0",
                                                           order=order), ps))
    torch.cuda.synchronize()
    return out


def test_radon_limited_angles_weights():
    print("radon limited angles", flush=True)
    N, Ns, na = 1200, int(0.3 * 1200), 180
    img = create_phantoms(N, F32)
    delta = np.pi * 3 / 4 / (na - 1)
    incorrect = np.linspace(0, np.pi * 3 / 4.0, na) + np.pi / 8
    correct = (incorrect, np.pi / 8 - delta / 2, np.pi * 7 / 8 + delta / 2)
    ps_c = _ps(gratopy.RADON, img.shape, correct, Ns, detector_width=2.0)
    ps_i = _ps(gratopy.RADON, img.shape, incorrect, Ns, detector_width=2.0)
    ps_w = _ps(gratopy.RADON, img.shape, incorrect, Ns, detector_width=2.0, angle_weights=delta)
    dims = (N, N, Ns, 3, 2)
    sino_c = _fwd(ps_c, img)
    assert_control(sino_c, dims, 340.93, "sino", "sinogram (correct weights)")
    assert_control(_fwd(ps_i, img), dims, 340.93, "sino", "sinogram (incorrect weights)")
    assert_control(_fwd(ps_w, img), dims, 340.93, "sino", "sinogram (scalar weights)")
    assert_control(_back(ps_i, sino_c), dims, 3453.203, "img", "backprojection (incorrect weights)")
    ps_i.set_angle_weights(ps_c.angle_weights)
    assert_control(_back(ps_i, sino_c), dims, 3385.137, "img", "backprojection (repaired weights)")


def test_radon_midpoint_shift():
    print("radon midpoint shift", flush=True)
    N = 1200
    img = create_phantoms(N, F32)
    ps = _ps(gratopy.RADON, img.shape, 360, 600, image_width=3.0, detector_width=2.0,
             midpoint_shift=[0.0, 0.4])
    dims = (N, N, 600, 360, 2)
    assert_control(img, dims, 2949.0, "img", "image")
    sino = _fwd(ps, img)
    assert_control(sino, dims, 2083.80, "sino", "shifted sinogram")
    assert_control(_back(ps, sino), dims, 6269.703, "img", "shifted backprojection")


def test_fan_limited_angles_two_sections():
    print("fan limited angles", flush=True)
    N = 1200
    img = create_phantoms(N, F32)
    Ns = int(0.3 * N)
    n1, a1, b1 = 180, np.pi / 8.0, np.pi / 8 + np.pi * 3 / 4.0
    angles1 = np.linspace(a1, b1, n1)
    d1 = (b1 - a1) / (n1 - 1) * 0.5
    n2, a2, b2 = 90, 5.0 / 4 * np.pi, 2 * np.pi
    d2 = (b2 - a2) / (n2 - 1) * 0.5
    angles2 = np.linspace(a2, b2, n2)[:-1]
    correct = [(angles1, a1 - d1, b1 + d1), (angles2, a2 - d2, b2 - d2)]
    incorrect = list(angles1) + list(angles2)
    geo = dict(image_width=2.0, R=5.0, RE=2.0, detector_width=6.0)
    ps_c = _ps(gratopy.FANBEAM, img.shape, correct, Ns, **geo)
    ps_i = _ps(gratopy.FANBEAM, img.shape, incorrect, Ns, **geo)
    dims = (N, N, Ns, 2, 2)
    sino_c = _fwd(ps_c, img)
    assert_control(sino_c, dims, -131.00, "sino", "limited-angle sinogram")
    assert_control(_back(ps_c, sino_c), dims, 19495.96, "img", "backprojection (correct weights)")
    assert_control(_back(ps_i, sino_c), dims, 23043.191, "img", "backprojection (incorrect weights)")


def test_fan_nonquadratic():
    print("fan non-square", flush=True)
    N = 1200
    full = create_phantoms(N, F32)
    N2 = int(N * 2 / 3.0)
    img = np.asfortranarray(full[:, 0:N2, :])
    ps = _ps(gratopy.FANBEAM, img.shape, 360, 600, R=5.0, RE=3.0, detector_width=5.0)
    dims = (N, N2, 600, 360, 2)
    sino = _fwd(ps, img)
    assert_control(sino, dims, 530.45, "sino", "non-square fan sinogram")
    assert_control(_back(ps, sino), dims, 16101.3542, "img", "non-square fan backprojection")


GEOMETRIC_ORIENTATION = [
    ((0, 0, 0), (-22.6691, 1760.90, 131.301, 1157.280)),
    ((50, 0, 0), (51.959, -61.3725, 85.7716, 1058.5918)),
    ((0, 20, 0), (-703.53, 881.399, 542.11, 1050.41)),
    ((0, -20, 0), (390.100, 173.108, 110.7186, 1202.997)),
    ((0, 0, 30), (935.812, 1775.13, 578.73, 951.169)),
    ((0, 0, -30), (51.3383, 1225.368, -322.945, 579.33)),
]


@pytest.mark.parametrize("params,expected", GEOMETRIC_ORIENTATION)
def test_geometric_orientation(params, expected):
    print(f"orientation {params}", flush=True)
    Nx, Ns, angles, width = 600, 300, 360, 94.39
    img1 = np.zeros((Nx, Nx), np.float32)
    img1[0 : int(Nx * 0.3)][:, 0 : int(Nx * 0.3)] = 1
    img2 = img1.copy()
    img2[0 : int(Nx * 0.3)][:, 0:Nx] = 1
    dshift, mshift = params[0], [params[1], params[2]]
    fan = _ps(gratopy.FANBEAM, img1.shape, angles, Ns, detector_width=170.0, RE=200.0, R=350,
              detector_shift=dshift, image_width=width, midpoint_shift=mshift)
    par = _ps(gratopy.RADON, img1.shape, angles, Ns, image_width=width, detector_width=width,
              detector_shift=dshift, midpoint_shift=mshift)
    dims = (Nx, Nx, Ns, angles, 1)
    assert_control(_fwd(par, img1, "C"), dims, expected[0], "sino", "parallel, image 1")
    assert_control(_fwd(par, img2, "C"), dims, expected[1], "sino", "parallel, image 2")
    assert_control(_fwd(fan, img1, "C"), dims, expected[2], "sino", "fanbeam, image 1")
    assert_control(_fwd(fan, img2, "C"), dims, expected[3], "sino", "fanbeam, image 2")


@pytest.mark.parametrize("geometry,expected", [
    ("radon", (7.1182, 6.5315, 1.03955)),
    ("fan", (7.1182, 233.96, 2690.00)),
])
@pytest.mark.parametrize("dtype", [F32, F64])
def test_sparse_matrix_control_numbers(geometry, expected, dtype):
    print(f"sparse matrix {geometry} {dtype}", flush=True)
    Nx, nd, angles = 150, 100, 30
    if geometry == "radon":
        ps = _ps(gratopy.RADON, (Nx, Nx), angles, nd)
    else:
        ps = _ps(gratopy.FANBEAM, (Nx, Nx), angles, nd, detector_width=400.0, R=752.0, RE=200.0)
    M = ps.create_sparse_matrix(dtype=dtype, order="F").tocsr()
    img = phantom(Nx, dtype).reshape(Nx * Nx, order="F")
    sino = M @ img
    back = M.T @ sino
    dims = (Nx, Nx, nd, angles, 1)
    assert_control(img.reshape(Nx, Nx, order="F"), dims, expected[0], "img", "image")
    assert_control(sino.reshape(nd, angles, order="F"), dims, expected[1], "sino", "sparse sinogram")
    assert_control(back.reshape(Nx, Nx, order="F"), dims, expected[2], "img", "sparse backprojection")


def test_radon_mass_preservation():
    """The mass of a square is carried into every parallel projection."""
    print("radon mass", flush=True)
    N = 300
    ps = _ps(gratopy.RADON, (N, N), 30, 170, detector_width=4.0, image_width=4.0)
    for dtype in (F32, F64):
        img = np.zeros((N, N), dtype)
        img[N // 3 : 2 * N // 3, N // 3 : 2 * N // 3] = 1
        sino = _fwd(ps, img)
        mass = img.sum() * ps.delta_x**2
        for a in range(30):
            assert abs(1 - mass / (sino[:, a].sum() * ps.delta_s)) < 1e-3, (dtype, a)


def test_fan_mass_preservation():
    """Fanbeam projection mass is about R/RE times the image mass."""
    print("fan mass", flush=True)
    N = 100
    r = 1 / 40.0 * np.sqrt(2)
    for nd in (50, 200, 800):
        ps = _ps(gratopy.FANBEAM, (N, N), 72, nd, detector_width=400.0 * r, R=1200.0 * r,
                 RE=200.0 * r, image_width=40.0 * r)
        sino = _fwd(ps, np.ones((N, N), F32))
        ratio = sino.sum() * ps.delta_s / 72 / (N * N * ps.delta_x**2)
        assert abs(ratio / (ps.R / ps.RE) - 1) < 0.1, (nd, ratio)
