"""Pins of the CPU oracle (CPU only, no GPU).

Every oracle kernel is checked against something it does not share code
with:

* the dense closed form of oracle/closed_form.py (geometry-derived,
  pixel-driven) on small grids, float64 to 1e-10 and float32 to 1e-5, across
  non-square images, shifts, reversed detectors, limited angles, 3-D stacks
  and both memory orders -- pins radon_fwd/back, fan_fwd/back;
* the reference test-suite's control numbers (rtol 1e-3), computed with the
  reference's own rng recipe and test images -- pins the same kernels plus
  the table builders end to end;
* invariants: adjointness in the weighted inner products, mass
  preservation, the sparse matrix (single-line kernels) reproducing the
  forward projection;
* closed forms for the pointwise kernels (weighting, axpy, TV updates) and a
  brute-force numpy evaluation of the TV updates.
"""

from __future__ import annotations

import numpy as np
import pytest

import gratopy_amd as gratopy
import oracle
from oracle import closed_form as cf
from tests.helpers import assert_control, create_phantoms, phantom, rel_err

F32, F64 = np.dtype("float32"), np.dtype("float64")


def _ps(geometry, img_shape, angles, n_det, **kw):
    return gratopy.ProjectionSettings(None, geometry, img_shape, angles, n_det, **kw)


def _oracle_fwd(ps, img, dtype, order="F"):
    t = ps.struct
    if ps.is_fan:
        return oracle.fan_fwd(np.require(img, dtype, order), t["ofs_dict"][dtype], t["sdpd_dict"][dtype],
                              t["geo_dict"][dtype], *ps.sinogram_shape, sino_order=order)
    return oracle.radon_fwd(np.require(img, dtype, order), t["ofs_dict"][dtype], t["geo_dict"][dtype],
                            *ps.sinogram_shape, sino_order=order)


def _oracle_back(ps, sino, dtype, order="F"):
    t = ps.struct
    if ps.is_fan:
        return oracle.fan_back(np.require(sino, dtype, order), t["ofs_dict"][dtype], t["sdpd_dict"][dtype],
                               t["geo_dict"][dtype], ps.img_shape, img_order=order)
    return oracle.radon_back(np.require(sino, dtype, order), t["ofs_dict"][dtype], t["geo_dict"][dtype],
                             ps.img_shape, img_order=order)


RADON_CASES = [
    dict(img_shape=(24, 24), angles=13, n_det=30),
    dict(img_shape=(20, 31), angles=17, n_det=25, detector_width=2.5, image_width=2.0),
    dict(img_shape=(26, 18), angles=11, n_det=19, midpoint_shift=(0.2, -0.1), detector_shift=0.15),
    dict(img_shape=(16, 16), angles=[(7, 0.3, 1.9)], n_det=21, detector_width=1.3),
]

FAN_CASES = [
    dict(img_shape=(24, 24), angles=19, n_det=30, R=10.0, RE=5.0, detector_width=6.0),
    dict(img_shape=(22, 15), angles=23, n_det=27, R=7.0, RE=3.0, detector_width=5.0, image_width=2.0),
    dict(img_shape=(20, 20), angles=14, n_det=25, R=9.0, RE=4.0, detector_width=6.0,
         midpoint_shift=(0.3, -0.2), detector_shift=0.4, image_width=2.0),
    dict(img_shape=(18, 18), angles=-12, n_det=22, R=8.0, RE=4.0, detector_width=7.0,
         reverse_detector=True, image_width=2.5),
]


def _radon_cf(ps, kw):
    A = cf.radon_matrix(ps.img_shape, ps.angles, ps.n_detectors, ps.detector_width, ps.image_width,
                        kw.get("midpoint_shift", (0.0, 0.0)), kw.get("detector_shift", 0.0))
    scale = cf.radon_scale(ps.img_shape, ps.n_detectors, ps.detector_width, ps.image_width)
    return A, scale


@pytest.mark.parametrize("case", range(len(RADON_CASES)))
def test_radon_matches_closed_form(case):
    kw = dict(RADON_CASES[case])
    shape, angles, nd = kw.pop("img_shape"), kw.pop("angles"), kw.pop("n_det")
    ps = _ps(gratopy.RADON, shape, angles, nd, **kw)
    rng = np.random.default_rng(case)
    img = rng.random(shape)
    sino = rng.random(ps.sinogram_shape)
    A, scale = _radon_cf(ps, RADON_CASES[case])
    ref_f = cf.radon_fwd(img, A, scale)
    ref_b = cf.radon_back(sino, A, ps.angle_weights)
    for dtype, tol in ((F64, 1e-10), (F32, 1e-5)):
        for order in ("F", "C"):
            assert rel_err(_oracle_fwd(ps, img, dtype, order), ref_f) < tol
            assert rel_err(_oracle_back(ps, sino, dtype, order), ref_b) < tol


@pytest.mark.parametrize("case", range(len(FAN_CASES)))
def test_fan_matches_closed_form(case):
    kw = dict(FAN_CASES[case])
    shape, angles, nd = kw.pop("img_shape"), kw.pop("angles"), kw.pop("n_det")
    ps = _ps(gratopy.FANBEAM, shape, angles, nd, **kw)
    A, sdpd, Rp, dxi, dx = cf.fan_matrix(
        ps.img_shape, ps.angles, ps.n_detectors, ps.detector_width, ps.R, ps.RE, ps.image_width,
        kw.get("midpoint_shift", (0.0, 0.0)), kw.get("detector_shift", 0.0), kw.get("reverse_detector", False))
    rng = np.random.default_rng(10 + case)
    img = rng.random(shape)
    sino = rng.random(ps.sinogram_shape)
    ref_f = cf.fan_fwd(img, A, sdpd, Rp, dxi, dx)
    ref_b = cf.fan_back(sino, A, sdpd, Rp, ps.angle_weights)
    for dtype, tol in ((F64, 1e-10), (F32, 1e-5)):
        for order in ("F", "C"):
            assert rel_err(_oracle_fwd(ps, img, dtype, order), ref_f) < tol
            assert rel_err(_oracle_back(ps, sino, dtype, order), ref_b) < tol


def test_closed_form_catches_planted_mistakes():
    """A plausible slip in the oracle's inputs (sign of the detector shift,
    transposed image) must fail the closed-form comparison."""
    kw = dict(detector_shift=0.15, midpoint_shift=(0.2, -0.1))
    ps = _ps(gratopy.RADON, (26, 18), 11, 19, **kw)
    A, scale = _radon_cf(ps, kw)
    img = np.random.default_rng(3).random((26, 18))
    ref = cf.radon_fwd(img, A, scale)
    assert rel_err(_oracle_fwd(ps, img, F64), ref) < 1e-10
    bad = _ps(gratopy.RADON, (26, 18), 11, 19, detector_shift=-0.15, midpoint_shift=(0.2, -0.1))
    assert rel_err(_oracle_fwd(bad, img, F64), ref) > 1e-2
    sq = _ps(gratopy.RADON, (20, 20), 11, 19)
    img2 = np.random.default_rng(4).random((20, 20))
    A2, s2 = _radon_cf(sq, {})
    assert rel_err(_oracle_fwd(sq, img2.T.copy(), F64), cf.radon_fwd(img2, A2, s2)) > 1e-2


def test_3d_stack_is_slicewise():
    ps = _ps(gratopy.FANBEAM, (20, 20), 15, 24, R=10.0, RE=5.0, detector_width=6.0)
    img = np.random.default_rng(5).random((20, 20, 3))
    for order in ("F", "C"):
        s3 = _oracle_fwd(ps, img, F64, order)
        b3 = _oracle_back(ps, s3, F64, order)
        for z in range(3):
            assert rel_err(s3[:, :, z], _oracle_fwd(ps, img[:, :, z], F64)) < 1e-14
            assert rel_err(b3[:, :, z], _oracle_back(ps, s3[:, :, z].copy(), F64)) < 1e-14


@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_oracle_adjointness(geometry):
    """<R x, y>_sino = <x, R^# y>_img with <.,.>_img = dx^2 sum and
    <.,.>_sino = ds sum_a w_a sum_s (getting_started.rst:91-131)."""
    if geometry == "radon":
        ps = _ps(gratopy.RADON, (60, 60), 45, 50)
    else:
        ps = _ps(gratopy.FANBEAM, (60, 60), 60, 50, detector_width=83.0, R=900.0, RE=300.0)
    rng = np.random.default_rng(7)
    for dtype, tol in ((F64, 1e-12), (F32, 1e-5)):
        for _ in range(5):
            x = rng.random(ps.img_shape).astype(dtype)
            y = rng.random(ps.sinogram_shape).astype(dtype)
            Rx = _oracle_fwd(ps, x, dtype)
            Ry = _oracle_back(ps, y, dtype)
            lhs = np.sum(Rx.astype(np.float64) * y * ps.angle_weights[None, :]) * ps.delta_s
            rhs = np.sum(x.astype(np.float64) * Ry) * ps.delta_x**2
            if geometry == "radon":
                # the forward kernel scales by float(dx)^2/float(ds) even in
                # double precision (radon.cl:82-83): adjoint up to that rounding
                lhs *= (ps.delta_x**2 / ps.delta_s) / cf.radon_scale(
                    ps.img_shape, ps.n_detectors, ps.detector_width, ps.image_width)
            assert abs(lhs - rhs) / abs(rhs) < tol


def test_oracle_mass_preservation_radon():
    """tests/test_radon.py:273-337: the mass of a square is carried into every projection."""
    N = 300
    ps = _ps(gratopy.RADON, (N, N), 30, 170, detector_width=4.0, image_width=4.0)
    img = np.zeros((N, N))
    img[N // 3 : 2 * N // 3, N // 3 : 2 * N // 3] = 1
    sino = _oracle_fwd(ps, img, F64)
    mass = img.sum() * ps.delta_x**2
    for a in range(30):
        assert abs(1 - mass / (sino[:, a].sum() * ps.delta_s)) < 1e-3


def test_oracle_mass_preservation_fan():
    """tests/test_fanbeam.py:280-355: projection mass is about R/RE times the image mass."""
    N = 100
    r = 1 / 40.0 * np.sqrt(2)
    for nd in (50, 200, 800):
        ps = _ps(gratopy.FANBEAM, (N, N), 72, nd, detector_width=400.0 * r, R=1200.0 * r,
                 RE=200.0 * r, image_width=40.0 * r)
        sino = _oracle_fwd(ps, np.ones((N, N)), F32)
        ratio = sino.sum() * ps.delta_s / 72 / (N * N * ps.delta_x**2)
        assert abs(ratio / (ps.R / ps.RE) - 1) < 0.1


# ---- control numbers of the reference test suite --------------------------------


def test_control_radon_projection():
    """tests/test_radon.py:126-148 (1200^2 two-slice phantom, 360 angles, 600 det)."""
    N, Ns, angles = 1200, 600, 360
    img = create_phantoms(N, F32)
    ps = _ps(gratopy.RADON, img.shape, angles, Ns, image_width=4.0, detector_width=4.0)
    dims = (N, N, Ns, angles, 2)
    assert_control(img, dims, 2949.3738, "img", "image")
    sino = _oracle_fwd(ps, img, F32)
    assert_control(sino, dims, 4737.367, "sino", "sinogram")
    back = _oracle_back(ps, sino, F32)
    assert_control(back, dims, 7427.70, "img", "backprojection")


@pytest.mark.parametrize("dtype", [F32, F64])
def test_control_radon_types_contiguity(dtype):
    """tests/test_radon.py:164-270 (600^2 phantom, 180 angles, 300 det, 4 orders)."""
    Nx, nd, angles = 600, 300, 180
    img = phantom(Nx, dtype)
    ps = _ps(gratopy.RADON, (Nx, Nx), angles, nd)
    dims = (Nx, Nx, nd, angles, 1)
    assert_control(img, dims, 7.89220, "img", "image")
    for o1 in ("F", "C"):
        for o2 in ("F", "C"):
            sino = _oracle_fwd(ps, np.require(img, dtype, o1), dtype, o2)
            assert_control(sino, dims, 6.69419, "sino", f"sinogram {o1}{o2}")
            back = _oracle_back(ps, sino, dtype, o1)
            assert_control(back, dims, 20.2049, "img", f"backprojection {o1}{o2}")


def test_control_fan_projection():
    """tests/test_fanbeam.py:128-150 (1200^2, 360 angles, 600 det, dw=400, RE=200, R=752)."""
    N, nd, angles = 1200, 600, 360
    img = create_phantoms(N, F32)
    ps = _ps(gratopy.FANBEAM, img.shape, angles, nd, detector_width=400.0, RE=200.0, R=752.0)
    dims = (N, N, nd, angles, 2)
    sino = _oracle_fwd(ps, img, F32)
    assert_control(sino, dims, 94031.43, "sino", "sinogram")
    back = _oracle_back(ps, sino, F32)
    assert_control(back, dims, 1482240.72, "img", "backprojection")


@pytest.mark.parametrize("dtype", [F32, F64])
def test_control_fan_types_contiguity(dtype):
    """tests/test_fanbeam.py:166-277."""
    Nx, nd, angles = 600, 300, 180
    img = phantom(Nx, dtype)
    ps = _ps(gratopy.FANBEAM, (Nx, Nx), angles, nd, detector_width=400.0, RE=200.0, R=752.0)
    dims = (Nx, Nx, nd, angles, 1)
    for o1 in ("F", "C"):
        for o2 in ("F", "C"):
            sino = _oracle_fwd(ps, np.require(img, dtype, o1), dtype, o2)
            assert_control(sino, dims, 416.107, "sino", f"sinogram {o1}{o2}")
            back = _oracle_back(ps, sino, dtype, o1)
            assert_control(back, dims, 7851.41, "img", f"backprojection {o1}{o2}")


def test_control_radon_limited_angles():
    """tests/test_radon.py:427-623."""
    N, Ns, na = 1200, int(0.3 * 1200), 180
    img = create_phantoms(N, F32)
    delta = np.pi * 3 / 4 / (na - 1)
    incorrect = np.linspace(0, np.pi * 3 / 4.0, na) + np.pi / 8
    correct = (incorrect, np.pi / 8 - delta / 2, np.pi * 7 / 8 + delta / 2)
    ps_c = _ps(gratopy.RADON, img.shape, correct, Ns, detector_width=2.0)
    ps_i = _ps(gratopy.RADON, img.shape, incorrect, Ns, detector_width=2.0)
    ps_c2 = _ps(gratopy.RADON, img.shape, incorrect, Ns, detector_width=2.0, angle_weights=delta)
    dims = (N, N, Ns, 3, 2)
    sino_c = _oracle_fwd(ps_c, img, F32)
    assert_control(sino_c, dims, 340.93, "sino", "sinogram (correct)")
    assert_control(_oracle_fwd(ps_i, img, F32), dims, 340.93, "sino", "sinogram (incorrect)")
    assert_control(_oracle_fwd(ps_c2, img, F32), dims, 340.93, "sino", "sinogram (weights)")
    assert_control(_oracle_back(ps_c, sino_c, F32), dims, 3385.137, "img", "back (correct)")
    assert_control(_oracle_back(ps_i, sino_c, F32), dims, 3453.203, "img", "back (incorrect)")
    ps_i.set_angle_weights(ps_c.angle_weights)
    assert_control(_oracle_back(ps_i, sino_c, F32), dims, 3385.137, "img", "back (repaired)")


def test_control_radon_nonquadratic_and_shift():
    """tests/test_radon.py:626-699 and 799-884."""
    N = 1200
    full = create_phantoms(N, F32)
    N2 = int(N * 2 / 3.0)
    img = np.asfortranarray(full[:, 0:N2, :])
    ps = _ps(gratopy.RADON, img.shape, 360, int(0.5 * N))
    dims = (N, N2, 600, 360, 2)
    assert_control(img, dims, 999.4965, "img", "image")
    sino = _oracle_fwd(ps, img, F32)
    assert_control(sino, dims, 2401.26, "sino", "sinogram")
    assert_control(_oracle_back(ps, sino, F32), dims, 3310.3464, "img", "backprojection")

    ps = _ps(gratopy.RADON, full.shape, 360, 600, image_width=3.0, detector_width=2.0,
             midpoint_shift=[0.0, 0.4])
    dims = (N, N, 600, 360, 2)
    sino = _oracle_fwd(ps, full, F32)
    assert_control(sino, dims, 2083.80, "sino", "shifted sinogram")
    assert_control(_oracle_back(ps, sino, F32), dims, 6269.703, "img", "shifted backprojection")


def test_control_fan_limited_shift_nonquadratic():
    """tests/test_fanbeam.py:452-675, 678-766, 1385-1475."""
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
    sino_c = _oracle_fwd(ps_c, img, F32)
    assert_control(sino_c, dims, -131.00, "sino", "limited sinogram")
    assert_control(_oracle_back(ps_c, sino_c, F32), dims, 19495.96, "img", "limited back (correct)")
    assert_control(_oracle_back(ps_i, sino_c, F32), dims, 23043.191, "img", "limited back (incorrect)")

    ps = _ps(gratopy.FANBEAM, img.shape, 360, 600, image_width=2.0, R=5.0, RE=3.0,
             detector_width=6.0, midpoint_shift=[0, 0.5])
    dims = (N, N, 600, 360, 2)
    sino = _oracle_fwd(ps, img, F32)
    assert_control(sino, dims, 297.880, "sino", "shifted fan sinogram")
    assert_control(_oracle_back(ps, sino, F32), dims, 12217.92, "img", "shifted fan backprojection")

    N2 = int(N * 2 / 3.0)
    img2 = np.asfortranarray(img[:, 0:N2, :])
    ps = _ps(gratopy.FANBEAM, img2.shape, 360, 600, R=5.0, RE=3.0, detector_width=5.0)
    dims = (N, N2, 600, 360, 2)
    sino = _oracle_fwd(ps, img2, F32)
    assert_control(sino, dims, 530.45, "sino", "non-square fan sinogram")
    assert_control(_oracle_back(ps, sino, F32), dims, 16101.3542, "img", "non-square fan back")


GEOMETRIC_ORIENTATION = [
    ((0, 0, 0), (-22.6691, 1760.90, 131.301, 1157.280)),
    ((50, 0, 0), (51.959, -61.3725, 85.7716, 1058.5918)),
    ((0, 20, 0), (-703.53, 881.399, 542.11, 1050.41)),
    ((0, -20, 0), (390.100, 173.108, 110.7186, 1202.997)),
    ((0, 0, 30), (935.812, 1775.13, 578.73, 951.169)),
    ((0, 0, -30), (51.3383, 1225.368, -322.945, 579.33)),
]


@pytest.mark.parametrize("params,expected", GEOMETRIC_ORIENTATION)
def test_control_geometric_orientation(params, expected):
    """tests/test_fanbeam.py:769-943 (C-ordered float32 images)."""
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
    assert_control(_oracle_fwd(par, img1, F32, "C"), dims, expected[0], "sino", "radon 1")
    assert_control(_oracle_fwd(par, img2, F32, "C"), dims, expected[1], "sino", "radon 2")
    assert_control(_oracle_fwd(fan, img1, F32, "C"), dims, expected[2], "sino", "fan 1")
    assert_control(_oracle_fwd(fan, img2, F32, "C"), dims, expected[3], "sino", "fan 2")


@pytest.mark.parametrize("geometry,expected", [
    ("radon", (6.5315, 1.03955)),
    ("fan", (233.96, 2690.00)),
])
def test_control_sparse_matrix(geometry, expected):
    """tests/test_radon.py:715-796, tests/test_fanbeam.py:1491-1579: the
    single-line kernels assemble a matrix whose products reproduce the
    control numbers and the forward projection."""
    import scipy.sparse

    Nx, nd, angles = 150, 100, 30
    if geometry == "radon":
        ps = _ps(gratopy.RADON, (Nx, Nx), angles, nd)
    else:
        ps = _ps(gratopy.FANBEAM, (Nx, Nx), angles, nd, detector_width=400.0, R=752.0, RE=200.0)
    t = ps.struct
    rows, cols, vals = [], [], []
    for x in range(Nx):
        for y in range(Nx):
            col = oracle.single_line(ps.is_fan, x, y, t["ofs_dict"][F64], t.get("sdpd_dict", {}).get(F64),
                                     t["geo_dict"][F64], ps.img_shape, nd, angles, F64)
            s_idx, a_idx = np.nonzero(col > 0)
            rows.append(s_idx + nd * a_idx)
            cols.append(np.full(len(s_idx), x + Nx * y))
            vals.append(col[s_idx, a_idx])
    M = scipy.sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                shape=(nd * angles, Nx * Nx)).tocsr()
    img = phantom(Nx, F64)
    flat = img.reshape(Nx * Nx, order="F")
    sino = (M @ flat).reshape(nd, angles, order="F")
    back = (M.T @ sino.reshape(-1, order="F")).reshape(Nx, Nx, order="F")
    dims = (Nx, Nx, nd, angles, 1)
    assert_control(sino, dims, expected[0], "sino", "sparse sinogram")
    assert_control(back, dims, expected[1], "img", "sparse backprojection")
    assert rel_err(sino, _oracle_fwd(ps, img, F64)) < 1e-12


# ---- pointwise kernels -----------------------------------------------------------


def test_pointwise_kernels_closed_form():
    rng = np.random.default_rng(11)
    for dtype in (F32, F64):
        sino = np.asfortranarray(rng.random((7, 5, 3)).astype(dtype))
        w = rng.random(5).astype(dtype) + 0.5
        assert np.array_equal(oracle.weight_sino(sino, w), sino * w[None, :, None])
        assert np.array_equal(oracle.weight_sino(sino, w, divide=True), sino / w[None, :, None])
        x, y = rng.random(50).astype(dtype), rng.random(50).astype(dtype)
        r = rng.random(50).astype(dtype)
        expect = r + dtype.type(0.25) * x
        oracle.equ_mul_add(r, 0.25, x)
        assert np.array_equal(r, expect)
        oracle.mul_add_add(r, -1.5, x, y)
        assert np.array_equal(r, dtype.type(-1.5) * x + y)


def _tv_numpy(u, v, sigma, zd):
    """Forward-difference gradient with Neumann boundary (zero last difference)."""
    g = np.zeros((3,) + u.shape)
    g[0, :-1] = u[1:] - u[:-1]
    g[1, :, :-1] = u[:, 1:] - u[:, :-1]
    if zd > 0:
        g[2, :, :, :-1] = (u[:, :, 1:] - u[:, :, :-1]) / zd
    vn = v + sigma * g
    norm = np.sqrt((vn**2).sum(0))
    return vn / np.maximum(norm, 1.0)


def _div_numpy(v, zd):
    """Negative adjoint of the gradient above."""
    d = np.zeros(v.shape[1:])
    for k in range(3 if zd > 0 else 2):
        vk = np.moveaxis(v[k], k, 0)
        dk = np.zeros_like(vk)
        dk[0] = vk[0]
        dk[1:-1] = vk[1:-1] - vk[:-2]
        dk[-1] = -vk[-2]
        d += np.moveaxis(dk, 0, k) / (zd if k == 2 else 1.0)
    return d


@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("zd", [0.0, 0.6])
def test_tv_div_is_minus_grad_transpose(order, zd):
    """<grad u, p> = -<u, div p> for the oracle's update_v / update_u pair
    (the identity the hand-written TV adjoint relies on)."""
    rng = np.random.default_rng(21)
    shape = (7, 6, 5)
    u = np.require(rng.random(shape), F64, order)
    p = rng.normal(size=4 * u.size)
    p[3::4] = 0
    # grad u via update_v on a zero dual variable with sigma = 1
    V = np.zeros(4 * u.size)
    oracle.tv_update_v(V, u, 1.0, zd)
    # div p via update_u with u_old = 0, tau = 1, no data term (then clamp is inactive: add a shift)
    shift = 100.0
    out = np.require(np.zeros(shape), F64, order)
    oracle.tv_update_u(out, np.require(np.full(shape, shift), F64, order), p,
                       np.require(np.zeros(shape), F64, order), 1.0, 0.0, zd)
    div = np.asarray(out) - shift
    lhs = float(np.dot(V, p))
    rhs = -float(np.sum(np.asarray(u) * div))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("order", ["F", "C"])
def test_tv_updates_match_numpy(order):
    """The TV kernels equal an independent numpy gradient/divergence
    (the divergence is minus the adjoint of the gradient)."""
    rng = np.random.default_rng(12)
    shape, zd = (9, 7, 4), 0.7
    u = np.require(rng.random(shape), F64, order)
    v0 = rng.normal(size=(3,) + shape) * 0.3
    # the dual variable holds v0[:, x, y, z] at 4*pos(x, y, z) + k
    idx = np.arange(u.size).reshape(shape, order=order)
    V = np.zeros(4 * u.size)
    for k in range(3):
        V[4 * idx + k] = v0[k]
    oracle.tv_update_v(V, u, 0.6, zd)
    oracle.tv_update_normv(V, shape, order)
    expect = _tv_numpy(np.asarray(u), v0, 0.6, zd)
    got = np.stack([V[4 * idx + k] for k in range(3)])
    assert rel_err(got, expect) < 1e-6  # sigma and z_distance are float32 in the kernel

    # adjointness of grad/div on the oracle kernels: <grad u, v> = -<u, div v>
    kstar = np.zeros(shape, order=order)
    unew = np.zeros(shape, order=order)
    uold = np.require(np.full(shape, 10.0), F64, order)
    oracle.tv_update_u(unew, uold, V, kstar, 1.0, 0.0, zd)
    div = np.asarray(unew) - 10.0
    assert rel_err(div, _div_numpy(got, np.float32(zd))) < 1e-6
    lam = np.require(rng.random((5, 4)), F64, "F")
    ku, f = np.require(rng.random((5, 4)), F64, "F"), np.require(rng.random((5, 4)), F64, "F")
    expect = (lam + np.float32(0.3) * (ku - f)) * np.float32(0.9)
    oracle.tv_update_lambda(lam, ku, f, 0.3, 0.9)
    assert np.allclose(lam, expect, rtol=1e-15, atol=0)
    a, b = rng.random(10), rng.random(10)
    e = 2 * a - b
    oracle.tv_update_extra(a, b)
    assert np.array_equal(b, e)
