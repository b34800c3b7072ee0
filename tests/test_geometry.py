"""Host-side geometry of the drop-in API (CPU only): angle parsing and angle
weights against the known answers of the reference's
tests/test_radon.py:887-1117, ProjectionSettings attributes, table layouts,
and the error types of the reference."""

from __future__ import annotations

import numpy as np
import pytest

import gratopy_amd as gratopy
from gratopy_amd.gratopy import check_compatibility

PI = np.pi


def _angle_cases():
    cases = []
    cases.append((100, np.linspace(0, PI, 101)[:-1], np.ones(100) * PI / 100))
    cases.append((np.linspace(0, PI, 101)[:-1], np.linspace(0, PI, 101)[:-1], np.ones(100) * PI / 100))
    mixed = (list(np.linspace(0, PI / 3, 101)[:-1]) + list(np.linspace(PI / 3, PI * 2.0 / 3, 51)[:-1])
             + list(np.linspace(PI * 2 / 3, PI, 101)[:-1]))
    mixed_expected = (list(np.arange(0, PI / 3 - 0.00001, PI / 300))
                      + list(np.arange(PI / 3, PI * 2 / 3 - 0.00001, PI / 150))
                      + list(np.arange(PI * 2 / 3, PI - 0.00001, PI / 300)))
    w = list(np.ones(100) * PI / 300) + list(np.ones(50) * PI / 150) + list(np.ones(100) * PI / 300)
    w[100] = (PI / 300 + PI / 150) * 0.5
    w[150] = (PI / 300 + PI / 150) * 0.5
    cases.append((mixed, mixed_expected, w))
    cases.append((np.array(mixed), mixed_expected, w))
    cases.append((
        [(list(np.linspace(0, PI / 2, 101)[:-1] + PI / 4 / 100), 0, PI / 2),
         (50, PI / 2, PI * 3 / 4),
         (np.linspace(PI * 3 / 4, PI, 51)[:-1] + PI / 4 / 100, PI * 3 / 4, PI)],
        list(np.arange(PI * 1 / 400, PI - 0.00001, PI / 200)), np.ones(200) * PI / 200))
    cases.append(([(50, 0, PI / 3), (50, PI * 2 / 3, PI)],
                  list(np.arange(PI / 300, PI / 3 - 0.00001, PI / 150))
                  + list(np.arange(PI * 201 / 300, PI - 0.00001, PI / 150)), np.ones(100) * PI / 150))
    cases.append(([(1, 0, PI)], PI / 2, [PI]))
    cases.append((np.linspace(0, 2 * PI, 101)[:-1], np.arange(0, 2 * PI - 0.00001, PI / 50),
                  list(np.ones(100) * PI / 100)))
    cases.append(([(np.linspace(0, 2 * PI, 101)[:-1] + PI / 100, 0, 2 * PI)],
                  np.arange(PI / 100, 2 * PI - 0.00001, PI / 50), list(np.ones(100) * PI / 50)))
    cases.append((-np.linspace(0, PI, 101)[:-1], np.arange(0, -(PI - 0.00001), -PI / 100),
                  list(np.ones(100) * PI / 100)))
    cases.append(((-100, 0, PI), np.flip(np.arange(PI / 200, PI - 0.00001, PI / 100)),
                  list(np.ones(100) * PI / 100)))
    cases.append((-100, np.flip(np.arange(0, (PI - 0.00001), PI / 100)), list(np.ones(100) * PI / 100)))
    cases.append((([0, 0, PI / 2, PI / 2, PI / 2, PI * 2 / 3, 0], 0, PI),
                  [0, 0, PI / 2, PI / 2, PI / 2, PI * 2 / 3, 0],
                  [PI / 12, PI / 12, (PI / 4 + PI / 12) / 3, (PI / 4 + PI / 12) / 3,
                   (PI / 4 + PI / 12) / 3, (PI / 12 + PI / 3), PI / 12]))
    cases.append(([0, 0, PI / 2, PI / 2, PI / 2, PI * 2 / 3, 0],
                  [0, 0, PI / 2, PI / 2, PI / 2, PI * 2 / 3, 0],
                  [(PI / 4 + PI / 6) / 3, (PI / 4 + PI / 6) / 3, (PI / 4 + PI / 12) / 3,
                   (PI / 4 + PI / 12) / 3, (PI / 4 + PI / 12) / 3, PI / 6 + PI / 12,
                   (PI / 4 + PI / 6) / 3]))
    return cases


ANGLE_CASES = _angle_cases()


@pytest.mark.parametrize("k", range(len(ANGLE_CASES)))
def test_angle_input_variants(k):
    """tests/test_radon.py:887-1085: angles and weights of every input form."""
    angles, expected, weights = ANGLE_CASES[k]
    ps = gratopy.ProjectionSettings(None, gratopy.PARALLEL, (300, 300), angles, 150,
                                    image_width=4.0, detector_width=4.0)
    assert np.linalg.norm(ps.angles - np.array(expected)) < 0.01
    assert np.linalg.norm(ps.angle_weights - np.array(weights)) < 0.001


@pytest.mark.parametrize("k", range(len(ANGLE_CASES)))
def test_angle_weight_overrides(k):
    """tests/test_radon.py:1087-1111: every angle_weights form is accepted
    for both geometries."""
    angles, _, weights = ANGLE_CASES[k]
    na = len(weights)
    for angle_weights in [1, 1.0, np.ones(na), list(np.ones(na)), None]:
        for geometry in (gratopy.PARALLEL, gratopy.FAN):
            ps = gratopy.ProjectionSettings(None, geometry, (300, 300), angles, 150,
                                            angle_weights=angle_weights, image_width=4.0, R=8, RE=4,
                                            detector_width=4.0, detector_shift=0)
            assert ps.n_angles == na
            if angle_weights is not None:
                assert np.allclose(ps.angle_weights, 1.0)


def test_angle_weights_length_mismatch_asserts():
    with pytest.raises(AssertionError):
        gratopy.ProjectionSettings(None, gratopy.RADON, (10, 10), 5, angle_weights=[1.0, 2.0])


def test_radon_settings_attributes_and_tables():
    ps = gratopy.ProjectionSettings(None, gratopy.RADON, (40, 30), 12, 50, detector_width=3.0,
                                    image_width=2.0, midpoint_shift=(0.1, 0.2), detector_shift=0.05)
    assert ps.is_parallel and not ps.is_fan
    assert ps.sinogram_shape == (50, 12) and ps.img_shape == (40, 30)
    assert ps.delta_x == 2.0 / 40 and ps.delta_s == 3.0 / 50
    assert ps.delta_ratio == ps.delta_s / ps.delta_x
    for dt in (np.float32, np.float64):
        ofs = ps.struct["ofs_dict"][np.dtype(dt)]
        assert ofs.shape == (8, 12) and ofs.flags.f_contiguous and ofs.dtype == dt
        geo = ps.struct["geo_dict"][np.dtype(dt)]
        assert list(geo[2:]) == [40, 30, 50, 12]
    ofs = ps.struct["ofs_dict"][np.dtype("float64")]
    rho = ps.delta_s / ps.delta_x
    assert np.allclose(np.hypot(ofs[0], ofs[1]), 1 / rho)
    flipped = ofs[5] == 1
    assert np.allclose(ofs[3][~flipped], 1 / ofs[0][~flipped])
    assert np.allclose(ofs[3][flipped], 1 / ofs[1][flipped])
    assert np.all(np.abs(ofs[0][~flipped]) >= np.abs(ofs[1][~flipped]))
    assert np.allclose(ofs[4], ps.angle_weights)


def test_fan_settings_tables():
    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM, (64, 64), 20, 90, R=300.0, RE=120.0,
                                    detector_width=100.0, image_width=60.0)
    ofs = ps.struct["ofs_dict"][np.dtype("float64")]
    geo = ps.struct["geo_dict"][np.dtype("float64")]
    unit = 64 / 60.0
    assert np.isclose(geo[0], 300 * unit) and np.isclose(geo[1], 120 * unit)
    assert np.isclose(geo[2], 100 * unit / 90) and np.isclose(geo[10], 60 / 64)
    assert np.allclose(np.hypot(ofs[0], ofs[1]), geo[2])        # detector step length
    assert np.allclose(np.hypot(ofs[2], ofs[3]), geo[1])        # origin -> source
    assert np.allclose(np.hypot(ofs[4], ofs[5]), geo[0] - geo[1])  # origin -> detector
    sdpd = ps.struct["sdpd_dict"][np.dtype("float64")][0]
    assert np.allclose(sdpd.min(), np.sqrt(geo[0] ** 2 + (0.5 * geo[2]) ** 2))


def test_fan_geometry_assertions():
    with pytest.raises(AssertionError, match="R"):
        gratopy.ProjectionSettings(None, gratopy.FANBEAM, (10, 10), 5)
    with pytest.raises(AssertionError, match="origin is not between"):
        gratopy.ProjectionSettings(None, gratopy.FANBEAM, (10, 10), 5, R=1.0, RE=2.0)
    with pytest.raises(AssertionError, match="source is not outside"):
        gratopy.ProjectionSettings(None, gratopy.FANBEAM, (10, 10), 5, R=10.0, RE=1.0, image_width=4.0)
    # the automatic image width keeps the image inside the fan
    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM, (10, 10), 5, R=10.0, RE=5.0, detector_width=4.0)
    slope = 2.0 / 10.0
    assert np.isclose(ps.image_width, 2 * slope * 5.0 / np.sqrt(1 + slope**2))


def test_fan_default_detector_count():
    """The reference divides by the raw n_detectors argument and fails on None
    (gratopy.py:1594); the drop-in resolves ceil(hypot(Nx, Ny)) first."""
    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM, (30, 40), 5, R=500.0, RE=200.0, detector_width=50.0)
    assert ps.n_detectors == 50 and ps.delta_s == 1.0


def test_check_compatibility_asserts():
    import torch

    ps = gratopy.ProjectionSettings(None, gratopy.RADON, (8, 6), 4, 10)
    img = torch.zeros(8, 6)
    sino = torch.zeros(10, 4)
    check_compatibility(img, sino, ps)
    with pytest.raises(AssertionError, match="data type"):
        check_compatibility(img.double(), sino, ps)
    with pytest.raises(AssertionError, match="dimensions of the sinogram"):
        check_compatibility(img, torch.zeros(9, 4), ps)
    with pytest.raises(AssertionError, match="dimensions of the image"):
        check_compatibility(torch.zeros(6, 8), sino, ps)
    with pytest.raises(AssertionError, match="third dimension"):
        check_compatibility(img, torch.zeros(10, 4, 3), ps)
    with pytest.raises(AssertionError, match="third dimension"):
        check_compatibility(torch.zeros(8, 6, 2), torch.zeros(10, 4, 3), ps)
    with pytest.raises(AssertionError):
        check_compatibility(torch.zeros(8, 6, 3), sino, ps)


def test_projection_requires_device_arrays():
    import torch

    ps = gratopy.ProjectionSettings(None, gratopy.RADON, (8, 6), 4, 10)
    with pytest.raises(TypeError, match="device array"):
        gratopy.forwardprojection(np.zeros((8, 6), np.float32), ps)
    with pytest.raises(TypeError, match="HIP device"):
        gratopy.forwardprojection(torch.zeros(8, 6), ps)


def test_geometry_value_and_constants():
    assert gratopy.RADON is gratopy.PARALLEL and gratopy.FAN is gratopy.FANBEAM
    ps = gratopy.ProjectionSettings(None, "fanbeam", (8, 8), 4, 10, R=100.0, RE=50.0)
    assert ps.is_fan and ps.geometry == gratopy.FANBEAM
    with pytest.raises(ValueError):
        gratopy.ProjectionSettings(None, "cone", (8, 8), 4, 10)
