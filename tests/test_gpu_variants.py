"""Every launch variant the autotuner can pick (forward bin groups K = 1, 2,
4, 8, 16, with or without 8 row bands, reading strip rows one pixel or 16
bytes per load; backprojection rows per lane RY = 1, 2, 4, 8, with or without
8 angle bands, or with sinogram windows staged in LDS for chunks of 8 or 16
angles in tiles 16 or 32 pixels wide) meets the parity bar against the
oracle: normwise 1e-5 in float32, 1e-12 in float64, on regular, shifted,
ragged and degenerate geometries (including fan bins so wide that a bin
group's union strip is irregular and the group falls back to per-bin rays),
and on detectors with many bins per pixel, where the staged backprojection
reads some or all angles from memory."""

from __future__ import annotations

import numpy as np
import pytest

import gratopy_amd as gratopy
from gratopy_amd import autotune
from tests.helpers import rel_err
from tests.test_gpu_projection import TOL, _hip, _oracle

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")

CASES = [
    ("radon", (40, 33), 17, 61, {}),
    ("radon", (5, 29), 3, 1, dict(detector_width=2.5, midpoint_shift=(0.3, -0.2), detector_shift=0.4)),
    ("fan", (40, 33), 17, 61, dict(R=160.0, RE=70.0, detector_width=120.0)),
    ("fan", (1, 5), 7, 2, dict(R=800.0, RE=400.0, detector_width=300.0)),
    ("fan", (5, 29), 3, 1, dict(R=800.0, RE=400.0, detector_width=300.0)),
    ("fan", (17, 3), 9, 40, dict(R=90.0, RE=40.0, detector_width=60.0, detector_shift=3.0)),
]

# Many detector bins per pixel: the staged backprojection's windows exceed its LDS
# width for every tile (75 bins per pixel), for some angles and tiles (15), or only
# near the source (fan, 4 to 12 bins per pixel), so those angles take the direct
# path.  At these detector counts float32 has a floor of its own: x*X + y*Y + off
# is ~Ns in size, so the reference algorithm in float32 (the oracle in float32)
# is ~1e-4 away from the float64 result.  float64 is held to 1e-12; float32 to
# 1.5 times the reference algorithm's own float32 distance from float64.
WIDE_CASES = [
    ("radon", (40, 33), 17, 3000, {}),
    ("radon", (40, 33), 17, 600, {}),
    ("fan", (48, 40), 23, 900, dict(R=120.0, RE=60.0, detector_width=300.0, image_width=48.0)),
]


def _check_wide(ps, x, adjoint, tag):
    kw = dict(in_order="C", out_order="C") if adjoint else {}
    truth = _oracle(ps, x, F64, adjoint)
    scale = max(np.abs(truth).max(), 1e-3)
    got64 = _hip(ps, x, F64, adjoint, **kw)
    assert np.abs(got64 - truth).max() <= TOL[F64] * scale, (tag, "float64")
    e_ref = np.abs(_oracle(ps, x, F32, adjoint) - truth).max() / scale
    e_hip = np.abs(_hip(ps, x, F32, adjoint, **kw) - truth).max() / scale
    assert e_hip <= max(TOL[F32], 1.5 * e_ref), (tag, "float32", e_hip, e_ref)


@pytest.fixture
def restore_variant():
    before = autotune.current_variant()
    yield
    autotune.set_variant(**before)


FWD_IDS = [autotune.variant_key(v) for v in autotune.FWD_VARIANTS]
BACK_IDS = [autotune.variant_key(v) for v in autotune.BACK_VARIANTS]


@pytest.mark.parametrize("variant", autotune.FWD_VARIANTS, ids=FWD_IDS)
def test_forward_variants(variant, restore_variant):
    print(f"variant {autotune.variant_key(variant)}", flush=True)
    autotune.set_variant(**variant)
    for geometry, shape, na, nd, kw in CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        img = np.random.default_rng(4).random(shape + (2,))
        for dtype in (F32, F64):
            ref = _oracle(ps, img, dtype, False)
            got = _hip(ps, img, dtype, False)
            scale = max(np.abs(ref).max(), 1e-3)
            assert np.abs(got - ref).max() <= TOL[dtype] * scale, (geometry, shape, variant, dtype)
    for geometry, shape, na, nd, kw in WIDE_CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        _check_wide(ps, np.random.default_rng(4).random(shape + (2,)), False, (geometry, nd, variant))


GROUP_VARIANTS = [v for v in autotune.FWD_VARIANTS if v["fwd_group"] > 1]


@pytest.mark.parametrize("variant", GROUP_VARIANTS,
                         ids=[autotune.variant_key(v) for v in GROUP_VARIANTS])
def test_forward_lane_mapping(variant, restore_variant):
    """Bin-group launches cover every (angle, group) pair once: whole columns
    of 64 groups per 4 angles, then the leftover groups of all angles packed
    into full waves.  Detector counts give whole columns with a leftover
    (1124), whole columns only (1024 for K = 16; 1024 / K groups otherwise)
    and leftovers only or mostly (200); angle counts below, at and past a
    block of 4, so the last block of either kind is partly past the end.
    With this many bins per pixel float32 sits at its own floor, so it is
    held to the reference algorithm's float32 distance (as the wide cases)."""
    print(f"variant {autotune.variant_key(variant)} lane mapping", flush=True)
    autotune.set_variant(**variant)
    shape = (24, 20)
    img = np.random.default_rng(6).random(shape + (2,))
    for geometry in ("radon", "fan"):
        for nd in (200, 1024, 1124):
            kw = dict(R=60.0, RE=25.0, detector_width=70.0) if geometry == "fan" else {}
            for na in (1, 3, 6):
                g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
                ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
                _check_wide(ps, img, False, (geometry, nd, na))


@pytest.mark.parametrize("variant", autotune.BACK_VARIANTS, ids=BACK_IDS)
def test_backward_variants(variant, restore_variant):
    print(f"variant {autotune.variant_key(variant)}", flush=True)
    autotune.set_variant(**variant)
    for geometry, shape, na, nd, kw in CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        sino = np.random.default_rng(5).random(ps.sinogram_shape + (2,))
        for dtype in (F32, F64):
            ref = _oracle(ps, sino, dtype, True)
            got = _hip(ps, sino, dtype, True, in_order="C", out_order="C")
            scale = max(np.abs(ref).max(), 1e-3)
            assert np.abs(got - ref).max() <= TOL[dtype] * scale, (geometry, shape, variant, dtype)
    for geometry, shape, na, nd, kw in WIDE_CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        sino = np.random.default_rng(5).random(ps.sinogram_shape + (2,))
        _check_wide(ps, sino, True, (geometry, nd, variant))


STAGED_VARIANTS = [v for v in autotune.BACK_VARIANTS if v["back_stage"]]

# Detectors narrower than the image, shifted, or both: most tiles' windows
# start before the padded sinogram row (negative indices) or end past it.
EDGE_CASES = [
    ("radon", (64, 48), 90, 16, dict(detector_width=0.5)),
    ("radon", (64, 48), 45, 24, dict(detector_width=0.8, detector_shift=0.6)),
    ("fan", (64, 48), 60, 20, dict(R=300.0, RE=150.0, detector_width=40.0, detector_shift=15.0)),
]


@pytest.mark.parametrize("variant", STAGED_VARIANTS,
                         ids=[autotune.variant_key(v) for v in STAGED_VARIANTS])
def test_staged_windows_past_the_row_ends(variant, restore_variant):
    """The window copies of the staged backprojection read the padded row
    through a buffer descriptor whose range check supplies the zeros outside
    it, negative indices included.  A build that let the compiler split such
    an offset into a negative register part and a positive instruction offset
    read zeros where the row has data (float32 error 3e-3 against the
    oracle)."""
    print(f"variant {autotune.variant_key(variant)} edge windows", flush=True)
    autotune.set_variant(**variant)
    for geometry, shape, na, nd, kw in EDGE_CASES:
        g = gratopy.RADON if geometry == "radon" else gratopy.FANBEAM
        ps = gratopy.ProjectionSettings("cuda:0", g, shape, na, nd, **kw)
        sino = np.random.default_rng(7).random(ps.sinogram_shape + (2,))
        for dtype in (F32, F64):
            ref = _oracle(ps, sino, dtype, True)
            got = _hip(ps, sino, dtype, True, in_order="C", out_order="C")
            scale = max(np.abs(ref).max(), 1e-3)
            assert np.abs(got - ref).max() <= TOL[dtype] * scale, (geometry, nd, variant, dtype)


_LARGE = {}


def _large_case(geometry):
    """512^2 random-ellipse image, 360 angles, 768 detectors (fanbeam: RE=400,
    R=800, detector width 1024, image width 512): float32 input, float64 and
    float32 oracle forward projections."""
    if geometry not in _LARGE:
        from gratopy_amd.phantom import random_ellipses

        if geometry == "fan":
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (512, 512), 360, 768, R=800.0,
                                            RE=400.0, detector_width=1024.0, image_width=512.0)
        else:
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (512, 512), 360, 768)
        img = random_ellipses(512, 15, seed=2, dtype=F32)
        _LARGE[geometry] = (ps, img, _oracle(ps, img, F64, False), _oracle(ps, img, F32, False))
    return _LARGE[geometry]


@pytest.mark.parametrize("variant", autotune.FWD_VARIANTS, ids=FWD_IDS)
@pytest.mark.parametrize("geometry", ["radon", "fan"])
def test_forward_variants_large_float32(geometry, variant, restore_variant):
    """At 768 detectors the float32 result of every variant is within 1e-5 of
    the float64 oracle, or within 1.2x the distance of the reference
    algorithm in float32 (the oracle in float32).  At this size both sit at
    the float32 floor (~1.4e-5 for Radon): the float32 geometry tables alone
    put ~3e-5 detector bins of error into x*X + y*Y + off."""
    name = autotune.variant_key(variant)
    print(f"variant {name} large {geometry}", flush=True)
    ps, img, truth, ref32 = _large_case(geometry)
    autotune.set_variant(**variant)
    got = _hip(ps, img, F32, False)
    e_hip, e_ref = rel_err(got, truth), rel_err(ref32, truth)
    print(f"{name} {geometry}: float32 error vs float64 oracle: hip {e_hip:.3e}, "
          f"reference algorithm {e_ref:.3e}", flush=True)
    assert e_hip <= max(TOL[F32], 1.2 * e_ref)


def test_autotuner_picks_a_variant_and_caches_it(tmp_path, monkeypatch, restore_variant):
    monkeypatch.setenv("GRATOPY_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    from gratopy_amd import _device as dev

    ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (96, 96), 60, 128, R=400.0, RE=200.0,
                                    detector_width=250.0)
    img = dev.to_device(np.asfortranarray(np.random.default_rng(1).random((96, 96)).astype(F32)), "cuda:0")
    sino = gratopy.forwardprojection(img, ps)
    best = autotune.tune(ps, img.clone(), sino.clone(), repeats=2)
    assert best["fwd_group"] in autotune.FWD_GROUPS and best["back_rows"] in autotune.BACK_ROWS
    chosen = {k: best[k] for k in autotune.KNOBS if k != "automatic"}
    assert autotune.current_variant() == {**chosen, "automatic": 0}
    assert set(best["fwd_ms"]) == set(FWD_IDS) and set(best["back_ms"]) == set(BACK_IDS)
    autotune.set_variant(**autotune.DEFAULTS)
    assert autotune.current_variant()["automatic"] == 1
    assert autotune.apply(ps, F32, 1)
    assert autotune.current_variant() == {**chosen, "automatic": 0}
    # all variants agree to rounding on the tuned geometry
    ref = dev.to_host(gratopy.forwardprojection(img, ps))
    for variant in autotune.FWD_VARIANTS:
        autotune.set_variant(**variant)
        assert rel_err(dev.to_host(gratopy.forwardprojection(img, ps)), ref) < 1e-5, variant
