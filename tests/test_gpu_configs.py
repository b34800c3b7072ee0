"""The benchmark configurations' workloads on the GPU against the oracle.

config 1  Radon 256^2, 180 angles in [0, pi), 256 detectors, 10-ellipse
          phantom: forward, backprojection and Landweber(100) on the noiseless
          sinogram (radon.cl:66-241, gratopy.py:2380-2446), float32 and float64.
config 2  fanbeam 512^2, 360 angles in [0, 2pi), 768 detectors, RE 400, R 800,
          detector width 1024 and image width 512 (pixel units), 15 ellipses,
          1% noise: CG(50) and TV(200, mu 1e-2) (gratopy.py:2449-2762), float32
          and float64.
config 3  the 64 uneven slices of bench.py (128..1024 px, 60..720 angles,
          ceil(1.5 Nx) detectors, Radon and fanbeam alternating, 0.5% noise):
          batched + graph-captured Landweber(100) and TV(100) bitwise equal to
          the per-slice functions for all 64 slices, and the batched results of
          the smallest and the largest slice of each geometry against the
          oracle's Landweber (5 iterations for the largest: 100 would cost the
          oracle minutes per slice).
config 5  the 32 x 2048^2 fanbeam stack of bench.py (1440 angles, 2560
          detectors, RE 2048, R 4096): TV batched + graph-captured bitwise equal
          to the plain per-slice loop over 10 iterations; full-size one-slice
          backprojection with all 1440 angles against the oracle in float32, and
          float64 forward (every 16th angle) and backprojection against 1e-12.
(config 4, the differentiation checks at 512^2, is in test_gpu_autograd.py and
test_gpu_tv_unrolled.py.)

Every run shares the same injected operator-norm estimate on both sides.
Float32 comparisons record three numbers (tests/helpers.py::record_parity):
e_hip_ref = HIP float32 vs the oracle in float32 (the reference's float32
arithmetic), e_hip = HIP float32 vs the float64 oracle on the same
float32-valued data (the truth), e_ref = the oracle's float32 vs that truth.
Float64 results are held to the oracle at the bars stated per test.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import bench
import gratopy_amd as gratopy
import oracle
from gratopy_amd import _device as dev
from gratopy_amd.batch import batch_landweber, batch_total_variation
from gratopy_amd.phantom import add_noise, random_ellipses
from oracle import solvers as osolv
from tests.helpers import record_parity, rel_err

pytestmark = pytest.mark.gpu

F32, F64 = np.dtype("float32"), np.dtype("float64")
DEV = "cuda:0"
_CACHE: dict = {}


@pytest.fixture
def pinned_variant():
    """One launch variant for every call (the bench's autotuned choice on the
    2048^2 stack): bitwise comparisons between a batched and a per-slice path
    need the same kernels and row bands on both sides, which the automatic
    shapes (chosen per call size) do not give."""
    from gratopy_amd import autotune

    before = autotune.current_variant()
    autotune.set_variant(fwd_group=16, fwd_bands=8, fwd_chunk=4, back_rows=8,
                         back_bands=1, back_stage=16, back_tile=32)
    yield
    autotune.set_variant(**before)


def _cached(key, fn):
    if key not in _CACHE:
        _CACHE[key] = fn()
    return _CACHE[key]


def _gpu_norm(ps):
    """||R|| from the GPU power iteration in float64 from a seeded start; the
    same value is handed to the HIP solvers and the oracle's."""
    np.random.seed(0)
    return float(gratopy.normest(ps, number_iterations=50, dtype="float64"))


def _three(hip32, ref32, truth):
    return {"e_hip_ref": rel_err(hip32, ref32), "e_hip": rel_err(hip32, truth),
            "e_ref": rel_err(ref32, truth)}


# ---------------------------------------------------------------------------
# config 1
# ---------------------------------------------------------------------------
def _config1():
    def make():
        ps = gratopy.ProjectionSettings(DEV, gratopy.RADON, (256, 256), 180, 256)
        g = osolv.OracleGeometry.from_settings(ps)
        img64 = np.asfortranarray(random_ellipses(256, 10, seed=101, dtype=F64))
        data32 = g.fwd(np.asfortranarray(img64.astype(F32)))  # the noiseless float32 sinogram
        return ps, g, img64, data32, _gpu_norm(ps)

    return _cached("config1", make)


def test_config1_projections():
    print("config1 projections", flush=True)
    ps, g, img64, data32, _ = _config1()
    img32 = np.asfortranarray(img64.astype(F32))
    s_truth = g.fwd(img32.astype(F64))
    s_ref = g.fwd(img32)
    s_hip = dev.to_host(gratopy.forwardprojection(dev.to_device(img32, DEV), ps))
    fwd = _three(s_hip, s_ref, s_truth)
    b_truth = g.back(data32.astype(F64))
    b_ref = g.back(data32)
    b_hip = dev.to_host(gratopy.backprojection(dev.to_device(data32, DEV), ps))
    back = _three(b_hip, b_ref, b_truth)
    f64 = {"fwd": rel_err(dev.to_host(gratopy.forwardprojection(dev.to_device(img32.astype(F64), DEV), ps)),
                          s_truth),
           "back": rel_err(dev.to_host(gratopy.backprojection(dev.to_device(data32.astype(F64), DEV), ps)),
                           b_truth)}
    record_parity("config1", "projections", fwd_f32=fwd, back_f32=back, f64=f64)
    assert fwd["e_hip_ref"] <= 1e-5 and back["e_hip_ref"] <= 1e-5, (fwd, back)
    assert f64["fwd"] <= 1e-12 and f64["back"] <= 1e-12, f64


def test_config1_landweber100():
    print("config1 landweber(100)", flush=True)
    ps, g, _, data32, norm = _config1()
    data64 = data32.astype(F64)
    truth = osolv.landweber(data64, g, 100, w=1.0, norm_estimate=norm)
    ref32 = osolv.landweber(data32, g, 100, w=1.0, norm_estimate=norm)
    hip32 = dev.to_host(gratopy.landweber(dev.to_device(data32, DEV), ps, 100, norm_estimate=norm))
    hip64 = dev.to_host(gratopy.landweber(dev.to_device(data64, DEV), ps, 100, norm_estimate=norm))
    e32 = _three(hip32, ref32, truth)
    e64 = rel_err(hip64, truth)
    record_parity("config1", "landweber100", f32=e32, f64=e64)
    assert e32["e_hip_ref"] <= 1e-5, e32
    assert e64 <= 1e-12, e64


# ---------------------------------------------------------------------------
# config 2
# ---------------------------------------------------------------------------
def _config2():
    def make():
        ps = gratopy.ProjectionSettings(DEV, gratopy.FANBEAM, (512, 512), 360, 768, RE=400.0, R=800.0,
                                        image_width=512.0, detector_width=1024.0)
        g = osolv.OracleGeometry.from_settings(ps)
        img64 = np.asfortranarray(random_ellipses(512, 15, seed=102, dtype=F64))
        clean = g.fwd(img64)
        data32 = np.asfortranarray(add_noise(clean, 1.0, np.random.default_rng(202)).astype(F32))
        return ps, g, data32, _gpu_norm(ps)

    return _cached("config2", make)


def _config2_tv(dtype):
    """(oracle, HIP) TV(200, mu 1e-2) on config 2's data in `dtype` (float64
    runs take the float32 data cast up, so they are also the truth of the
    float32 runs)."""
    ps, g, data32, norm = _config2()
    data = data32 if dtype == F32 else data32.astype(F64)
    ref = _cached(("config2", "tv", dtype.name),
                  lambda: osolv.total_variation(data, g, 1e-2, 200, norm_estimate=norm))
    hip = gratopy.total_variation(dev.to_device(data, DEV), ps, 1e-2, 200, norm_estimate=norm)
    return ref, dev.to_host(hip)


@pytest.mark.timeout(600)
def test_config2_tv_float64():
    print("config2 tv(200) float64", flush=True)
    ref, hip = _config2_tv(F64)
    e = rel_err(hip, ref)
    record_parity("config2", "tv200_f64", e_hip_ref=e)
    assert e <= 1e-12, e


@pytest.mark.timeout(600)
def test_config2_tv_float32():
    print("config2 tv(200) float32", flush=True)
    truth, _ = _config2_tv(F64)
    ref32, hip32 = _config2_tv(F32)
    e = _three(hip32, ref32, truth)
    record_parity("config2", "tv200_f32", **e)
    assert e["e_hip"] <= max(1e-5, 1.5 * e["e_ref"]), e


CG_ITERS = (1, 2, 3, 5, 50)


def _residual(g, x, f64):
    """||R x - f|| / ||f|| with the float64 oracle operator."""
    r = g.fwd(np.asfortranarray(np.asarray(x, F64))) - f64
    return float(np.linalg.norm(r) / np.linalg.norm(f64))


@pytest.mark.timeout(600)
def test_config2_cg_float64():
    """CG on these normal equations is chaotic.  Its iterates develop
    high-frequency components, on which two float64 implementations of the
    projections differ by ~5e-13 (the cancellation in the detector
    coordinate, ~Ns in size, leaves ~1e-13 of a bin in both; smooth inputs
    agree to ~1e-14).  CG then amplifies such differences about tenfold
    every few iterations, as it does a relative 1e-14 perturbation of the
    data (the oracle moves by ~1e-3 after 50 iterations).  So: the first
    five iterations agree to 1e-12; after 50, the HIP solution is no further
    from the oracle's than 10x the oracle's own spread under that data
    perturbation, and its data residual is within 1% of the oracle's."""
    print("config2 cg float64 at 1, 2, 3, 5 and 50 iterations", flush=True)
    ps, g, data32, _ = _config2()
    d64 = data32.astype(F64)
    xs = _cached(("config2", "cg_iterates"), lambda: osolv.conjugate_gradients_iterates(d64, g, CG_ITERS))
    pert = np.asfortranarray(d64 * (1 + 1e-14 * np.random.default_rng(5).standard_normal(d64.shape)))
    xp = osolv.conjugate_gradients_iterates(pert, g, (50,))
    out = {}
    for n in CG_ITERS:
        hip = dev.to_host(gratopy.conjugate_gradients(dev.to_device(d64, DEV), ps, n))
        out[n] = {"e_hip_ref": rel_err(hip, xs[n])}
    out[50]["e_oracle_perturbed"] = rel_err(xp[50], xs[50])
    out["residual50"] = {"hip": _residual(g, hip, d64), "oracle": _residual(g, xs[50], d64)}
    record_parity("config2", "cg_f64", **{str(k): v for k, v in out.items()})
    for n in CG_ITERS[:-1]:
        assert out[n]["e_hip_ref"] <= 1e-12, (n, out[n])
    assert out[50]["e_hip_ref"] <= 10 * out[50]["e_oracle_perturbed"], out[50]
    r = out["residual50"]
    assert abs(r["hip"] - r["oracle"]) <= 1e-2 * r["oracle"], r


@pytest.mark.timeout(600)
def test_config2_cg_float32():
    """Float32 CG(50) is as chaotic: HIP and the reference algorithm in
    float32 each end a few per cent from the float64 solution.  The HIP run
    must be no further from it than 1.5x the reference algorithm's float32
    run, and its data residual within 5% of the float64 solution's."""
    print("config2 cg(50) float32", flush=True)
    ps, g, data32, _ = _config2()
    d64 = data32.astype(F64)
    truth = _cached(("config2", "cg_iterates"), lambda: osolv.conjugate_gradients_iterates(d64, g, CG_ITERS))[50]
    ref32 = _cached(("config2", "cg", "float32"), lambda: osolv.conjugate_gradients(data32, g, 50))
    hip32 = dev.to_host(gratopy.conjugate_gradients(dev.to_device(data32, DEV), ps, 50))
    e = _three(hip32, ref32, truth)
    res = {"hip": _residual(g, hip32, d64), "oracle32": _residual(g, ref32, d64),
           "truth": _residual(g, truth, d64)}
    record_parity("config2", "cg50_f32", residual=res, **e)
    assert e["e_hip"] <= 1.5 * e["e_ref"], e
    assert res["hip"] <= 1.05 * res["truth"], res


# ---------------------------------------------------------------------------
# config 3
# ---------------------------------------------------------------------------
def _config3():
    def make():
        mix = bench.uneven_mix()
        settings = [ps for ps, _ in mix]
        rng = np.random.default_rng(11)
        datas = []
        for ps, ph in mix:
            s = dev.to_host(gratopy.forwardprojection(dev.to_device(np.asfortranarray(ph), DEV), ps))
            datas.append(np.asfortranarray(add_noise(s, 0.5, rng)))
        norms = [_gpu_norm(ps) for ps in settings]
        return settings, datas, norms

    return _cached("config3", make)


@pytest.mark.timeout(600)
def test_config3_batched_equals_per_slice(pinned_variant):
    print("config3 batched vs per-slice, 64 slices", flush=True)
    settings, datas, norms = _config3()
    sinos = [dev.to_device(d, DEV) for d in datas]
    lw_b = batch_landweber(sinos, settings, 100, norm_estimates=norms, graph=True)
    lw_p = [gratopy.landweber(f, ps, 100, norm_estimate=n) for f, ps, n in zip(sinos, settings, norms)]
    assert all(torch.equal(a, b) for a, b in zip(lw_b, lw_p))
    tv_b = batch_total_variation(sinos, settings, 1e-2, 100, norm_estimates=norms, graph=True)
    tv_p = [gratopy.total_variation(f, ps, 1e-2, 100, norm_estimate=n)
            for f, ps, n in zip(sinos, settings, norms)]
    assert all(torch.equal(a, b) for a, b in zip(tv_b, tv_p))
    record_parity("config3", "batched_vs_per_slice", slices=len(settings), landweber100_bitwise=True,
                  tv100_bitwise=True)


def _extremes(settings):
    """Indices of the smallest and the largest slice (by pixel-angle work) of
    each geometry."""
    out = {}
    for fan in (False, True):
        idx = [i for i, ps in enumerate(settings) if ps.is_fan == fan]
        work = lambda i: settings[i].img_shape[0] * settings[i].img_shape[1] * settings[i].sinogram_shape[1]
        out["fan" if fan else "radon"] = (min(idx, key=work), max(idx, key=work))
    return out


@pytest.mark.timeout(600)
def test_config3_batched_against_oracle():
    settings, datas, norms = _config3()
    sinos = [dev.to_device(d, DEV) for d in datas]
    picks = _extremes(settings)
    runs = {100: batch_landweber(sinos, settings, 100, norm_estimates=norms, graph=True),
            5: batch_landweber(sinos, settings, 5, norm_estimates=norms, graph=True)}
    for geom, (small, large) in picks.items():
        for i, n_iter in ((small, 100), (large, 5)):
            ps = settings[i]
            print(f"config3 oracle {geom} slice {i} {ps.img_shape} {ps.sinogram_shape} "
                  f"landweber({n_iter})", flush=True)
            g = osolv.OracleGeometry.from_settings(ps)
            truth = osolv.landweber(datas[i].astype(F64), g, n_iter, norm_estimate=norms[i])
            ref32 = osolv.landweber(datas[i], g, n_iter, norm_estimate=norms[i])
            e = _three(dev.to_host(runs[n_iter][i]), ref32, truth)
            record_parity("config3", f"landweber{n_iter}_{geom}_slice{i}", img=list(ps.img_shape),
                          sino=list(ps.sinogram_shape), **e)
            assert e["e_hip"] <= max(1e-5, 1.5 * e["e_ref"]), (i, e)


# ---------------------------------------------------------------------------
# config 5
# ---------------------------------------------------------------------------
def _settings5(angles=1440):
    fan, kw = bench.projection_kwargs("fanbeam-2048x32")
    return gratopy.ProjectionSettings(DEV, gratopy.FANBEAM, (2048, 2048), angles, 2560, **kw)


@pytest.mark.timeout(600)
def test_config5_tv_batched_equals_plain_loop(pinned_variant):
    print("config5 TV(10) batched+graph vs plain loop, 32 x 2048^2", flush=True)
    ps = _settings5()
    img = dev.to_device(bench.make_stack(32, 2048, 30), DEV, order="F")
    sino = gratopy.forwardprojection(img, ps)
    del img
    f = dev.to_device(np.asfortranarray(add_noise(dev.to_host(sino), 1.0, np.random.default_rng(7))), DEV)
    del sino
    norm = _gpu_norm(ps)
    ub = gratopy.total_variation(f, ps, 5e-3, 10, slice_thickness=0.0, norm_estimate=norm, graph=True)
    same = all(torch.equal(ub[:, :, z], gratopy.total_variation(f[:, :, z], ps, 5e-3, 10, norm_estimate=norm))
               for z in range(32))
    record_parity("config5", "tv10_batched_graph_vs_plain", slices=32, bitwise=bool(same))
    assert same


@pytest.mark.timeout(600)
@pytest.mark.parametrize("data", ["phantom", "random"])
def test_config5_full_size_backprojection_against_oracle(data):
    """All 1440 angles of one 2048^2 slice.  "phantom": the sinogram of a
    30-ellipse phantom (smooth along the detector, as measured data are);
    "random": independent uniform values per bin, where every error of the
    detector coordinate turns fully into an error of the interpolated value.
    The float32 tables alone put ~1e-5 of error into the reference's own
    float32 result on random data (e_ref)."""
    print(f"config5 backprojection, 2048^2, all 1440 angles, {data} sinogram", flush=True)
    ps = _settings5()
    g = osolv.OracleGeometry.from_settings(ps)
    if data == "random":
        sino32 = np.asfortranarray(np.random.default_rng(21).random(ps.sinogram_shape).astype(F32))
    else:
        img = dev.to_device(np.asfortranarray(bench.make_stack(1, 2048, 30, dtype=F64)[:, :, 0]), DEV)
        sino32 = np.asfortranarray(dev.to_host(gratopy.forwardprojection(img, ps)).astype(F32))
    truth = g.back(sino32.astype(F64))
    ref32 = g.back(sino32)
    hip32 = dev.to_host(gratopy.backprojection(dev.to_device(sino32, DEV), ps))
    hip64 = dev.to_host(gratopy.backprojection(dev.to_device(sino32.astype(F64), DEV), ps))
    e = _three(hip32, ref32, truth)
    e64 = rel_err(hip64, truth)
    record_parity("config5", f"back_2048_1440_{data}", f32=e, f64=e64)
    assert e["e_hip"] <= max(1e-5, 1.25 * e["e_ref"]), e
    assert e64 <= 1e-12, e64


@pytest.mark.timeout(600)
def test_config5_full_size_forward_float64():
    print("config5 forward float64, 2048^2, every 16th of 1440 angles", flush=True)
    ps = _settings5()
    img = np.asfortranarray(bench.make_stack(1, 2048, 30, dtype=F64)[:, :, 0])
    got = dev.to_host(gratopy.forwardprojection(dev.to_device(img, DEV), ps))[:, ::16]
    sub = _settings5(angles=np.asarray(ps.angles)[::16])
    truth = osolv.OracleGeometry.from_settings(sub).fwd(img)
    e64 = rel_err(got, truth)
    record_parity("config5", "fwd_2048_f64_every16th", f64=e64)
    assert e64 <= 1e-12, e64
