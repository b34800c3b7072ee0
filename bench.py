"""Benchmark of the projection hot path on one MI355X.

Workload (default): a stack of 32 float32 slices of 2048 x 2048 random-
ellipse phantoms (30 ellipses each, one seed per slice) in Fortran order,
fanbeam geometry with 1440 angles in [0, 2pi), 2560 detectors, RE = 2048,
R = 4096, image width 2048 and detector width 4096 (pixel units).  One step
is one forward projection plus one backprojection of the whole stack.

Metric: fwd+back GUPS = Nx*Ny*Na*Nz / (t_fwd + t_back) / 1e9 with inputs
resident in HBM.  Prints one JSON line (see ``main``).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X
# VALU issue peak of MI355X in lane-operations per second: the published FP32
# vector peak, 157.3 TFLOP/s, counts two flops per FMA lane-operation
# (256 CUs x 128 lanes x 2.4 GHz)
VALU_PEAK = 157.3e12 / 2

WORKLOADS = {
    # name: (Nz, N, Na, Ns, RE, R, image_width, detector_width, ellipses); RE and R are
    # None for parallel beam (angles in [0, pi), detector spanning the image diagonal)
    "fanbeam-2048x32": (32, 2048, 1440, 2560, 2048.0, 4096.0, 2048.0, 4096.0, 30),
    "fanbeam-512": (1, 512, 360, 768, 400.0, 800.0, 512.0, 1024.0, 15),
    "fanbeam-2048": (1, 2048, 1440, 2560, 2048.0, 4096.0, 2048.0, 4096.0, 30),
    "radon-2048": (1, 2048, 1440, 2560, None, None, 2048.0, 2048.0 * 2 ** 0.5, 30),
    "radon-2048x32": (32, 2048, 1440, 2560, None, None, 2048.0, 2048.0 * 2 ** 0.5, 30),
}
DEFAULT_WORKLOAD = "fanbeam-2048x32"


def projection_kwargs(workload):
    """(geometry is fanbeam, ProjectionSettings keyword arguments) of a workload."""
    _, _, _, _, RE, R, iw, dw, _ = WORKLOADS[workload]
    fan = R is not None
    kw = dict(image_width=iw, detector_width=dw)
    if fan:
        kw.update(R=R, RE=RE)
    return fan, kw


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-angles", type=int, default=None,
                   help="angles of the CPU-baseline sample (one slice; default: all)")
    p.add_argument("--unrolled", action="store_true",
                   help="only time the unrolled TV(100) forward and backward passes, 16 x 512^2 "
                        "float64, Radon and fanbeam")
    p.add_argument("--cpu-configs", action="store_true",
                   help="only time configurations 1 and 2 in full on the CPU oracle and the GPU")
    p.add_argument("--fwd-group", type=int, default=None, help="detector bins per lane (forward)")
    p.add_argument("--back-rows", type=int, default=None, help="pixels per lane (backprojection)")
    p.add_argument("--fwd-bands", type=int, default=None, help="row bands of the forward (1 or 8)")
    p.add_argument("--back-bands", type=int, default=None, help="angle bands of the backprojection")
    p.add_argument("--back-stage", type=int, default=None,
                   help="angles per LDS staging chunk of the backprojection (0 = unstaged, 8, 16)")
    p.add_argument("--back-tile", type=int, default=None,
                   help="tile width (16 or 32 pixels) of the staged backprojection")
    p.add_argument("--fwd-chunk", type=int, default=None,
                   help="forward strip reads: 1 pixel or 4 (16 bytes) per load")
    p.add_argument("--fwd-compact", type=int, default=None,
                   help="1: banded forward launches only the bin groups that meet each band")
    p.add_argument("--tune", action="store_true",
                   help="time every launch variant on the workload first and keep the fastest "
                        "(default: the library's automatic launch shapes, which on the default "
                        "workload are the autotuner's choice)")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the batched-vs-plain solver measurements")
    p.add_argument("--uneven-only", action="store_true",
                   help="only the batched-vs-plain Landweber and TV measurements on the 64 uneven "
                        "slices (with launch-variant flags: those variants, no batch autotuning)")
    p.add_argument("--tv-iterations", type=int, default=500,
                   help="TV iterations of the batched + graph and the plain per-slice paths on the stack")
    p.add_argument("--tv-ref-iterations", type=int, default=100,
                   help="TV iterations of the reference-structured per-slice loop on the stack")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run an in-place scale of 1 GiB (reads and writes 1 GiB) three times, "
                        "to calibrate the FETCH_SIZE / WRITE_SIZE counters")
    return p.parse_args()


def calibrate_copy():
    """x *= 0.5 over 2^28 float32 values (1 GiB read + 1 GiB written per launch)."""
    import torch

    from gratopy_amd import _hip

    _hip.load_library()
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    s = torch.cuda.current_stream()
    for _ in range(3):
        _hip.call("gt_scale", s.cuda_stream, _hip.dtype_code(x.dtype), x.data_ptr(), 0.5, n)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "gt_scale in place", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "launches": 3}), flush=True)


def add_noise(sino, percent, rng):
    """f + (p/100) * (||f||_2 / sqrt(n)) * N(0, 1) on the host."""
    scale = percent / 100.0 * np.linalg.norm(sino) / np.sqrt(sino.size)
    return (sino + scale * rng.standard_normal(sino.shape)).astype(sino.dtype)


def _timed(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def tv_reference_structured(f, ps, mu, n_iter, norm_estimate):
    """The TV loop in the reference's structure (gratopy/gratopy.py:2744-2759):
    seven launches per iteration through the one-to-one entry points
    (update_v, update_NormV, forward projection, update_lambda,
    backprojection, update_u, update_extra), scalars from the host, no
    fusion, no graph.  2-D slice, slice_thickness 0.  Returns u."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd import _hip

    lib = _hip.load_library()
    dtype = dev.to_np_dtype(f.dtype)
    Nx, Ny = ps.img_shape
    Ns, Na = ps.sinogram_shape
    dc, oc = _hip.dtype_code(dtype), dev.order_code(dev.order_of(f))
    U = dev.zeros(ps.img_shape, dtype, order="F", device=f.device)
    U_ = dev.zeros(ps.img_shape, dtype, order="F", device=f.device)
    V = torch.zeros(4 * Nx * Ny, dtype=f.dtype, device=f.device)
    normV = dev.zeros(ps.img_shape, dtype, order="F", device=f.device)
    Lamb = dev.zeros(ps.sinogram_shape, dtype, order="F", device=f.device)
    KU = dev.zeros(ps.sinogram_shape, dtype, order="F", device=f.device)
    KSTAR = dev.zeros(ps.img_shape, dtype, order="F", device=f.device)
    sigma = 1.0 / np.sqrt(13) * 10.0
    tau = 1.0 / np.sqrt(13) / 10.0
    mu = mu * norm_estimate**2 * ps.delta_x
    mu = mu / (sigma + mu)
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(n_iter):
        _hip.check(lib.gt_tv_update_v(st, dc, V.data_ptr(), U_.data_ptr(), sigma, 0.0, Nx, Ny, 1, oc),
                   "gt_tv_update_v")
        _hip.check(lib.gt_tv_update_normv(st, dc, V.data_ptr(), normV.data_ptr(), Nx, Ny, 1, oc),
                   "gt_tv_update_normv")
        gratopy.forwardprojection(U_, ps, sino=KU)
        _hip.check(lib.gt_tv_update_lambda(st, dc, Lamb.data_ptr(), KU.data_ptr(), f.data_ptr(),
                                           sigma / norm_estimate, mu, Ns, Na, 1, oc),
                   "gt_tv_update_lambda")
        gratopy.backprojection(Lamb, ps, img=KSTAR)
        _hip.check(lib.gt_tv_update_u(st, dc, U_.data_ptr(), U.data_ptr(), V.data_ptr(), KSTAR.data_ptr(),
                                      tau, 1.0 / norm_estimate, 0.0, Nx, Ny, 1, oc), "gt_tv_update_u")
        _hip.check(lib.gt_tv_update_extra(st, dc, U_.data_ptr(), U.data_ptr(), Nx * Ny),
                   "gt_tv_update_extra")
        U, U_ = U_, U
    return U


def bench_tv_stack(ps, img, n_iter, n_ref):
    """TV (mu=5e-3, 1% noise, slice_thickness 0) on the whole stack, three
    ways: batched (all slices per launch) and graph-captured, and the plain
    per-slice loop of total_variation (fused dual and primal updates, eager
    launches), both for n_iter iterations; the reference-structured per-slice
    loop (seven launches per iteration, host scalars, no graph) for n_ref
    iterations (its slices/s is quoted per n_iter iterations)."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev

    Nz = img.shape[2]
    sino = gratopy.forwardprojection(img, ps)
    host = add_noise(dev.to_host(sino), 1.0, np.random.default_rng(7))
    del sino
    f = dev.to_device(np.asfortranarray(host), img.device)
    del host
    norm = gratopy.normest(ps)
    mu = 5e-3
    slices = [f[:, :, z] for z in range(Nz)]  # Fortran-ordered 2-D views
    run_batched = lambda: gratopy.total_variation(f, ps, mu, n_iter, slice_thickness=0.0,
                                                  norm_estimate=norm, graph=True)
    run_plain = lambda: [gratopy.total_variation(s, ps, mu, n_iter, norm_estimate=norm) for s in slices]
    run_ref = lambda: [tv_reference_structured(s, ps, mu, n_ref, norm) for s in slices]
    gratopy.total_variation(f, ps, mu, 2, slice_thickness=0.0, norm_estimate=norm, graph=True)  # warm-up
    tb, ub = _timed(run_batched)
    print(f"tv stack: batched + graph {tb:.1f} s", file=sys.stderr, flush=True)
    tp, up = _timed(run_plain)
    print(f"tv stack: plain loop {tp:.1f} s", file=sys.stderr, flush=True)
    same_plain = all(torch.equal(ub[:, :, z], up[z]) for z in range(Nz))
    diff_plain = max(float((ub[:, :, z] - up[z]).abs().max()) for z in range(Nz))
    del up
    tr, ur = _timed(run_ref)
    print(f"tv stack: reference-structured loop {tr:.1f} s", file=sys.stderr, flush=True)
    del ub
    # the reference-structured loop against the batched path at the same iteration count
    ub = gratopy.total_variation(f, ps, mu, n_ref, slice_thickness=0.0, norm_estimate=norm, graph=True)
    diff_ref = max(float((ub[:, :, z] - ur[z]).abs().max()) for z in range(Nz))
    Nx, Ny, Nz = (int(n) for n in img.shape)
    stack = f"{Nz} x {Nx}^2 {'fanbeam' if ps.is_fan else 'radon'} stack"
    return {"metric": f"TV({n_iter}) slices/s, {stack}", "batched_graph": round(Nz / tb, 4),
            "plain_per_slice": round(Nz / tp, 4),
            "reference_structured": round(Nz / (tr / n_ref * n_iter), 4),
            "speedup_vs_plain": round(tp / tb, 3),
            "speedup_vs_reference_structured": round((tr / n_ref) / (tb / n_iter), 3),
            "iterations": n_iter, "reference_structured_iterations_timed": n_ref,
            "ms_per_iteration_batched": round(tb / n_iter * 1e3, 2),
            "ms_per_iteration_plain": round(tp / n_iter * 1e3, 2),
            "ms_per_iteration_reference_structured": round(tr / n_ref * 1e3, 2),
            "bitwise_batched_vs_plain": same_plain,
            "max_abs_diff_batched_vs_plain": diff_plain,
            "note": "each path takes its own automatic launch shapes (the per-slice calls sum "
                    "a 2048^2 slice's backprojection over 8 angle bands, the launches over the "
                    "whole stack do not), so the sums are ordered differently; under one pinned "
                    "shape the two are bitwise equal (tests/test_gpu_configs.py)",
            f"max_abs_diff_batched_vs_reference_structured_at_{n_ref}": diff_ref}


def uneven_mix(n=64, seed=3):
    """64 slices: Nx from {128,...,1024}, angles from {60,120,360,720},
    Ns = ceil(1.5 Nx), alternating Radon / fanbeam (RE = Nx, R = 2 Nx,
    image width Nx, detector width 2 Nx), random-ellipse phantoms,
    0.5% noise."""
    import gratopy_amd as gratopy
    from gratopy_amd.phantom import random_ellipses

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.choice([128, 256, 384, 512, 768, 1024]))
        Na = int(rng.choice([60, 120, 360, 720]))
        Ns = int(np.ceil(1.5 * N))
        if i % 2 == 0:
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (N, N), Na, Ns)
        else:
            ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (N, N), Na, Ns, RE=float(N),
                                            R=2.0 * N, image_width=float(N), detector_width=2.0 * N)
        out.append((ps, random_ellipses(N, 10, seed=500 + i, dtype=np.float32)))
    return out


def uneven_problem(tune=True):
    """The 64-slice uneven mix with noisy sinograms, norm estimates, the
    pixel-angle work of one fwd+back over all slices, and the launch variant
    (the batch autotuner's choice, installed, unless tune is False)."""
    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd import autotune

    mix = uneven_mix()
    settings = [ps for ps, _ in mix]
    rng = np.random.default_rng(11)
    sinos = []
    for ps, ph in mix:
        s = dev.to_host(gratopy.forwardprojection(dev.to_device(np.asfortranarray(ph), "cuda:0"), ps))
        sinos.append(dev.to_device(np.asfortranarray(add_noise(s, 0.5, rng)), "cuda:0"))
    norms = [gratopy.normest(ps) for ps in settings]
    work = sum(ps.img_shape[0] * ps.img_shape[1] * ps.sinogram_shape[1] for ps in settings)
    if not tune:
        return settings, sinos, norms, work, {**autotune.current_variant(), "source": "fixed"}
    scratch_i = [dev.empty(ps.img_shape, np.float32, order="F", device="cuda:0") for ps in settings]
    scratch_s = [dev.empty(ps.sinogram_shape, np.float32, order="F", device="cuda:0") for ps in settings]
    tuned = autotune.tune_batch(settings, scratch_i, scratch_s)
    return settings, sinos, norms, work, {**tuned, "source": "autotuner"}


class _automatic_shapes:
    """Within the block every call picks its own automatic launch shape; the
    knobs in force before (the batch autotuner's choice) come back after."""

    def __enter__(self):
        from gratopy_amd import autotune

        self.before = autotune.current_variant()
        autotune.set_variant(automatic=1)

    def __exit__(self, *exc):
        from gratopy_amd import autotune

        autotune.set_variant(**self.before)


def bench_uneven_landweber(problem, n_iter=100):
    """Landweber(100) on the 64-slice uneven mix: one batched plan, graph
    captured (the batch autotuner's launch variant), vs the plain per-slice
    loop of landweber calls, each call with its own automatic launch shape.
    The plain loop is also run once under the batch's variant, where the two
    must agree bit for bit."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd.batch import batch_landweber

    settings, sinos, norms, work, tuned = problem
    run_b = lambda: batch_landweber(sinos, settings, n_iter, norm_estimates=norms, graph=True)
    run_p = lambda: [gratopy.landweber(f, ps, n_iter, norm_estimate=n)
                     for f, ps, n in zip(sinos, settings, norms)]
    run_b()
    tb, ub = _timed(run_b)
    same = all(torch.equal(a, b) for a, b in zip(ub, run_p()))
    with _automatic_shapes():
        run_p()
        tp, _ = _timed(run_p)
    n = len(settings)
    return {"metric": "Landweber(100) slices/s, 64 uneven slices (config of 128..1024 px, 60..720 angles)",
            "batched_graph": round(n / tb, 3), "plain_per_slice": round(n / tp, 3),
            "speedup": round(tp / tb, 3), "fwd+back GUPS batched": round(work * n_iter / tb / 1e9, 2),
            "bitwise_equal_under_one_variant": same,
            "tuning": {"batched": {k: tuned[k] for k in ("fwd_group", "fwd_bands", "fwd_chunk",
                                                          "back_rows", "back_stage", "back_tile",
                                                          "source")},
                       "plain_per_slice": "automatic per call"}}


def bench_uneven_tv(problem, n_iter=100, mu=1e-2):
    """TV(100) (mu = 1e-2) on the 64-slice uneven mix: one batched plan, graph
    captured (the batch autotuner's launch variant), vs two per-slice loops,
    each call with its own automatic launch shape: total_variation (fused dual
    and primal updates, five launches per iteration, eager) and the
    reference-structured loop (seven launches per iteration through the
    one-to-one entry points, host scalars, no graph).  total_variation is also
    run once under the batch's variant, where it must agree bit for bit."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd.batch import batch_total_variation

    settings, sinos, norms, work, _ = problem
    run_b = lambda: batch_total_variation(sinos, settings, mu, n_iter, norm_estimates=norms, graph=True)
    run_p = lambda: [gratopy.total_variation(f, ps, mu, n_iter, norm_estimate=n)
                     for f, ps, n in zip(sinos, settings, norms)]
    run_r = lambda: [tv_reference_structured(f, ps, mu, n_iter, n) for f, ps, n in zip(sinos, settings, norms)]
    run_b()
    tb, ub = _timed(run_b)
    same = all(torch.equal(a, b) for a, b in zip(ub, run_p()))
    with _automatic_shapes():
        run_p()
        tp, _ = _timed(run_p)
        run_r()
        tr, ur = _timed(run_r)
    diff_r = max(float((a - b).abs().max() / max(float(b.abs().max()), 1e-30)) for a, b in zip(ub, ur))
    n = len(settings)
    return {"metric": "TV(100) slices/s, 64 uneven slices (config of 128..1024 px, 60..720 angles)",
            "batched_graph": round(n / tb, 3), "plain_per_slice": round(n / tp, 3),
            "reference_structured": round(n / tr, 3),
            "speedup": round(tp / tb, 3), "speedup_vs_reference_structured": round(tr / tb, 3),
            "fwd+back GUPS batched": round(work * n_iter / tb / 1e9, 2),
            "bitwise_equal_under_one_variant": same,
            "max_rel_diff_vs_reference_structured": diff_r}


def profile_value(name, key):
    """The number on the line 'key: N' of the committed rocprofv3 summary
    profiles/<name>, or None when the file or the line is absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of the `kernel` direction on `workload` from the
    committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary, or None."""
    return profile_value(f"pmc_{kernel}_{workload}.txt", "hbm_bytes_per_launch")


def roofline(workload, direction, seconds, work, elem=4):
    """Roofline of one direction of the projection (all its kernels, `seconds`
    per launch, measured live).  The bound is VALU issue: lane-operations
    (SQ_INSTS_VALU x 64, from the committed rocprofv3 --pmc summary) per
    second against the FP32 vector peak.  Also reported: HBM traffic from the
    FETCH_SIZE / WRITE_SIZE counters against 8 TB/s, and the algorithmic
    figure of two interpolation taps per pixel-angle, whose bytes are mostly
    served from LDS and the caches, not HBM."""
    valu = profile_value(f"pmc_sq_{workload}.txt", f"valu_wave_insts_per_launch {direction}")
    traffic = measured_traffic(workload, direction)
    algo = 2.0 * elem * work
    out = {}
    if valu is not None:
        lane_ops = valu * 64 / seconds
        out = {"bound": "valu", "kernel": direction, "achieved": round(lane_ops / 1e12, 3),
               "peak": round(VALU_PEAK / 1e12, 3), "unit": "T lane-op/s",
               "frac": round(lane_ops / VALU_PEAK, 4), "traffic": traffic,
               "valu_lane_ops_per_pixel_angle": round(valu * 64 / work, 2)}
    elif traffic is not None:
        out = {"bound": "hbm", "kernel": direction, "achieved": round(traffic / seconds / 1e9, 2),
               "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(traffic / seconds / HBM_PEAK, 4),
               "traffic": traffic}
    else:
        out = {"bound": None, "kernel": direction, "traffic": None,
               "note": f"no committed counter summary for {workload}"}
    if traffic is not None:
        out["hbm"] = {"achieved": round(traffic / seconds / 1e9, 2), "peak": HBM_PEAK / 1e9,
                      "unit": "GB/s", "frac": round(traffic / seconds / HBM_PEAK, 4)}
    out["algorithmic"] = {"bytes_per_launch": algo, "achieved": round(algo / seconds / 1e9, 1),
                          "unit": "GB/s", "of_hbm_peak": round(algo / seconds / HBM_PEAK, 3),
                          "note": "2 taps x element size per pixel-angle; taps come from LDS/L1/L2"}
    return out


def launch_shapes(lib, fan, N, Ns, Na, Nz):
    """The launch shapes the library takes for the stack (gt_projection_shape)."""
    import ctypes

    out = {}
    for adjoint, name, keys in ((0, "forward", ("bins_per_lane", "row_bands", "pixels_per_load",
                                                "compacted")),
                                (1, "backprojection", ("pixels_per_lane", "angle_bands",
                                                       "staged_angles", "tile_width"))):
        shape = (ctypes.c_int * 4)()
        if lib.gt_projection_shape(int(fan), adjoint, N, N, Ns, Na, Nz, shape) == 0:
            out[name] = dict(zip(keys, list(shape)))
    return out


def make_stack(Nz, N, n_ellipses, dtype=np.float32):
    from gratopy_amd.phantom import random_ellipses

    stack = np.empty((N, N, Nz), dtype=dtype, order="F")
    for z in range(Nz):
        stack[:, :, z] = random_ellipses(N, n_ellipses, seed=1000 + z, dtype=dtype)
    return stack


def cpu_baseline(fan, ps_kwargs, N, n_angles, n_det, slice_img):
    """The oracle's timing build (the reference kernels restated in C,
    OpenMP over output elements, FMA contraction allowed, AVX-512 where the
    host has it) on one slice of the stack with all its angles; returns
    pixel-angle updates per second (the rate of the whole stack, whose slices
    are independent) and the two directions' times."""
    import gratopy_amd as gratopy
    import oracle

    geometry = gratopy.FANBEAM if fan else gratopy.RADON
    ps = gratopy.ProjectionSettings(None, geometry, (N, N), n_angles, n_det, **ps_kwargs)
    t = ps.struct
    dt = np.dtype("float32")
    img = np.asfortranarray(slice_img)
    with oracle.timing_build() as build:
        t0 = time.perf_counter()
        if fan:
            sino = oracle.fan_fwd(img, t["ofs_dict"][dt], t["sdpd_dict"][dt], t["geo_dict"][dt], n_det,
                                  n_angles)
        else:
            sino = oracle.radon_fwd(img, t["ofs_dict"][dt], t["geo_dict"][dt], n_det, n_angles)
        t1 = time.perf_counter()
        if fan:
            oracle.fan_back(sino, t["ofs_dict"][dt], t["sdpd_dict"][dt], t["geo_dict"][dt], (N, N))
        else:
            oracle.radon_back(sino, t["ofs_dict"][dt], t["geo_dict"][dt], (N, N))
        t2 = time.perf_counter()
    work = N * N * n_angles
    return work / (t2 - t0), (t1 - t0), (t2 - t1), build.variant


def cpu_threads():
    """OpenMP threads the oracle runs with."""
    n = os.environ.get("OMP_NUM_THREADS")
    return int(n) if n and n.isdigit() else (os.cpu_count() or 1)


def cpu_configs():
    """Configurations 1 and 2 in full, timed on the CPU (the oracle's timing
    build under the solver loops of oracle/solvers.py) and on the GPU, same
    inputs and injected norm estimate; one JSON object."""
    import torch

    import gratopy_amd as gratopy
    import oracle
    from gratopy_amd import _device as dev
    from gratopy_amd.phantom import random_ellipses
    from oracle import solvers as osolv

    out = {"cores": cpu_threads()}

    def both(label, cpu_fn, gpu_fn):
        with oracle.timing_build() as build:
            t0 = time.perf_counter()
            cpu_fn()
            tc = time.perf_counter() - t0
        gpu_fn()  # warm-up
        tg, _ = _timed(gpu_fn)
        out[label] = {"cpu_s": round(tc, 3), "gpu_s": round(tg, 5), "gpu_over_cpu": round(tc / tg, 1),
                      "cpu_build": build.variant}
        print(f"{label}: cpu {tc:.2f} s, gpu {tg * 1e3:.2f} ms", file=sys.stderr, flush=True)

    # configuration 1: Radon 256^2, 180 angles, 256 detectors, float32
    ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (256, 256), 180, 256)
    g = osolv.OracleGeometry.from_settings(ps)
    img = np.asfortranarray(random_ellipses(256, 10, seed=101, dtype=np.float32))
    sino = g.fwd(img)
    img_d, sino_d = dev.to_device(img, "cuda:0"), dev.to_device(sino, "cuda:0")
    np.random.seed(0)
    norm = float(gratopy.normest(ps))
    both("config1_forward", lambda: g.fwd(img), lambda: gratopy.forwardprojection(img_d, ps, sino=sino_d.clone()))
    both("config1_backprojection", lambda: g.back(sino), lambda: gratopy.backprojection(sino_d, ps))
    both("config1_landweber100", lambda: osolv.landweber(sino, g, 100, norm_estimate=norm),
         lambda: gratopy.landweber(sino_d, ps, 100, norm_estimate=norm))
    # configuration 2: fanbeam 512^2, 360 angles, 768 detectors, 1% noise
    ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (512, 512), 360, 768, RE=400.0, R=800.0,
                                    image_width=512.0, detector_width=1024.0)
    g = osolv.OracleGeometry.from_settings(ps)
    clean = g.fwd(np.asfortranarray(random_ellipses(512, 15, seed=102, dtype=np.float64)))
    data32 = np.asfortranarray(add_noise(clean, 1.0, np.random.default_rng(202)).astype(np.float32))
    np.random.seed(0)
    norm = float(gratopy.normest(ps, dtype="float64"))
    for dtype in (np.float32, np.float64):
        data = np.asfortranarray(data32.astype(dtype))
        d = dev.to_device(data, "cuda:0")
        name = np.dtype(dtype).name
        both(f"config2_cg50_{name}", lambda: osolv.conjugate_gradients(data, g, 50),
             lambda: gratopy.conjugate_gradients(d, ps, 50))
        both(f"config2_tv200_{name}", lambda: osolv.total_variation(data, g, 1e-2, 200, norm_estimate=norm),
             lambda: gratopy.total_variation(d, ps, 1e-2, 200, norm_estimate=norm))
    torch.cuda.synchronize()
    return out


def bench_unrolled(geometry="fan", nz=16, n_iter=100, checkpoint=10):
    """The differentiable unrolled TV at the differentiation configuration's
    size: 512^2, 360 angles, 768 detectors, float64, a batch of `nz` slices,
    `n_iter` iterations with checkpoints every `checkpoint`: seconds for the
    forward pass and for the backward pass (recomputation + adjoint)."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd.phantom import random_ellipses
    from gratopy_amd.tv_unrolled import tv_unrolled

    if geometry == "fan":
        ps = gratopy.ProjectionSettings("cuda:0", gratopy.FANBEAM, (512, 512), 360, 768, R=800.0,
                                        RE=400.0, detector_width=1024.0, image_width=512.0)
    else:
        ps = gratopy.ProjectionSettings("cuda:0", gratopy.RADON, (512, 512), 360, 768)
    imgs = np.stack([random_ellipses(512, 15, seed=900 + z, dtype=np.float64) for z in range(nz)], -1)
    f = gratopy.forwardprojection(dev.to_device(np.asfortranarray(imgs), "cuda:0"), ps)
    np.random.seed(0)
    norm = float(gratopy.normest(ps, dtype="float64"))
    run = lambda x: tv_unrolled(x, ps, 1e-2, n_iter, norm_estimate=norm, checkpoint=checkpoint)
    x = f.clone().requires_grad_(True)
    run(x).sum().backward()  # warm-up
    x = f.clone().requires_grad_(True)
    tf, u = _timed(lambda: run(x))
    tb, _ = _timed(lambda: u.sum().backward())
    W = 512 * 512 * 360 * nz
    return {"metric": f"unrolled TV({n_iter}) forward + backward, {nz} x 512^2 {geometry}, float64",
            "forward_s": round(tf, 3), "backward_s": round(tb, 3),
            "backward_over_forward": round(tb / tf, 2), "checkpoint": checkpoint,
            "projections_per_pass": {"forward": 2 * n_iter, "backward": 4 * n_iter},
            "fwd+back GUPS forward pass": round(2 * n_iter * W / 2 / tf / 1e9, 1)}


def main():
    args = parse()
    # the JSON line is the only output on stdout; the solvers' progress
    # messages and everything else go to stderr
    json_out = sys.stdout
    sys.stdout = sys.stderr
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (slices would "
                                   "shard across GPUs without communication; no multi-GPU path "
                                   "is built)"}), file=json_out)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.cpu_configs:
        print(json.dumps({"cpu_configs": cpu_configs()}), file=json_out, flush=True)
        return
    if args.unrolled:
        print(json.dumps({"unrolled": [bench_unrolled("fan"), bench_unrolled("radon")]}), file=json_out,
              flush=True)
        return

    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd import _hip
    from gratopy_amd import build as hip_build

    hip_build.ensure_built()
    lib = _hip.load_library()
    from gratopy_amd import autotune

    flags = {k: getattr(args, k) for k in ("fwd_group", "back_rows", "fwd_bands", "back_bands",
                                           "back_stage", "back_tile", "fwd_chunk", "fwd_compact")}
    fixed = any(v is not None for v in flags.values())
    if fixed:  # the given knobs; the others keep their values; no automatic choice
        autotune.set_variant(**flags)
    if args.uneven_only:
        problem = uneven_problem(tune=not fixed)
        out = {"batched": {"landweber_uneven": bench_uneven_landweber(problem),
                           "tv_uneven": bench_uneven_tv(problem)}}
        print(json.dumps(out), file=json_out, flush=True)
        return
    Nz, N, Na, Ns, RE, R, iw, dw, n_ell = WORKLOADS[args.workload]
    fan, ps_kwargs = projection_kwargs(args.workload)
    device = "cuda:0"
    geometry = gratopy.FANBEAM if fan else gratopy.RADON
    ps = gratopy.ProjectionSettings(device, geometry, (N, N), Na, Ns, **ps_kwargs)

    host = make_stack(Nz, N, n_ell)
    img = dev.to_device(host, device, order="F")
    sino = dev.empty((Ns, Na, Nz), np.float32, order="F", device=device)
    back = dev.empty((N, N, Nz), np.float32, order="F", device=device)
    stream = torch.cuda.current_stream()
    tuning_source = "fixed"
    tuned = None
    if not args.tune and not fixed:
        tuning_source = "automatic"
    elif not fixed:
        back.copy_(img)
        tuned = autotune.tune(ps, back, sino, repeats=2, store=False)
        tuning_source = "autotuner"
        print(f"autotuned: {tuned}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        gratopy.forwardprojection(img, ps, sino=sino)
        gratopy.backprojection(sino, ps, img=back)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        gratopy.forwardprojection(img, ps, sino=sino)
        ev[k][1].record(stream)
        gratopy.backprojection(sino, ps, img=back)
        ev[k][2].record(stream)
        print(f"step {k} enqueued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    t_fwd = np.array([ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)]) / 1e3
    t_back = np.array([ev[k][1].elapsed_time(ev[k][2]) for k in range(args.steps)]) / 1e3
    step = float(np.median(t_fwd + t_back))
    work = N * N * Na * Nz
    gups = work / step / 1e9

    # dominant launch: the slower of the two directions (all kernels of that
    # direction: image transpose or sinogram staging, the projection kernel
    # and, when banded, the ordered band reduction); algorithmic bytes =
    # 2 interpolation taps x 4 B per pixel-angle per direction
    fwd_med, back_med = float(np.median(t_fwd)), float(np.median(t_back))
    dom_name, dom_t = ("forward", fwd_med) if fwd_med >= back_med else ("backprojection", back_med)

    cpu = None
    if not args.no_cpu_baseline:
        n_ang = Na if args.cpu_angles is None else min(args.cpu_angles, Na)
        rate, tf, tb, variant = cpu_baseline(fan, ps_kwargs, N, n_ang, Ns, host[:, :, 0])
        cpu = {"value": round(rate / 1e9, 4), "unit": "GUPS", "cores": cpu_threads(), "kind": "port",
               "sample": f"oracle timing build ({variant}, OpenMP, FMA) {'fan' if fan else 'radon'} "
                         f"fwd+back of 1 of the {Nz} slices, {N}x{N}, {n_ang} of {Na} angles, {Ns} det "
                         f"(fwd {tf:.2f}s, back {tb:.2f}s); slices are independent, so the stack's "
                         f"rate is the slice's"}

    out = {
        "metric": ("fwd+back GUPS (pixel*angle/s) fanbeam 2048^2" if args.workload == DEFAULT_WORKLOAD
                   else f"fwd+back GUPS (pixel*angle/s) {'fanbeam' if fan else 'radon'} {N}^2 x {Nz}"),
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "slices": Nz, "image": [N, N], "angles": Na,
                   "detectors": Ns, "RE": RE, "R": R, "image_width": iw, "detector_width": dw,
                   "order": "F"},
        "tuning": {**launch_shapes(lib, fan, N, Ns, Na, Nz), "source": tuning_source,
                   "trial_ms": None if tuned is None else {"fwd": tuned["fwd_ms"], "back": tuned["back_ms"]}},
        "fwd_ms": round(fwd_med * 1e3, 3),
        "back_ms": round(back_med * 1e3, 3),
        "wall_s": round(wall, 3),
        "roofline": roofline(args.workload, dom_name, dom_t, work),
        "gups_target_50pct_of_hbm_model": 250.0,
        "cpu_baseline": cpu,
    }
    if not args.no_batched:
        del sino, back
        torch.cuda.empty_cache()
        out["batched"] = {"tv_stack": bench_tv_stack(ps, img, args.tv_iterations, args.tv_ref_iterations)}
        print("tv stack measured", file=sys.stderr, flush=True)
        del img
        torch.cuda.empty_cache()
        problem = uneven_problem()
        out["batched"]["landweber_uneven"] = bench_uneven_landweber(problem)
        print("uneven landweber measured", file=sys.stderr, flush=True)
        out["batched"]["tv_uneven"] = bench_uneven_tv(problem)
        print("uneven tv measured", file=sys.stderr, flush=True)
    print(json.dumps(out), file=json_out, flush=True)


if __name__ == "__main__":
    main()
