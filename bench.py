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

WORKLOADS = {
    # name: (Nz, N, Na, Ns, RE, R, image_width, detector_width, ellipses); RE and R are
    # None for parallel beam (angles in [0, pi), detector spanning the image diagonal)
    "fanbeam-2048x32": (32, 2048, 1440, 2560, 2048.0, 4096.0, 2048.0, 4096.0, 30),
    "fanbeam-512": (1, 512, 360, 768, 400.0, 800.0, 512.0, 1024.0, 15),
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
    p.add_argument("--cpu-angles", type=int, default=90,
                   help="angles of the bounded CPU-baseline sample (one slice)")
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
    p.add_argument("--auto", action="store_true",
                   help="no autotuning: the library's automatic launch shapes (its default)")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the batched-vs-plain solver measurements")
    p.add_argument("--uneven-only", action="store_true",
                   help="only the batched-vs-plain Landweber and TV measurements on the 64 uneven "
                        "slices (with launch-variant flags: those variants, no batch autotuning)")
    p.add_argument("--tv-iterations", type=int, default=6,
                   help="TV iterations timed per path on the 2048^2 stack (rate extrapolated to 500)")
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


def bench_tv_stack(ps, img, n_iter, nominal=500):
    """TV (mu=5e-3, 1% noise, slice_thickness 0) on the whole stack: batched
    (all slices per launch) and graph-captured vs the plain per-slice loop
    (one slice per call, eager launches).  Both paths run n_iter iterations;
    slices/s is extrapolated to `nominal` iterations per slice."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev

    Nz = img.shape[2]
    sino = gratopy.forwardprojection(img, ps)
    host = add_noise(dev.to_host(sino), 1.0, np.random.default_rng(7))
    f = dev.to_device(np.asfortranarray(host), img.device)
    norm = gratopy.normest(ps)
    mu = 5e-3
    slices = [f[:, :, z] for z in range(Nz)]  # Fortran-ordered 2-D views
    run_batched = lambda: gratopy.total_variation(f, ps, mu, n_iter, slice_thickness=0.0,
                                                  norm_estimate=norm, graph=True)
    run_plain = lambda: [gratopy.total_variation(s, ps, mu, n_iter, norm_estimate=norm) for s in slices]
    run_batched()  # capture / warm-up
    tb, ub = _timed(run_batched)
    run_plain()
    tp, up = _timed(run_plain)
    diff = max(float((ub[:, :, z] - up[z]).abs().max()) for z in range(Nz))
    rate_b = Nz / (tb / n_iter * nominal)
    rate_p = Nz / (tp / n_iter * nominal)
    Nx, Ny, Nz = (int(n) for n in img.shape)
    stack = f"{Nz} x {Nx}^2 {'fanbeam' if ps.is_fan else 'radon'} stack"
    return {"metric": f"TV(500) slices/s, {stack}", "batched_graph": round(rate_b, 4),
            "plain_per_slice": round(rate_p, 4), "speedup": round(rate_b / rate_p, 3),
            "iterations_timed": n_iter, "ms_per_iteration_batched": round(tb / n_iter * 1e3, 2),
            "ms_per_iteration_plain": round(tp / n_iter * 1e3, 2),
            "max_abs_diff_batched_vs_plain": diff}


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


def bench_uneven_landweber(problem, n_iter=100):
    """Landweber(100) on the 64-slice uneven mix: one batched plan, graph
    captured, vs the plain per-slice loop of landweber calls."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd.batch import batch_landweber

    settings, sinos, norms, work, tuned = problem
    run_b = lambda: batch_landweber(sinos, settings, n_iter, norm_estimates=norms, graph=True)
    run_p = lambda: [gratopy.landweber(f, ps, n_iter, norm_estimate=n)
                     for f, ps, n in zip(sinos, settings, norms)]
    run_b()
    tb, ub = _timed(run_b)
    run_p()
    tp, up = _timed(run_p)
    same = all(torch.equal(a, b) for a, b in zip(ub, up))
    n = len(settings)
    return {"metric": "Landweber(100) slices/s, 64 uneven slices (config of 128..1024 px, 60..720 angles)",
            "batched_graph": round(n / tb, 3), "plain_per_slice": round(n / tp, 3),
            "speedup": round(tp / tb, 3), "fwd+back GUPS batched": round(work * n_iter / tb / 1e9, 2),
            "bitwise_equal": same,
            "tuning": {k: tuned[k] for k in ("fwd_group", "fwd_bands", "fwd_chunk", "back_rows",
                                             "back_stage", "back_tile", "source")}}


def bench_uneven_tv(problem, n_iter=100, mu=1e-2):
    """TV(100) (mu = 1e-2) on the 64-slice uneven mix: one batched plan, graph
    captured, vs the plain per-slice loop of total_variation calls (the
    reference's structure: seven launches per iteration with host scalars)."""
    import torch

    import gratopy_amd as gratopy
    from gratopy_amd.batch import batch_total_variation

    settings, sinos, norms, work, _ = problem
    run_b = lambda: batch_total_variation(sinos, settings, mu, n_iter, norm_estimates=norms, graph=True)
    run_p = lambda: [gratopy.total_variation(f, ps, mu, n_iter, norm_estimate=n)
                     for f, ps, n in zip(sinos, settings, norms)]
    run_b()
    tb, ub = _timed(run_b)
    run_p()
    tp, up = _timed(run_p)
    same = all(torch.equal(a, b) for a, b in zip(ub, up))
    n = len(settings)
    return {"metric": "TV(100) slices/s, 64 uneven slices (config of 128..1024 px, 60..720 angles)",
            "batched_graph": round(n / tb, 3), "plain_per_slice": round(n / tp, 3),
            "speedup": round(tp / tb, 3), "fwd+back GUPS batched": round(work * n_iter / tb / 1e9, 2),
            "bitwise_equal": same}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (profiles/pmc_<kernel>_<workload>.txt, line
    'hbm_bytes_per_launch: N'), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def make_stack(Nz, N, n_ellipses, dtype=np.float32):
    from gratopy_amd.phantom import random_ellipses

    stack = np.empty((N, N, Nz), dtype=dtype, order="F")
    for z in range(Nz):
        stack[:, :, z] = random_ellipses(N, n_ellipses, seed=1000 + z, dtype=dtype)
    return stack


def cpu_baseline(fan, ps_kwargs, N, n_angles, n_det, slice_img):
    """Oracle (C restatement of the reference kernels, OpenMP) on one slice
    with a subset of the angles; returns pixel-angle updates per second."""
    import gratopy_amd as gratopy
    import oracle

    geometry = gratopy.FANBEAM if fan else gratopy.RADON
    ps = gratopy.ProjectionSettings(None, geometry, (N, N), n_angles, n_det, **ps_kwargs)
    t = ps.struct
    dt = np.dtype("float32")
    img = np.asfortranarray(slice_img)
    oracle.lib()
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
    return work / (t2 - t0), (t1 - t0), (t2 - t1)


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

    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd import _hip
    from gratopy_amd import build as hip_build

    hip_build.ensure_built()
    lib = _hip.load_library()
    from gratopy_amd import autotune

    flags = {k: getattr(args, k) for k in ("fwd_group", "back_rows", "fwd_bands", "back_bands",
                                           "back_stage", "back_tile", "fwd_chunk")}
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
    if args.auto and not fixed:
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
    algo_bytes = 8.0 * work
    achieved = algo_bytes / dom_t

    cpu = None
    if not args.no_cpu_baseline:
        n_ang = min(args.cpu_angles, Na)
        rate, tf, tb = cpu_baseline(fan, ps_kwargs, N, n_ang, Ns, host[:, :, 0])
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": rate / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
               "sample": f"oracle {'fan' if fan else 'radon'} fwd+back, 1 slice {N}x{N}, "
                         f"{n_ang} of {Na} angles, {Ns} det (fwd {tf:.2f}s, back {tb:.2f}s)"}

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
        "tuning": {"fwd_group": lib.gt_tune_get(0), "fwd_bands": lib.gt_tune_get(2),
                   "back_rows": lib.gt_tune_get(1), "back_bands": lib.gt_tune_get(3),
                   "back_stage": lib.gt_tune_get(4), "back_tile": lib.gt_tune_get(5),
                   "fwd_chunk": lib.gt_tune_get(6), "automatic": lib.gt_tune_get(7),
                   "source": tuning_source,
                   "trial_ms": None if tuned is None else {"fwd": tuned["fwd_ms"], "back": tuned["back_ms"]}},
        "fwd_ms": round(fwd_med * 1e3, 3),
        "back_ms": round(back_med * 1e3, 3),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(args.workload, dom_name)},
        "cpu_baseline": cpu,
    }
    if not args.no_batched:
        del sino, back
        torch.cuda.empty_cache()
        out["batched"] = {"tv_stack": bench_tv_stack(ps, img, args.tv_iterations)}
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
