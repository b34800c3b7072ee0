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
    # name: (Nz, N, Na, Ns, RE, R, image_width, detector_width, ellipses)
    "fanbeam-2048x32": (32, 2048, 1440, 2560, 2048.0, 4096.0, 2048.0, 4096.0, 30),
    "fanbeam-512": (1, 512, 360, 768, 400.0, 800.0, 512.0, 1024.0, 15),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="fanbeam-2048x32", choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-angles", type=int, default=90,
                   help="angles of the bounded CPU-baseline sample (one slice)")
    p.add_argument("--fwd-group", type=int, default=None, help="detector bins per lane (forward)")
    p.add_argument("--back-rows", type=int, default=None, help="pixels per lane (backprojection)")
    return p.parse_args()


def make_stack(Nz, N, n_ellipses, dtype=np.float32):
    from gratopy_amd.phantom import random_ellipses

    stack = np.empty((N, N, Nz), dtype=dtype, order="F")
    for z in range(Nz):
        stack[:, :, z] = random_ellipses(N, n_ellipses, seed=1000 + z, dtype=dtype)
    return stack


def cpu_baseline(ps_kwargs, N, n_angles, n_det, slice_img):
    """Oracle (C restatement of the reference kernels, OpenMP) on one slice
    with a subset of the angles; returns pixel-angle updates per second."""
    import gratopy_amd as gratopy
    import oracle

    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM, (N, N), n_angles, n_det, **ps_kwargs)
    t = ps.struct
    dt = np.dtype("float32")
    img = np.asfortranarray(slice_img)
    oracle.lib()
    t0 = time.perf_counter()
    sino = oracle.fan_fwd(img, t["ofs_dict"][dt], t["sdpd_dict"][dt], t["geo_dict"][dt], n_det, n_angles)
    t1 = time.perf_counter()
    oracle.fan_back(sino, t["ofs_dict"][dt], t["sdpd_dict"][dt], t["geo_dict"][dt], (N, N))
    t2 = time.perf_counter()
    work = N * N * n_angles
    return work / (t2 - t0), (t1 - t0), (t2 - t1)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (slices shard "
                                   "across GPUs without communication, see DESIGN.md)"}))
        sys.exit(2)

    import torch

    import gratopy_amd as gratopy
    from gratopy_amd import _device as dev
    from gratopy_amd import _hip
    from gratopy_amd import build as hip_build

    hip_build.ensure_built()
    lib = _hip.load_library()
    if args.fwd_group is not None:
        _hip.check(lib.gt_tune_set(0, args.fwd_group), "gt_tune_set")
    if args.back_rows is not None:
        _hip.check(lib.gt_tune_set(1, args.back_rows), "gt_tune_set")
    Nz, N, Na, Ns, RE, R, iw, dw, n_ell = WORKLOADS[args.workload]
    ps_kwargs = dict(R=R, RE=RE, image_width=iw, detector_width=dw)
    device = "cuda:0"
    ps = gratopy.ProjectionSettings(device, gratopy.FANBEAM, (N, N), Na, Ns, **ps_kwargs)

    host = make_stack(Nz, N, n_ell)
    img = dev.to_device(host, device, order="F")
    sino = dev.empty((Ns, Na, Nz), np.float32, order="F", device=device)
    back = dev.empty((N, N, Nz), np.float32, order="F", device=device)
    stream = torch.cuda.current_stream()

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

    # dominant launch: the slower of the two directions; algorithmic bytes =
    # 2 interpolation taps x 4 B per pixel-angle per direction
    fwd_med, back_med = float(np.median(t_fwd)), float(np.median(t_back))
    dom_name, dom_t = ("fan forward", fwd_med) if fwd_med >= back_med else ("fan backprojection", back_med)
    algo_bytes = 8.0 * work
    achieved = algo_bytes / dom_t

    cpu = None
    if not args.no_cpu_baseline:
        n_ang = min(args.cpu_angles, Na)
        rate, tf, tb = cpu_baseline(ps_kwargs, N, n_ang, Ns, host[:, :, 0])
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": rate / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
               "sample": f"oracle fan fwd+back, 1 slice {N}x{N}, {n_ang} of {Na} angles, {Ns} det "
                         f"(fwd {tf:.2f}s, back {tb:.2f}s)"}

    out = {
        "metric": "fwd+back GUPS (pixel*angle/s) fanbeam 2048^2",
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
        "tuning": {"fwd_group": lib.gt_tune_get(0), "back_rows": lib.gt_tune_get(1)},
        "fwd_ms": round(fwd_med * 1e3, 3),
        "back_ms": round(back_med * 1e3, 3),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
