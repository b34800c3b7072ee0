"""Launch-shape autotuner for the projection kernels.

The forward kernels assign K adjacent detector bins to a lane (K in 1, 2,
4, 8; K = 1 is the per-ray kernel), optionally split the strip walk into 8
row bands (one per XCD, partial sums added in band order), and the
backprojection assigns RY pixels along y to a lane (RY in 1, 2, 4, 8),
optionally over 8 angle bands on the same pattern, or stages each block's
sinogram windows in LDS for chunks of 8 or 16 angles (tiles 16 or 32 pixels
wide).
Batched launches use K and RY only.  :func:`tune` times every variant on the
caller's own geometry and array shapes with HIP events, keeps the fastest,
and remembers it per shape key in a JSON cache
(``$GRATOPY_AMD_TUNE_CACHE``, default ``~/.cache/gratopy_amd/tune.json``).
:func:`apply` installs the cached choice for a shape before a run.  The
variants compute the same sums (to rounding), so tuning never changes
results beyond rounding.
"""

from __future__ import annotations

import json
import os
import threading

import numpy as np

from . import _device as dev
from . import _hip

FWD_GROUPS = (1, 2, 4, 8)
BACK_ROWS = (1, 2, 4, 8)
FWD_BANDS = (1, 8)
BACK_STAGES = (0, 8, 16)
BACK_TILES = (16, 32)
# forward variants: (bins per lane, row bands); bands need bin groups
FWD_VARIANTS = tuple((k, 1) for k in FWD_GROUPS) + tuple((k, 8) for k in FWD_GROUPS if k > 1)
# backprojection variants: (rows per lane, angle bands, LDS staging chunk, staged tile width)
BACK_VARIANTS = (tuple((r, 1, 0, 16) for r in BACK_ROWS) + tuple((r, 8, 0, 16) for r in BACK_ROWS)
                 + tuple((r, 1, s, t) for s in BACK_STAGES if s for t in BACK_TILES
                         for r in BACK_ROWS))
KEY_FWD, KEY_BACK, KEY_BANDS, KEY_BACK_BANDS, KEY_BACK_STAGE, KEY_BACK_TILE = 0, 1, 2, 3, 4, 5

_lock = threading.Lock()


def cache_path() -> str:
    return os.environ.get(
        "GRATOPY_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "gratopy_amd", "tune.json")
    )


def shape_key(projectionsetting, dtype, nz: int) -> str:
    ps = projectionsetting
    geom = "fan" if ps.is_fan else "radon"
    Nx, Ny = ps.img_shape
    Ns, Na = ps.sinogram_shape
    return f"{geom}:{np.dtype(dtype).name}:{Nx}x{Ny}:{Ns}x{Na}:z{nz}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def set_variant(fwd_group: int | None = None, back_rows: int | None = None,
                fwd_bands: int | None = None, back_bands: int | None = None,
                back_stage: int | None = None, back_tile: int | None = None) -> None:
    lib = _hip.load_library()
    if fwd_group is not None:
        _hip.check(lib.gt_tune_set(KEY_FWD, int(fwd_group)), "gt_tune_set")
    if back_rows is not None:
        _hip.check(lib.gt_tune_set(KEY_BACK, int(back_rows)), "gt_tune_set")
    if fwd_bands is not None:
        _hip.check(lib.gt_tune_set(KEY_BANDS, int(fwd_bands)), "gt_tune_set")
    if back_bands is not None:
        _hip.check(lib.gt_tune_set(KEY_BACK_BANDS, int(back_bands)), "gt_tune_set")
    if back_stage is not None:
        _hip.check(lib.gt_tune_set(KEY_BACK_STAGE, int(back_stage)), "gt_tune_set")
    if back_tile is not None:
        _hip.check(lib.gt_tune_set(KEY_BACK_TILE, int(back_tile)), "gt_tune_set")


def current_variant() -> dict:
    lib = _hip.load_library()
    return {"fwd_group": lib.gt_tune_get(KEY_FWD), "back_rows": lib.gt_tune_get(KEY_BACK),
            "fwd_bands": lib.gt_tune_get(KEY_BANDS), "back_bands": lib.gt_tune_get(KEY_BACK_BANDS),
            "back_stage": lib.gt_tune_get(KEY_BACK_STAGE), "back_tile": lib.gt_tune_get(KEY_BACK_TILE)}


def _fwd_key(k: int, b: int) -> str:
    return f"K{k}" if b == 1 else f"K{k}xB{b}"


def _back_key(r: int, b: int, s: int = 0, t: int = 16) -> str:
    if s:
        return f"RY{r}xS{s}T{t}"
    return f"RY{r}" if b == 1 else f"RY{r}xB{b}"


def _time(fn, repeats: int) -> float:
    torch = dev.torch
    fn()  # warm-up (also compiles nothing: kernels are precompiled)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(repeats):
        start.record()
        fn()
        stop.record()
        stop.synchronize()
        times.append(start.elapsed_time(stop))
    return float(np.median(times))


def tune(projectionsetting, img, sino, repeats: int = 3, store: bool = True) -> dict:
    """Time all variants on (img, sino) and install the fastest.

    ``img`` and ``sino`` are device arrays of the shapes and orders the caller
    will use; their contents are overwritten.
    """
    from .gratopy import backprojection, forwardprojection

    ps = projectionsetting
    fwd_times = {}
    for k, b in FWD_VARIANTS:
        set_variant(fwd_group=k, fwd_bands=b)
        fwd_times[(k, b)] = _time(lambda: forwardprojection(img, ps, sino=sino), repeats)
    back_times = {}
    for r, b, s, t in BACK_VARIANTS:
        set_variant(back_rows=r, back_bands=b, back_stage=s, back_tile=t)
        back_times[(r, b, s, t)] = _time(lambda: backprojection(sino, ps, img=img), repeats)
    k, b = min(fwd_times, key=fwd_times.get)
    r, rb, rs, rt = min(back_times, key=back_times.get)
    best = {"fwd_group": k, "fwd_bands": b, "back_rows": r, "back_bands": rb, "back_stage": rs,
            "back_tile": rt,
            "fwd_ms": {_fwd_key(*v): t for v, t in fwd_times.items()},
            "back_ms": {_back_key(*v): t for v, t in back_times.items()}}
    set_variant(best["fwd_group"], best["back_rows"], best["fwd_bands"], best["back_bands"],
                best["back_stage"], best["back_tile"])
    if store:
        nz = int(img.shape[2]) if img.dim() > 2 else 1
        with _lock:
            table = _load()
            table[shape_key(ps, dev.to_np_dtype(img.dtype), nz)] = {
                "fwd_group": best["fwd_group"], "fwd_bands": best["fwd_bands"],
                "back_rows": best["back_rows"], "back_bands": best["back_bands"],
                "back_stage": best["back_stage"], "back_tile": best["back_tile"]}
            _store(table)
    return best


def tune_batch(settings, imgs, sinos, repeats: int = 2) -> dict:
    """Time all variants on a batch (the tile lists depend on the variant, so
    one plan is built per variant) and install the fastest.  ``imgs`` and
    ``sinos`` are overwritten.  Batched launches use no bands, so the bands
    are switched off: single-slice calls then run the same kernels as the
    batch."""
    from .batch import BatchPlan

    set_variant(fwd_bands=1, back_bands=1, back_stage=0)

    def timed(op):
        plan = BatchPlan(settings, imgs, sinos)
        try:
            return _time(lambda: plan.launch(op), repeats)
        finally:
            plan.destroy()

    fwd_times, back_times = {}, {}
    for k in FWD_GROUPS:
        set_variant(fwd_group=k)
        fwd_times[k] = timed(_hip.BATCH_FORWARD)
    for r in BACK_ROWS:
        set_variant(back_rows=r)
        back_times[r] = timed(_hip.BATCH_BACKWARD)
    best = {"fwd_group": min(fwd_times, key=fwd_times.get), "back_rows": min(back_times, key=back_times.get),
            "fwd_ms": fwd_times, "back_ms": back_times}
    set_variant(best["fwd_group"], best["back_rows"])
    return best


def apply(projectionsetting, dtype, nz: int = 1) -> bool:
    """Install the cached variant for this shape; False when none is cached."""
    entry = _load().get(shape_key(projectionsetting, dtype, nz))
    if not entry:
        return False
    set_variant(entry.get("fwd_group"), entry.get("back_rows"), entry.get("fwd_bands", 1),
                entry.get("back_bands", 1), entry.get("back_stage", 0), entry.get("back_tile", 16))
    return True
