"""Launch-shape autotuner for the projection kernels.

By default every call picks its launch shape from its size (``automatic``).
Knobs (the C ABI's ``GT_TUNE_*`` keys): the forward kernels assign K
adjacent detector bins to a lane (``fwd_group`` K in 1, 2, 4, 8, 16, 32; K = 1 is
the per-ray kernel), optionally split the strip walk into 8 or 32 row bands
(``fwd_bands``; with 8, one band per XCD; 32 give small images four times the
blocks; partial sums added in band order), read strip rows 16 bytes per
load (``fwd_chunk`` = 4; always with K = 32).  The backprojection assigns
RY pixels along y to a lane (``back_rows`` in 1, 2, 4, 8), sums all angles per
tile or splits them into 8 angle bands (``back_bands``), and reads its taps
from memory or stages each tile's sinogram windows in LDS for chunks of 8 or
16 angles (``back_stage``) in tiles 16 or 32 pixels wide (``back_tile``).
Batched launches take all but the angle bands and take at most 8 row bands.  :func:`tune` times every variant on the
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

# knob name -> GT_TUNE_* key of the C ABI; "automatic" (GT_TUNE_AUTO) = 1 lets
# every call pick its launch shape from its size and ignore the other knobs
KNOBS = {"fwd_group": 0, "back_rows": 1, "fwd_bands": 2, "back_bands": 3, "back_stage": 4,
         "back_tile": 5, "fwd_chunk": 6, "automatic": 7, "fwd_compact": 8}
# the library's values at load time
DEFAULTS = {"fwd_group": 4, "back_rows": 1, "fwd_bands": 1, "back_bands": 1, "back_stage": 0,
            "back_tile": 16, "fwd_chunk": 1, "automatic": 1, "fwd_compact": 1}

FWD_GROUPS = (1, 2, 4, 8, 16)
BACK_ROWS = (1, 2, 4, 8)
FWD_BANDS = (1, 4, 8, 16, 32)
FWD_CHUNKS = (1, 4)
BACK_STAGES = (0, 8, 16)
BACK_TILES = (16, 32)
# forward variants; bands and chunked reads need bin groups; 32 bins per lane
# come with chunked reads only
# (banded ones compacted: only the bin groups that meet a band are launched,
# bitwise the same result)
FWD_VARIANTS = ({"fwd_group": 1, "fwd_bands": 1, "fwd_chunk": 1, "fwd_compact": 1},) + tuple(
    {"fwd_group": k, "fwd_bands": b, "fwd_chunk": c, "fwd_compact": 1}
    for k in FWD_GROUPS if k > 1 for b in FWD_BANDS for c in FWD_CHUNKS) + tuple(
    {"fwd_group": 32, "fwd_bands": b, "fwd_chunk": 4, "fwd_compact": 1} for b in FWD_BANDS)
# backprojection variants: unstaged (with or without angle bands), staged, and
# staged in chunks of 16 angles over 8 angle bands
BACK_VARIANTS = tuple(
    {"back_rows": r, "back_bands": b, "back_stage": 0, "back_tile": 16}
    for b in (1, 8) for r in BACK_ROWS) + tuple(
    {"back_rows": r, "back_bands": 1, "back_stage": s, "back_tile": t}
    for s in BACK_STAGES if s for t in BACK_TILES for r in BACK_ROWS) + tuple(
    {"back_rows": r, "back_bands": 8, "back_stage": 16, "back_tile": t}
    for t in BACK_TILES for r in BACK_ROWS)

# tune() skips forward variants with more than 8 row bands whose band partials
# would exceed this many bytes of workspace
MAX_BAND_PARTIAL_BYTES = 8 << 30

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


def set_variant(**knobs) -> None:
    """Set launch-shape knobs (names of :data:`KNOBS`; None leaves a knob).
    Setting any knob switches the automatic choice off unless ``automatic``
    is given as well."""
    lib = _hip.load_library()
    for name, value in knobs.items():
        if name not in KNOBS:
            raise ValueError(f"unknown tuning knob {name!r}")
        if value is not None:
            _hip.check(lib.gt_tune_set(KNOBS[name], int(value)), "gt_tune_set")
    if "automatic" not in knobs and any(v is not None for v in knobs.values()):
        _hip.check(lib.gt_tune_set(KNOBS["automatic"], 0), "gt_tune_set")


def current_variant() -> dict:
    lib = _hip.load_library()
    return {name: lib.gt_tune_get(key) for name, key in KNOBS.items()}


def variant_key(v: dict) -> str:
    """Short name of a variant, e.g. ``K4xB8xC4`` or ``RY8xS16T32``."""
    if "fwd_group" in v:
        key = f"K{v['fwd_group']}"
        if v.get("fwd_bands", 1) > 1:
            key += f"xB{v['fwd_bands']}"
        if v.get("fwd_chunk", 1) > 1:
            key += f"xC{v['fwd_chunk']}"
        return key
    key = f"RY{v['back_rows']}"
    if v.get("back_stage", 0):
        key += f"xS{v['back_stage']}T{v.get('back_tile', 16)}"
    if v.get("back_bands", 1) > 1:
        key += f"xB{v['back_bands']}"
    return key


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
    will use; their contents are overwritten.  Returns the chosen knobs plus
    the trial times (ms) per variant under ``fwd_ms`` and ``back_ms``.
    """
    from .gratopy import backprojection, forwardprojection

    ps = projectionsetting
    # band partials take bands * Ns * Na * Nz elements of workspace; more than
    # 8 bands only pay where the launch is small, so large stacks skip them
    partial = int(sino.numel()) * sino.element_size()
    fwd_times, back_times = {}, {}
    for v in FWD_VARIANTS:
        if v["fwd_bands"] > 8 and v["fwd_bands"] * partial > MAX_BAND_PARTIAL_BYTES:
            continue
        set_variant(**v)
        fwd_times[variant_key(v)] = (_time(lambda: forwardprojection(img, ps, sino=sino), repeats), v)
    for v in BACK_VARIANTS:
        set_variant(**v)
        back_times[variant_key(v)] = (_time(lambda: backprojection(sino, ps, img=img), repeats), v)
    best_fwd = min(fwd_times.values(), key=lambda tv: tv[0])[1]
    best_back = min(back_times.values(), key=lambda tv: tv[0])[1]
    knobs = {**best_fwd, **best_back}
    set_variant(**knobs)
    if store:
        nz = int(img.shape[2]) if img.dim() > 2 else 1
        with _lock:
            table = _load()
            table[shape_key(ps, dev.to_np_dtype(img.dtype), nz)] = knobs
            _store(table)
    return {**knobs, "fwd_ms": {k: t for k, (t, _) in fwd_times.items()},
            "back_ms": {k: t for k, (t, _) in back_times.items()}}


def tune_batch(settings, imgs, sinos, repeats: int = 2) -> dict:
    """Time the batch variants (forward K, row bands and chunked reads;
    backprojection RY, unstaged or LDS-staged) on a batch and install the
    fastest.  The tile lists depend on the variant, so one plan is built per
    variant.  ``imgs`` and ``sinos`` are overwritten.  Batched backprojections
    use no angle bands, so those are switched off: single-slice calls then run
    the same kernels as the batch."""
    from .batch import BatchPlan

    set_variant(back_bands=1)

    def timed(op):
        plan = BatchPlan(settings, imgs, sinos)
        try:
            return _time(lambda: plan.launch(op), repeats)
        finally:
            plan.destroy()

    fwd_times, back_times = {}, {}
    for v in FWD_VARIANTS:
        # batch plans take at most 8 row bands and 16 bins per lane
        if v["fwd_bands"] > 8 or v["fwd_group"] > 16:
            continue
        set_variant(**v)
        fwd_times[variant_key(v)] = (timed(_hip.BATCH_FORWARD), v)
    for v in BACK_VARIANTS:
        if v["back_bands"] == 1:
            set_variant(**v)
            back_times[variant_key(v)] = (timed(_hip.BATCH_BACKWARD), v)
    knobs = {**min(fwd_times.values(), key=lambda tv: tv[0])[1],
             **min(back_times.values(), key=lambda tv: tv[0])[1]}
    set_variant(**knobs)
    return {**knobs, "fwd_ms": {k: t for k, (t, _) in fwd_times.items()},
            "back_ms": {k: t for k, (t, _) in back_times.items()}}


def apply(projectionsetting, dtype, nz: int = 1) -> bool:
    """Install the cached variant for this shape (knobs a cache entry lacks
    take their defaults); False when none is cached."""
    entry = _load().get(shape_key(projectionsetting, dtype, nz))
    if not entry:
        return False
    set_variant(**{name: entry.get(name, DEFAULTS[name]) for name in KNOBS if name != "automatic"})
    return True
