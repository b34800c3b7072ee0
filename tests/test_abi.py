"""The C ABI of libgratopy_hip.so on the CPU: the library loads, exports
every function include/gratopy_hip.h declares, matches the ctypes
prototypes, and rejects invalid arguments with a status code and message
before touching the device."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

from gratopy_amd import _hip
from gratopy_amd import build as hip_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gratopy_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    hip_build.ensure_built()
    return _hip.load_library()


def test_header_declares_the_bound_prototypes():
    assert _declared() == sorted(_hip.PROTOTYPES)


def test_library_exports_every_declared_symbol(lib):
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_empty_error(lib):
    assert lib.gt_version() >= 10000
    buf = ctypes.create_string_buffer(8)
    lib.gt_last_error(buf, 8)


def test_invalid_arguments_are_rejected_without_a_device(lib):
    # NULL arrays
    rc = lib.gt_radon_fwd(None, 0, None, None, None, None, 4, 4, 4, 4, 1, 0, 0, None, 0)
    assert rc == _hip.GT_EINVAL and "NULL" in _hip.last_error()
    # bad dtype
    rc = lib.gt_fan_back(None, 7, 1, 2, 3, 4, 5, 4, 4, 4, 4, 1, 0, 0, None, 0)
    assert rc == _hip.GT_EINVAL and "dtype" in _hip.last_error()
    # non-positive sizes
    rc = lib.gt_radon_back(None, 0, 1, 2, 3, 4, 0, 4, 4, 4, 1, 0, 0, 8, 1 << 20)
    assert rc == _hip.GT_EINVAL and "positive" in _hip.last_error()
    # workspace too small
    need = lib.gt_projection_workspace_bytes(0, 1, 0, 8, 8, 8, 8, 1, 0)
    assert need > 0
    rc = lib.gt_radon_back(None, 0, 1, 2, 3, 4, 8, 8, 8, 8, 1, 0, 0, 8, need - 1)
    assert rc == _hip.GT_ENOSPC and "workspace" in _hip.last_error()
    # aliasing output and input
    rc = lib.gt_radon_fwd(None, 0, 5, 5, 3, 4, 4, 4, 4, 4, 1, 0, 0, 8, 1 << 20)
    assert rc == _hip.GT_EINVAL and "alias" in _hip.last_error()
    # pixel outside the image for the single-line kernel
    rc = lib.gt_single_line(None, 0, 0, 1, 9, 0, 2, None, 3, 4, 4, 4, 4, 0)
    assert rc == _hip.GT_EINVAL
    with pytest.raises(ValueError, match="dtype"):
        _hip.check(lib.gt_weight_sino(None, 5, 1, 2, 3, 4, 4, 1, 0, 0), "gt_weight_sino")


def test_workspace_sizes(lib):
    es = 4
    # forward, F-ordered image: one transposed copy
    assert lib.gt_projection_workspace_bytes(0, 0, 0, 100, 60, 80, 30, 1, 0) >= es * 100 * 60
    # forward, C-ordered 3-D image: two copies
    assert lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 3, 1) >= 2 * es * 100 * 60 * 3
    # backward: padded sinogram + per-angle table
    assert lib.gt_projection_workspace_bytes(1, 1, 1, 100, 60, 80, 30, 2, 0) >= 8 * (84 * 30 * 2 + 8 * 30)


def test_tuning_knobs_accept_their_values_and_reject_others(lib):
    from gratopy_amd import autotune

    before = autotune.current_variant()
    valid = {"fwd_group": (1, 2, 4, 8, 16, 32), "back_rows": (1, 2, 4, 8),
             "fwd_bands": (1, 2, 4, 8, 16, 32), "back_bands": (1, 2, 4, 8), "back_stage": (0, 8, 16),
             "back_tile": (16, 32), "fwd_chunk": (1, 4), "automatic": (0, 1), "fwd_compact": (0, 1)}
    try:
        assert set(valid) == set(autotune.KNOBS)
        for name, values in valid.items():
            key = autotune.KNOBS[name]
            for v in values:
                assert lib.gt_tune_set(key, v) == _hip.GT_OK and lib.gt_tune_get(key) == v
            for v in (-1, 3, 5, 64):
                if v not in values:
                    assert lib.gt_tune_set(key, v) == _hip.GT_EINVAL, (name, v)
                    assert lib.gt_tune_get(key) == values[-1]
        assert lib.gt_tune_set(99, 1) == _hip.GT_EINVAL and lib.gt_tune_get(99) == -1
        # the chunked forward needs two row-padded copies in its workspace
        autotune.set_variant(fwd_group=8, fwd_bands=1, fwd_chunk=4)
        assert lib.gt_tune_get(autotune.KNOBS["automatic"]) == 0
        padded = lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 1, 0)
        assert padded >= 4 * ((100 + 4) * 60 + (60 + 4) * 100)
        autotune.set_variant(fwd_chunk=1)
        assert lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 1, 0) < padded
        # 32 bins per lane always read 16 bytes per load (padded copies)
        autotune.set_variant(fwd_group=32)
        assert lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 1, 0) == padded
        # a compacted banded forward adds its group plan to the band partials
        autotune.set_variant(fwd_group=8, fwd_bands=8, fwd_chunk=4, fwd_compact=0)
        plain = lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 1, 0)
        assert plain >= padded + 4 * 8 * 80 * 30
        autotune.set_variant(fwd_compact=1)
        assert lib.gt_projection_workspace_bytes(1, 0, 0, 100, 60, 80, 30, 1, 0) >= plain + 8 * 8 * 30
    finally:
        autotune.set_variant(**before)


def test_automatic_workspace_follows_the_automatic_shape(lib):
    """In automatic mode the workspace query sizes what the chosen shape uses:
    a small parallel-beam forward reads row-padded copies and a small fanbeam
    one does not; the backprojection keeps angle-band partials only below
    4096 * 2048 pixels; tiny calls (below 2e6 pixel-angles) use no padded
    copies and no backprojection bands."""
    from gratopy_amd import autotune

    before = autotune.current_variant()
    try:
        autotune.set_variant(automatic=1)
        ws = lib.gt_projection_workspace_bytes
        es, N, Ns, Na = 4, 100, 150, 300
        tiny = 30  # angles: 3e5 pixel-angles
        # F order: one y-fastest copy and the forward's 8 band partials
        assert ws(0, 0, 0, N, N, Ns, tiny, 1, 0) < es * ((N + 4) * N + 8 * Ns * tiny)
        assert ws(0, 1, 0, N, N, Ns, tiny, 1, 0) < es * N * N  # no band partials
        radon_f, fan_f = ws(0, 0, 0, N, N, Ns, Na, 1, 0), ws(1, 0, 0, N, N, Ns, Na, 1, 0)
        assert radon_f >= es * ((N + 4) * N * 2 + 8 * Ns * Na)
        assert fan_f < es * (N + 4) * N * 2 + es * 8 * Ns * Na
        small = ws(1, 1, 0, N, N, Ns, Na, 1, 0)
        assert small >= es * 8 * N * N
        big_n = 4096  # 1.7e7 pixels: staged, no bands
        big = ws(1, 1, 0, big_n, big_n, Ns, 10, 1, 0)
        assert big < es * 2 * big_n * big_n
        assert ws(0, 1, 0, big_n, big_n, Ns, 10, 1, 0) == big
    finally:
        autotune.set_variant(**before)


def test_batch_item_layout_matches_the_header(tmp_path):
    """ctypes' gt_batch_item has the C struct's size and field offsets (gcc on
    the header)."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "gratopy_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(gt_batch_item),"
        " offsetof(gt_batch_item, ofs), offsetof(gt_batch_item, img),"
        " offsetof(gt_batch_item, field), offsetof(gt_batch_item, p));return 0;}\n")
    exe = tmp_path / "layout"
    import subprocess

    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    B = _hip.BatchItem
    assert got == [ctypes.sizeof(B), B.ofs.offset, B.img.offset, B.field.offset, B.p.offset]


def test_batch_arguments_rejected_without_a_device(lib):
    plan = ctypes.c_void_p()
    rc = lib.gt_batch_plan_create(ctypes.byref(plan), 0, 0, None)
    assert rc == _hip.GT_EINVAL and "at least one" in _hip.last_error()
    items = (_hip.BatchItem * 2)()
    for it in items:
        it.geometry, it.Nx, it.Ny, it.Ns, it.Na = 0, 8, 8, 8, 4
        it.ofs, it.geo, it.img, it.sino = 1, 2, 3, 4
    items[1].geometry = 1  # fanbeam without sdpd
    rc = lib.gt_batch_plan_create(ctypes.byref(plan), 0, 2, ctypes.cast(items, ctypes.c_void_p))
    assert rc == _hip.GT_EINVAL and "sdpd" in _hip.last_error()
    items[1].geometry = 5
    rc = lib.gt_batch_plan_create(ctypes.byref(plan), 0, 2, ctypes.cast(items, ctypes.c_void_p))
    assert rc == _hip.GT_EINVAL and "geometry" in _hip.last_error()
    assert plan.value is None
    assert lib.gt_batch_launch(None, None, 0, 0.0) == _hip.GT_EINVAL
    assert lib.gt_batch_plan_destroy(None) == _hip.GT_OK
    assert lib.gt_batch_plan_bytes(None) == 0


def test_noalloc_rejects_calls_beyond_its_array(lib):
    # a 4096^2 x 8 float64 fan backprojection needs ~0.5 GiB of workspace; the
    # no-allocation entry point holds 128 MiB and refuses before any HIP call
    need = lib.gt_projection_workspace_bytes(1, 1, 1, 4096, 4096, 4096, 2048, 8, 0)
    assert need > (128 << 20)
    p = [ctypes.c_void_p(16 * k) for k in range(1, 6)]
    rc = lib.gt_projection_noalloc(None, 1, 1, 1, *p, 4096, 4096, 4096, 2048, 8, 0, 0)
    assert rc == _hip.GT_ENOSPC and "no-allocation" in _hip.last_error()
    rc = lib.gt_projection_noalloc(None, 1, 2, 1, *p, 64, 64, 64, 64, 1, 0, 0)
    assert rc == _hip.GT_EINVAL and "adjoint" in _hip.last_error()


def test_projection_shape_reports_the_automatic_choice(lib):
    from gratopy_amd import autotune

    before = autotune.current_variant()
    try:
        autotune.set_variant(automatic=1)
        shape = (ctypes.c_int * 4)()
        # the benchmark stacks: 32 x 2048^2, 1440 angles, 2560 detectors
        assert lib.gt_projection_shape(1, 0, 2048, 2048, 2560, 1440, 32, shape) == _hip.GT_OK
        assert list(shape) == [16, 4, 4, 1]
        assert lib.gt_projection_shape(0, 0, 2048, 2048, 2560, 1440, 32, shape) == _hip.GT_OK
        assert list(shape) == [16, 4, 4, 1]
        # one such slice: 16 bands (too few blocks with 4); one 512^2 fan slice: 32
        assert lib.gt_projection_shape(1, 0, 2048, 2048, 2560, 1440, 1, shape) == _hip.GT_OK
        assert list(shape) == [16, 16, 4, 1]
        assert lib.gt_projection_shape(1, 0, 512, 512, 768, 360, 1, shape) == _hip.GT_OK
        assert list(shape) == [8, 32, 4, 0]
        assert lib.gt_projection_shape(1, 1, 2048, 2048, 2560, 1440, 32, shape) == _hip.GT_OK
        assert list(shape) == [8, 1, 16, 32]
        assert lib.gt_projection_shape(0, 1, 2048, 2048, 2560, 1440, 32, shape) == _hip.GT_OK
        assert list(shape) == [8, 1, 16, 16]
        # tiny calls: 4 bins per lane, per-pixel reads, no compaction; unstaged back
        assert lib.gt_projection_shape(0, 0, 32, 32, 48, 30, 1, shape) == _hip.GT_OK
        assert list(shape) == [4, 8, 1, 0]
        assert lib.gt_projection_shape(0, 1, 32, 32, 48, 30, 1, shape) == _hip.GT_OK
        assert list(shape)[2] == 0
        # explicit knobs are reported as set
        autotune.set_variant(fwd_group=2, fwd_bands=32, fwd_chunk=1, fwd_compact=0)
        assert lib.gt_projection_shape(0, 0, 2048, 2048, 2560, 1440, 32, shape) == _hip.GT_OK
        assert list(shape) == [2, 32, 1, 0]
        assert lib.gt_projection_shape(0, 0, 0, 2048, 2560, 1440, 32, shape) == _hip.GT_EINVAL
        assert lib.gt_projection_shape(0, 0, 8, 8, 8, 8, 1, None) == _hip.GT_EINVAL
    finally:
        autotune.set_variant(**before)
