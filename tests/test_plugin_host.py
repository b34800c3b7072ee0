"""Host side of operator kernels supplied as HIP sources: the kernel spec's
naming ABI and content signature (the reference's
gratopy/operator/opencl.py:55-155), run-time compilation for gfx950 (hiprtc
needs no GPU), the kernel list read from the code object, compile errors
carrying the compiler log, and the argument rules of the launcher."""

from __future__ import annotations

import ctypes
import shutil
from pathlib import Path

import numpy as np
import pytest
import torch

from gratopy_amd import _hip
from gratopy_amd.operator import HipKernelSpec, Radon, invalidate_kernel_cache
from gratopy_amd.operator.hip import (
    PLUGIN_INCLUDE_DIR,
    _PROGRAM_CACHE,
    _ctypes_argument,
    builtin_kernel_spec,
    code_object_kernels,
    compile_plugin,
    launch_plugin_kernel,
)

KERNELS = Path(__file__).parent / "kernels"
AFFINE = KERNELS / "affine.hip"
AFFINE_ALT = KERNELS / "affine_alt.hip"
RADON_BRUTEFORCE = KERNELS / "radon_bruteforce.hip"


def _abi_names(base):
    return {f"{base}{ad}_{t}_{o}{i}" for ad in ("", "_ad") for t in ("float", "double")
            for o in "fc" for i in "fc"}


def test_kernel_name_follows_the_reference_abi():
    spec = HipKernelSpec.from_path(AFFINE, base_name="affine")
    assert spec.kernel_name(np.float32, "F", "C") == "affine_float_fc"
    assert spec.kernel_name(np.float64, "c", "f", adjoint=True) == "affine_ad_double_cf"


def test_signature_depends_on_content_base_name_and_options(tmp_path):
    path = tmp_path / "k.hip"
    shutil.copy(AFFINE, path)
    a = HipKernelSpec.from_path(path, base_name="affine")
    assert a.paths == (str(path.resolve()),)
    assert a.signature == HipKernelSpec.from_path(path, base_name="affine").signature
    assert a.signature != HipKernelSpec.from_path(path, base_name="other").signature
    assert a.signature != HipKernelSpec.from_path(path, "affine", build_options=("-DX=1",)).signature
    before = a.signature
    shutil.copy(AFFINE_ALT, path)
    assert a.signature != before


def test_builtin_and_plugin_specs():
    assert builtin_kernel_spec("radon").is_builtin
    assert Radon(image_domain=8, angles=4).kernel_spec.is_builtin
    spec = HipKernelSpec.from_path(AFFINE, base_name="affine")
    assert not spec.is_builtin
    code = spec.build_code()
    assert code.startswith(f'#line 1 "{AFFINE.resolve()}"')
    opts = spec.compile_options()
    assert opts[0] == f"-I{PLUGIN_INCLUDE_DIR}" and f"-I{KERNELS.resolve()}" in opts
    two = HipKernelSpec.from_paths([AFFINE, RADON_BRUTEFORCE], base_name="radon",
                                   build_options=("-DFOO",))
    assert two.compile_options()[-1] == "-DFOO"
    assert two.build_code().count("#line 1") == 2


@pytest.mark.parametrize("path,base", [(AFFINE, "affine"), (AFFINE_ALT, "affine"),
                                       (RADON_BRUTEFORCE, "radon")])
def test_plugins_compile_for_gfx950_with_every_abi_variant(path, base):
    spec = HipKernelSpec.from_path(path, base_name=base)
    image = compile_plugin(spec.build_code(), f"{base}.hip", "gfx950", spec.compile_options())
    assert image[:4] == b"\x7fELF"
    assert set(code_object_kernels(image)) == _abi_names(base)


def test_compile_error_carries_the_compiler_log():
    src = 'extern "C" __global__ void broken_float_ff(float *o) { o[0] = not_declared; }'
    with pytest.raises(_hip.HipLibraryError, match="not_declared"):
        compile_plugin(src, "broken.hip", "gfx950")


def test_build_options_reach_the_compiler():
    src = ('#ifndef SCALE\n#error SCALE missing\n#endif\n'
           'extern "C" __global__ void s_float_ff(float *o) { o[0] = SCALE; }')
    with pytest.raises(_hip.HipLibraryError, match="SCALE missing"):
        compile_plugin(src, "s.hip", "gfx950")
    image = compile_plugin(src, "s.hip", "gfx950", ("-DSCALE=2.0f",))
    assert code_object_kernels(image) == ["s_float_ff"]


def test_code_object_kernels_rejects_non_elf():
    with pytest.raises(ValueError):
        code_object_kernels(b"not an elf file at all" * 4)


def test_kernel_argument_types_are_explicit():
    assert isinstance(_ctypes_argument(np.int32(3)), ctypes.c_int32)
    assert isinstance(_ctypes_argument(np.float32(1.5)), ctypes.c_float)
    assert isinstance(_ctypes_argument(np.float64(1.5)), ctypes.c_double)
    assert _ctypes_argument(np.int64(-2)).value == -2
    with pytest.raises(TypeError, match="numpy scalars"):
        _ctypes_argument(3)
    with pytest.raises(TypeError, match="numpy scalars"):
        _ctypes_argument(2.0)
    with pytest.raises(TypeError, match="device arrays"):
        _ctypes_argument(torch.zeros(3))


def test_launch_rejects_more_than_three_axes():
    with pytest.raises(ValueError, match="at most 3 axes"):
        launch_plugin_kernel(0, 0, (2, 2, 2, 2))


def test_hip_operators_define_the_execution_pipeline():
    from gratopy_amd.operator.base import Operator
    from gratopy_amd.operator.hip import _HipOperator, invalidate_kernel_cache as inv

    assert _HipOperator.apply_to is not Operator.apply_to
    assert Radon.apply_to is not Operator.apply_to
    assert not [c for c in inv.__code__.co_consts if hasattr(c, "co_name")]  # no nested defs


def test_invalidate_kernel_cache_empties_the_cache():
    _PROGRAM_CACHE[("cuda:0", "dummy")] = object()
    invalidate_kernel_cache()
    assert not _PROGRAM_CACHE
