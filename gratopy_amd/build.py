"""Build libgratopy_hip.so in-tree with hipcc for gfx950.

``python -m gratopy_amd.build`` (or ``build_library()``) compiles every
source under ``gratopy_amd/csrc`` into one shared library next to this file.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgratopy_hip.so")
ARCH = os.environ.get("GRATOPY_HIP_ARCH", "gfx950")

SOURCES = ("runtime.hip", "elementwise.hip", "projection.hip", "batch.hip", "plugin.hip")

# hiprtc compiles operator kernels supplied at run time (plugin.hip)
LIBS = ["-lhiprtc"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off: the compiler emits plain VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def build_library(verbose: bool = False) -> str:
    """Compile the HIP sources (always, even if a library exists)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs, *LIBS]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, OUT)
    return OUT


def ensure_built() -> str:
    """Build only if the library is absent (tests and benchmarks use this)."""
    if not os.path.exists(OUT):
        build_library()
    return OUT


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
