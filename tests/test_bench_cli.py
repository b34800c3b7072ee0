"""bench.py on the CPU: the one-GPU contract and the workload table (no GPU
work runs here)."""

from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert proc.returncode != 0
    line = json.loads(proc.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


@pytest.mark.parametrize("workload", sorted(bench.WORKLOADS))
def test_workloads_build_settings_of_their_geometry(workload):
    import gratopy_amd as gratopy

    Nz, N, Na, Ns, RE, R, iw, dw, _ = bench.WORKLOADS[workload]
    fan, kw = bench.projection_kwargs(workload)
    assert fan == workload.startswith("fanbeam")
    assert fan == (R is not None and RE is not None)
    # the same keyword arguments on a small image: geometry tables build on the CPU
    n = 32
    scaled = {k: v * n / N for k, v in kw.items()}
    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM if fan else gratopy.RADON, (n, n), 8,
                                    max(8, Ns * n // N), **scaled)
    assert ps.is_fan == fan
    assert bench.DEFAULT_WORKLOAD in bench.WORKLOADS
