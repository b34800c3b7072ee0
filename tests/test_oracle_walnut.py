"""The oracle's solver loops against the reference's walnut control numbers
(tests/test_fanbeam.py:1094-1239 of the reference), on the CPU.

The CG number 0.75674 is reproduced by the oracle in float32 (the
reference's own arithmetic); the oracle in float64 gives 0.759492 on the
same data.  That float64 value is what the HIP float32 CG is held to.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import gratopy_amd as gratopy
from oracle import solvers as osolv
from tests.helpers import assert_control, control_number, phantom

pytestmark = pytest.mark.slow

WALNUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walnut_sinogram.npy")


def _walnut_data():
    ps = gratopy.ProjectionSettings(None, gratopy.FANBEAM, img_shape=(600, 600),
                                    angles=-np.linspace(0, 2 * np.pi, 121)[:-1], detector_width=114.8,
                                    R=300.0, RE=110.0, n_detectors=328, detector_shift=0.27,
                                    reverse_detector=True)
    g = osolv.OracleGeometry.from_settings(ps)
    walnut = np.load(WALNUT)
    walnut /= np.mean(walnut)
    ph_sino = g.fwd(np.require(phantom((600, 600), np.float32), np.float32, "F"), "F")
    sino = np.zeros(ps.sinogram_shape + (2,))
    sino[:, :, 0] = walnut / np.max(walnut)
    sino[:, :, 1] = ph_sino / np.max(ph_sino)
    return g, sino


def test_walnut_cg_float32_reproduces_the_reference_and_float64_value():
    g, sino = _walnut_data()
    dims = (600, 600, 328, 120, 2)
    u32 = osolv.conjugate_gradients(np.require(sino, np.float32, "F"), g, 50)
    assert_control(u32, dims, 0.75674, "img", "oracle float32 CG", rtol=1e-4)
    u64 = osolv.conjugate_gradients(np.require(sino, np.float64, "F"), g, 50)
    assert abs(control_number(u64, dims, "img") - 0.7594917441) < 1e-8
