"""Write tests/golden/walnut_sinogram.npy from the reference's data file.

The walnut sinogram (Hamalainen et al., "Tomographic X-ray data of a walnut",
arXiv:1502.04064, CC BY 4.0) is the input of the reference's solver tests
(tests/test_fanbeam.py:1094-1382), read there with matplotlib.image.imread.
This script reads the same PNG the same way and stores the float32 array, so
the GPU box never needs /root/reference.  Run from the repository root:
    python tests/golden/make_walnut.py
"""

import os

import matplotlib.image as mpimg
import numpy as np

SRC = "/root/reference/tests/walnut_sinogram.png"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "walnut_sinogram.npy")

if __name__ == "__main__":
    data = mpimg.imread(SRC)
    assert data.dtype == np.float32 and data.shape == (328, 120), (data.dtype, data.shape)
    np.save(DST, data)
    print(DST, data.shape, data.dtype, float(data.min()), float(data.max()))
