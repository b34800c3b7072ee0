"""gratopy_amd: MI355X-native drop-in for gratopy's projection hot path.

Mirrors the public surface of gratopy/__init__.py:21-48: the transforms and
solvers run as hand-written gfx950 HIP kernels behind a C ABI
(include/gratopy_hip.h), with torch tensors as device arrays.
"""

from ._version import __version__
from .gratopy import RADON, PARALLEL, FAN, FANBEAM, VERSION
from .gratopy import (
    forwardprojection,
    backprojection,
    ProjectionSettings,
    landweber,
    conjugate_gradients,
    total_variation,
    normest,
    weight_sinogram,
)
from .phantom import ct_shepp_logan as phantom, easy_phantom

from . import operator  # noqa: E402

from .gratopy import (  # internal functions, exported like the reference
    radon,
    radon_ad,
    radon_struct,
    fanbeam,
    fanbeam_ad,
    fanbeam_struct,
    read_angles,
)
