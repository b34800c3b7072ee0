from .projection import Radon, Fanbeam
from .base import IDENTITY, ZERO
from .hip import HipKernelSpec, invalidate_kernel_cache
