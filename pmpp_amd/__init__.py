"""pmpp_amd: an MI355X-native (gfx950 HIP) particle-mesh force step with PM++'s interfaces.

Layout mirrors the reference package: ``pmpp_amd.core`` (Configuration),
``pmpp_amd.nbody`` (Particles, gravity), ``pmpp_amd.cic`` (scatter/gather and the kernel
seam), ``pmpp_amd.numerics`` (real FFTs).  All arithmetic of the force step runs in
``libpmpp_amd.so`` (C ABI in ``include/pmpp_amd.h``); there is no CPU fallback.
"""

__version__ = "0.1.0"

from . import _lib  # noqa: F401
from .core import Configuration  # noqa: F401
from .nbody import Particles, gravity  # noqa: F401
from .cic import scatter, gather  # noqa: F401
