"""CIC mass assignment and interpolation on the HIP library (drop-in for ``pmpp.cic``)."""

from .kernels import hip_gather, hip_gather_bwd, hip_scatter, hip_scatter_bwd  # noqa: F401
from .scatter import scatter, _scatter  # noqa: F401
from .gather import gather, _gather  # noqa: F401
