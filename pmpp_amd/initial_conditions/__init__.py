"""Initial conditions on the particle grid (PM++ ``pmpp.initial_conditions``): white noise,
linear modes and LPT on the HIP library."""

from ..cosmology.models import LCDM, SimpleLCDM  # noqa: F401
from .lpt import lpt, lpt_vjp, particles_2lpt, second_order_source  # noqa: F401
from .modes import KGrid, linear_modes, white_noise  # noqa: F401
