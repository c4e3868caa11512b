"""Background expansion, linear growth and the linear power spectrum on the host (float64)."""

from .background import E2, H_deriv, Omega_m_a, GrowthTable  # noqa: F401
from .models import LCDM, SimpleLCDM  # noqa: F401
