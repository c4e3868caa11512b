"""Background expansion and linear growth on the host (float64), for the N-body factors."""

from .background import E2, H_deriv, Omega_m_a, GrowthTable  # noqa: F401
