"""P(k) and cross-correlation metrics on the HIP library (PM++ ``pmpp.analysis``)."""

from .power_spectrum import (  # noqa: F401
    delta_to_cross_correlation, delta_to_pk, delta_to_pk_vjp, density_to_cross_correlation,
    density_to_cross_correlation_vjp, density_to_pk, density_to_pk_vjp, particles_to_pk, shell_statistics,
)
