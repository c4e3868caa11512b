"""Cosmological parameters (the subset of PM++'s ``cosmology/models.py:105-262`` the PM path
uses): flat or curved LCDM with w0-wa dark energy, normalised by sigma8 (the reference's
``Cosmology.from_sigma8``; ``A_s_1e9`` is derived from it)."""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class LCDM:
    Omega_m: float = 0.3
    Omega_b: float = 0.05
    h: float = 0.7
    n_s: float = 0.96
    sigma8: float = 0.8
    T_cmb: float = 2.7255
    # curvature and w0-wa dark energy of the background (cosmology/models.py:16-52)
    Omega_k: float = 0.0
    w_0: float = -1.0
    w_a: float = 0.0

    @property
    def Omega_de(self) -> float:
        return 1.0 - self.Omega_m - self.Omega_k

    @property
    def Omega_c(self) -> float:
        return self.Omega_m - self.Omega_b

    @property
    def A_s_1e9(self) -> float:
        """Primordial amplitude A_s / 1e-9 that yields ``sigma8`` (Eisenstein-Hu transfer)."""
        from .boltzmann import scalar_amplitude

        return scalar_amplitude(self)


SimpleLCDM = LCDM
