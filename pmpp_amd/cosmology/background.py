"""Hubble rate, matter fraction and the tabulated 1st/2nd-order growth functions (float64, host).

Same quantities as PM++'s ``cosmology/models.py:16-100`` (``E2``, ``H_deriv``,
``Omega_m_a``) and ``cosmology/boltzmann.py:178-289`` (``growth_integ`` / ``growth``): the
growth ODE in ln a for G_m = D_m / a^m, started from G = (1, 0, 3/7, 0) at
a_ic = cbrt(eps)/2 and tabulated at the configuration's ``growth_a`` nodes; ``growth`` is
a^m times the linear interpolation of that table, exactly as the reference evaluates it, so
kick and drift factors at mid-step scale factors interpolate the same way.  The ODE is
integrated here by an adaptive Dormand-Prince 5(4) pair (Dormand & Prince 1980, with the
4th-order embedded error estimate) at rtol = atol = 1e-11.
"""

from __future__ import annotations

import math

import numpy as np


def E2(a, cosmo):
    """H^2 / H_0^2 = Om a^-3 + Ok a^-2 + Ode a^{-3(1+w0+wa)} exp(-3 wa (1-a))."""
    a = np.asarray(a, dtype=np.float64)
    de = a ** (-3.0 * (1 + cosmo.w_0 + cosmo.w_a)) * np.exp(-3.0 * cosmo.w_a * (1 - a))
    return cosmo.Omega_m * a ** -3 + cosmo.Omega_k * a ** -2 + cosmo.Omega_de * de


def H_deriv(a, cosmo):
    """d ln H / d ln a = a (dE2/da) / (2 E2)."""
    a = np.asarray(a, dtype=np.float64)
    q = -3.0 * (1 + cosmo.w_0 + cosmo.w_a)
    de = a ** q * np.exp(-3.0 * cosmo.w_a * (1 - a))
    a_dE2 = -3.0 * cosmo.Omega_m * a ** -3 - 2.0 * cosmo.Omega_k * a ** -2 \
        + cosmo.Omega_de * de * (q + 3.0 * cosmo.w_a * a)
    return 0.5 * a_dE2 / E2(a, cosmo)


def Omega_m_a(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    return cosmo.Omega_m * a ** -3 / E2(a, cosmo)


# Dormand-Prince 5(4) tableau
_C = (0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0)
_A = (
    (),
    (1 / 5,),
    (3 / 40, 9 / 40),
    (44 / 45, -56 / 15, 32 / 9),
    (19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729),
    (9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656),
    (35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84),
)
_B5 = (35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0)
_B4 = (5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40)


def _dopri(f, y0, ts, rtol=1e-11, atol=1e-11):
    """Integrate y' = f(t, y) from ts[0], returning y at every ts (increasing)."""
    y = np.array(y0, dtype=np.float64)
    t = float(ts[0])
    out = [y.copy()]
    h = 1e-3
    for t_end in ts[1:]:
        t_end = float(t_end)
        while t < t_end:
            h = min(h, t_end - t)
            k = [f(t, y)]
            for i in range(1, 7):
                yi = y + h * sum(a * kk for a, kk in zip(_A[i], k))
                k.append(f(t + _C[i] * h, yi))
            y5 = y + h * sum(b * kk for b, kk in zip(_B5, k))
            y4 = y + h * sum(b * kk for b, kk in zip(_B4, k))
            scale = atol + rtol * np.maximum(np.abs(y), np.abs(y5))
            err = math.sqrt(float(np.mean(((y5 - y4) / scale) ** 2)))
            if err <= 1.0:
                t += h
                y = y5
            fac = 0.9 * err ** -0.2 if err > 0 else 5.0
            h *= min(5.0, max(0.2, fac))
        out.append(y.copy())
    return np.array(out)


class GrowthTable:
    """D_m^(n) / a^m, m = 1, 2 and n = 0, 1, 2 (derivatives in ln a), at ``growth_a``."""

    def __init__(self, cosmo, growth_a, a_lpt_step):
        self.cosmo = cosmo
        self.growth_a = np.asarray(growth_a, dtype=np.float64)
        a_ic = 0.5 * np.cbrt(np.finfo(np.float64).eps)
        if a_ic >= a_lpt_step:
            a_ic = 0.1 * a_lpt_step
        a = self.growth_a.copy()
        a[0] = a_ic
        lna = np.log(a)

        def ode(t, G):
            aa = math.exp(t)
            h = float(H_deriv(aa, cosmo))
            of = 1.5 * float(Omega_m_a(aa, cosmo))
            g1, g1p, g2, g2p = G
            return np.array([g1p, -(3 + h - of) * g1 - (4 + h) * g1p,
                             g2p, of * g1 * g1 - (8 + 2 * h - of) * g2 - (6 + h) * g2p])

        G = _dopri(ode, (1.0, 0.0, 3.0 / 7.0, 0.0), lna)
        d = np.array([ode(t, g) for t, g in zip(lna, G)])
        tab = np.empty((2, 3, len(a)))
        for m, (g, gp, gpp) in enumerate(((G[:, 0], G[:, 1], d[:, 1]), (G[:, 2], G[:, 3], d[:, 3])), start=1):
            tab[m - 1, 0] = g
            tab[m - 1, 1] = m * g + gp
            tab[m - 1, 2] = m * m * g + 2 * m * gp + gpp
        self.table = tab

    def growth(self, a, order=1, deriv=0):
        a = np.asarray(a, dtype=np.float64)
        return a ** order * np.interp(a, self.growth_a, self.table[order - 1][deriv])

    def G_D(self, a):
        """Growth of the Zel'dovich canonical velocity, in H_0 (nbody/integrator.py:114-116)."""
        return a ** 2 * np.sqrt(E2(a, self.cosmo)) * self.growth(a, 1, 1)

    def G_K(self, a):
        """Growth of the Zel'dovich acceleration, in H_0^2 (nbody/integrator.py:119-122)."""
        return a ** 3 * E2(a, self.cosmo) * (self.growth(a, 1, 2) + (2 + H_deriv(a, self.cosmo)) * self.growth(a, 1, 1))

    def drift_factor(self, a_vel, a_prev, a_next):
        return float((self.growth(a_next) - self.growth(a_prev)) / self.G_D(a_vel))

    def kick_factor(self, a_acc, a_prev, a_next):
        return float((self.G_D(a_next) - self.G_D(a_prev)) / self.G_K(a_acc))
