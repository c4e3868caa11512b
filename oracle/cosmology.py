"""Float64 restatement of PM++'s background and growth functions used by the N-body factors.

Test infrastructure only (the product computes its own factors in ``pmpp_amd.cosmology``).
Follows (read as text, never imported):

* ``src/pmpp/cosmology/models.py:16-100``   -- ``E2`` (w0/wa dark energy, curvature),
  ``H_deriv`` = d ln H / d ln a, ``Omega_m_a``.
* ``src/pmpp/cosmology/boltzmann.py:178-246`` -- ``growth_integ``: the growth ODE in ln a
  for G_m = D_m / a^m (m = 1, 2) from G = (1, 0, 3/7, 0) at a_ic = cbrt(eps)/2, tabulated at
  ``conf.growth_a`` (node 0 is a = 0 and holds the a_ic value), with
  D_m'/a^m = m G + G' and D_m''/a^m = m^2 G + 2 m G' + G''.
* ``src/pmpp/cosmology/boltzmann.py:253-289`` -- ``growth``: a^m times the linear
  interpolation of that table.
* ``src/pmpp/core/configuration.py:526-566`` -- ``a_lpt``, ``a_nbody``, ``growth_a``.
* ``src/pmpp/nbody/integrator.py:114-170``    -- G_D, G_K, drift and kick factors.

The ODE is integrated with scipy's DOP853 at rtol 1e-12 (the reference's adaptive
Dormand-Prince runs at rtol = atol = sqrt(eps64) ~ 1.5e-8, so tables agree to ~1e-8).

Pinned in ``tests/test_oracle_cosmology.py`` by the Einstein-de Sitter closed forms
(G_1 = 1, G_2 = 3/7 exactly when Omega_m = 1) and, for LCDM, Heath's integral
D_1 ∝ H(a) ∫_0^a da' / (a' H(a'))^3 for the linear growth.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.integrate import solve_ivp


def E2(a, Om, Ok=0.0, w0=-1.0, wa=0.0):
    a = np.asarray(a, dtype=np.float64)
    Ode = 1.0 - Om - Ok
    de = a ** (-3 * (1 + w0 + wa)) * np.exp(-3 * wa * (1 - a))
    return Om * a ** -3 + Ok * a ** -2 + Ode * de


def _dE2_da(a, Om, Ok=0.0, w0=-1.0, wa=0.0):
    a = np.asarray(a, dtype=np.float64)
    Ode = 1.0 - Om - Ok
    de = a ** (-3 * (1 + w0 + wa)) * np.exp(-3 * wa * (1 - a))
    dde = de * (-3 * (1 + w0 + wa) / a + 3 * wa)
    return -3 * Om * a ** -4 - 2 * Ok * a ** -3 + Ode * dde


def H_deriv(a, Om, **kw):
    """d ln H / d ln a = a E2'(a) / (2 E2(a))  (models.py:54-73)."""
    return 0.5 * a * _dE2_da(a, Om, **kw) / E2(a, Om, **kw)


def Omega_m_a(a, Om, **kw):
    return Om * np.asarray(a, dtype=np.float64) ** -3 / E2(a, Om, **kw)


def a_grids(a_start=1 / 64, a_stop=1.0, a_nbody_maxstep=1 / 64, a_lpt_maxstep=1 / 128):
    """(a_lpt, a_nbody, growth_a) of the reference configuration (configuration.py:526-566)."""
    n_lpt = math.ceil(a_start / a_lpt_maxstep)
    n_nbody = math.ceil((a_stop - a_start) / a_nbody_maxstep)
    a_lpt = np.linspace(0.0, a_start, n_lpt + 1)
    a_nbody = np.linspace(a_start, a_stop, n_nbody + 1)
    return a_lpt, a_nbody, np.concatenate((a_lpt, a_nbody[1:]))


def growth_table(growth_a, Om, a_lpt_step, **kw):
    """Table (2 orders, 3 derivatives, len(growth_a)) of D_m^(n) / a^m (boltzmann.py:178-246)."""
    eps = np.finfo(np.float64).eps
    a_ic = 0.5 * np.cbrt(eps)
    if a_ic >= a_lpt_step:
        a_ic = 0.1 * a_lpt_step
    a = np.array(growth_a, dtype=np.float64)
    a[0] = a_ic
    lna = np.log(a)

    def ode(t, G):
        aa = math.exp(t)
        h = H_deriv(aa, Om, **kw)
        of = 1.5 * Omega_m_a(aa, Om, **kw)
        G1, G1p, G2, G2p = G
        G1pp = -(3 + h - of) * G1 - (4 + h) * G1p
        G2pp = of * G1 ** 2 - (8 + 2 * h - of) * G2 - (6 + h) * G2p
        return [G1p, G1pp, G2p, G2pp]

    sol = solve_ivp(ode, (lna[0], lna[-1]), [1.0, 0.0, 3.0 / 7.0, 0.0], method="DOP853", t_eval=lna,
                    rtol=1e-12, atol=1e-14)
    G = sol.y.T  # (num_a, 4)
    d = np.array([ode(t, g) for t, g in zip(lna, G)])
    out = np.empty((2, 3, len(a)))
    for m, (g, gp, gpp) in enumerate(((G[:, 0], G[:, 1], d[:, 1]), (G[:, 2], G[:, 3], d[:, 3])), start=1):
        out[m - 1, 0] = g
        out[m - 1, 1] = m * g + gp
        out[m - 1, 2] = m * m * g + 2 * m * gp + gpp
    return out


def growth(a, table, growth_a, order=1, deriv=0):
    """a^order times the linear interpolation of the table (boltzmann.py:253-289)."""
    a = np.asarray(a, dtype=np.float64)
    return a ** order * np.interp(a, growth_a, table[order - 1][deriv])


def G_D(a, table, growth_a, Om, **kw):
    """Growth of the ZA canonical velocity (integrator.py:114-116)."""
    return a ** 2 * np.sqrt(E2(a, Om, **kw)) * growth(a, table, growth_a, 1, 1)


def G_K(a, table, growth_a, Om, **kw):
    """Growth of the ZA acceleration (integrator.py:119-122)."""
    return a ** 3 * E2(a, Om, **kw) * (growth(a, table, growth_a, 1, 2)
                                       + (2 + H_deriv(a, Om, **kw)) * growth(a, table, growth_a, 1, 1))


def drift_factor(a_vel, a_prev, a_next, table, growth_a, Om, **kw):
    """(D(a_next) - D(a_prev)) / G_D(a_vel)  (integrator.py:125-146)."""
    num = growth(a_next, table, growth_a) - growth(a_prev, table, growth_a)
    return num / G_D(a_vel, table, growth_a, Om, **kw)


def kick_factor(a_acc, a_prev, a_next, table, growth_a, Om, **kw):
    """(G_D(a_next) - G_D(a_prev)) / G_K(a_acc)  (integrator.py:149-170)."""
    num = G_D(a_next, table, growth_a, Om, **kw) - G_D(a_prev, table, growth_a, Om, **kw)
    return num / G_K(a_acc, table, growth_a, Om, **kw)
