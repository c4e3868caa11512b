"""Linear matter power spectrum on the host (float64): the Eisenstein-Hu transfer fit and the
primordial normalisation of PM++'s ``cosmology/boltzmann.py``.

* ``transfer_fit`` -- Eisenstein & Hu (1998, ApJ 496, 605) fitting formulae, with baryon
  wiggles (their eqs. 2-24) or the no-wiggle form (eqs. 26-31), as ``boltzmann.py:33-137``;
  k in [h/Mpc] (the length unit is 1 Mpc/h, ``core/configuration.py:201-208``).
* ``transfer_table`` -- tabulation at ``transfer_k`` = [0, logspace(-4, 3)] with a 1/128
  decade step (``configuration.py:509-523``), linearly interpolated by ``transfer``.
* ``linear_power`` -- P(k) = 0.32 A_s k_pivot (k/k_pivot)^n_s (pi (c/H_0)^2 T(k) / Omega_m)^2
  D(a)^2 (``boltzmann.py:428-482``), D the growth normalised in matter domination.
* ``sigma8`` -- sqrt of the top-hat variance at R = 8 Mpc/h and a = 1.  PM++ evaluates it with
  mcfit's FFTLog ``TophatVar`` (not installed here); this is a trapezoid rule in ln k on the
  same interpolated power, extended with the table's end values (agreement with FFTLog is
  ~1e-5 in sigma8, not bit parity).  ``A_s_1e9`` inverts it: sigma8 is linear in sqrt(A_s)
  (``models.py:184-203``).
"""

from __future__ import annotations

import functools
import math

import numpy as np

MPC_SI = 3.0856775815e22
H0_SI = 1e5 / MPC_SI
C_SI = 299792458.0
C_OVER_H0 = C_SI / (MPC_SI * H0_SI)  # c / H_0 in Mpc/h
K_PIVOT_MPC = 0.05

TRANSFER_LGK_MIN = -4.0
TRANSFER_LGK_MAX = 3.0
TRANSFER_LGK_MAXSTEP = 1 / 128


def transfer_k():
    """[0, logspace(lgk_min, lgk_max)] in h/Mpc (configuration.py:509-523)."""
    num = 1 + math.ceil((TRANSFER_LGK_MAX - TRANSFER_LGK_MIN) / TRANSFER_LGK_MAXSTEP) + 1
    k = np.logspace(TRANSFER_LGK_MIN, TRANSFER_LGK_MAX, num=num - 1)
    return np.concatenate(([0.0], k))


def transfer_fit(k, cosmo, nowiggle=False):
    """Eisenstein & Hu (1998) matter transfer function at k [h/Mpc]."""
    k = np.asarray(k, dtype=np.float64) * cosmo.h  # 1/Mpc
    t2 = (cosmo.T_cmb / 2.7) ** 2
    h2 = cosmo.h ** 2
    w_m = cosmo.Omega_m * h2
    w_b = cosmo.Omega_b * h2
    zero_baryon = w_b <= 0
    w_b_safe = 1e-8 if zero_baryon else w_b
    f_b = cosmo.Omega_b / cosmo.Omega_m
    f_c = (cosmo.Omega_m - cosmo.Omega_b) / cosmo.Omega_m

    z_eq = 2.50e4 * w_m / t2 ** 2
    k_eq = 7.46e-2 * w_m / t2
    b1 = 0.313 * w_m ** -0.419 * (1 + 0.607 * w_m ** 0.674)
    b2 = 0.238 * w_m ** 0.223
    z_d = 1291 * w_m ** 0.251 / (1 + 0.659 * w_m ** 0.828) * (1 + b1 * w_b ** b2)
    R_d = 31.5 * w_b_safe / t2 ** 2 * (1e3 / z_d)
    R_eq = 31.5 * w_b_safe / t2 ** 2 * (1e3 / z_eq)
    s = 2 / (3 * k_eq) * np.sqrt(6 / R_eq) * np.log((np.sqrt(1 + R_d) + np.sqrt(R_eq + R_d)) / (1 + np.sqrt(R_eq)))
    k_silk = 1.6 * w_b ** 0.52 * w_m ** 0.73 * (1 + (10.4 * w_m) ** -0.95) if w_b > 0 else 1.0

    if nowiggle:
        alpha_g = 1 - 0.328 * np.log(431 * w_m) * f_b + 0.38 * np.log(22.3 * w_m) * f_b ** 2
        gamma_ratio = alpha_g + (1 - alpha_g) / (1 + (0.43 * k * s) ** 4)
        q = k / (13.41 * k_eq * gamma_ratio)
        L0 = np.log(2 * math.e + 1.8 * q)
        C0 = 14.2 + 731 / (1 + 62.5 * q)
        return L0 / (L0 + C0 * q ** 2)

    def t0_tilde(kk, alpha_c, beta_c):
        q = kk / (13.41 * k_eq)
        L = np.log(math.e + 1.8 * beta_c * q)
        C = 14.2 / alpha_c + 386 / (1 + 69.9 * q ** 1.08)
        return L / (L + C * q ** 2)

    if zero_baryon:
        return t0_tilde(k, 1.0, 1.0)
    a1 = (46.9 * w_m) ** 0.670 * (1 + (32.1 * w_m) ** -0.532)
    a2 = (12.0 * w_m) ** 0.424 * (1 + (45.0 * w_m) ** -0.582)
    alpha_c = a1 ** -f_b * a2 ** -f_b ** 3
    bb1 = 0.944 / (1 + (458 * w_m) ** -0.708)
    bb2 = (0.395 * w_m) ** -0.0266
    beta_c = 1 / (1 + bb1 * (f_c ** bb2 - 1))
    f = 1 / (1 + (k * s / 5.4) ** 4)
    T_c = f * t0_tilde(k, 1.0, beta_c) + (1 - f) * t0_tilde(k, alpha_c, beta_c)
    y = (1 + z_eq) / (1 + z_d)
    x = np.sqrt(1 + y)
    G = y * (-6 * x + (2 + 3 * y) * np.log((x + 1) / (x - 1)))
    alpha_b = 2.07 * k_eq * s * (1 + R_d) ** -0.75 * G
    beta_node = 8.41 * w_m ** 0.435
    beta_b = 0.5 + f_b + (3 - 2 * f_b) * np.sqrt(1 + (17.2 * w_m) ** 2)
    ks = k * s
    T_b = (t0_tilde(k, 1.0, 1.0) / (1 + (ks / 5.2) ** 2)
           + alpha_b * ks ** 3 / (beta_b ** 3 + ks ** 3) * np.exp(-(k / k_silk) ** 1.4)) \
        * np.sinc(ks ** 2 / (np.pi * np.cbrt(beta_node ** 3 + ks ** 3)))
    return f_c * T_c + f_b * T_b


@functools.lru_cache(maxsize=32)
def _table(key):
    omega_m, omega_b, h, t_cmb, nowiggle = key

    class _C:
        pass

    c = _C()
    c.Omega_m, c.Omega_b, c.h, c.T_cmb = omega_m, omega_b, h, t_cmb
    k = transfer_k()
    return k, transfer_fit(k, c, nowiggle)


def transfer_table(cosmo, nowiggle=False):
    """(transfer_k, T(transfer_k)) for ``cosmo`` (cached)."""
    return _table((float(cosmo.Omega_m), float(cosmo.Omega_b), float(cosmo.h), float(cosmo.T_cmb), bool(nowiggle)))


def k_pivot(cosmo):
    """The primordial pivot scale in h/Mpc."""
    return K_PIVOT_MPC / cosmo.h


def power_coef(cosmo, A_s_1e9):
    """0.32 A_s k_pivot (pi (c/H_0)^2 / Omega_m)^2, the prefactor of ``linear_power``."""
    return 0.32 * A_s_1e9 * 1e-9 * k_pivot(cosmo) * (math.pi * C_OVER_H0 ** 2 / cosmo.Omega_m) ** 2


def linear_power(k, cosmo, A_s_1e9=None, D=1.0, nowiggle=False):
    """Linear P(k) in (Mpc/h)^3 at k [h/Mpc], scaled by D^2 (boltzmann.py:428-482)."""
    if A_s_1e9 is None:
        A_s_1e9 = scalar_amplitude(cosmo, nowiggle)
    k = np.asarray(k, dtype=np.float64)
    tk, tT = transfer_table(cosmo, nowiggle)
    T = np.interp(k, tk, tT)
    kp = k_pivot(cosmo)
    with np.errstate(divide="ignore"):
        prim = np.where(k > 0, (np.where(k > 0, k, 1.0) / kp) ** cosmo.n_s, 0.0)
    return power_coef(cosmo, A_s_1e9) * prim * T * T * D * D


def sigma_tophat(R, cosmo, A_s_1e9, D=1.0, nowiggle=False):
    """rms linear overdensity in a top-hat sphere of radius R [Mpc/h]."""
    lnk = np.linspace(np.log(1e-6), np.log(1e4), 20001)
    k = np.exp(lnk)
    x = k * R
    w = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    small = x < 1e-3
    w[small] = 1 - x[small] ** 2 / 10
    integrand = k ** 3 * linear_power(k, cosmo, A_s_1e9, D, nowiggle) * w * w / (2 * np.pi ** 2)
    return float(np.sqrt(np.trapezoid(integrand, lnk)))


def scalar_amplitude(cosmo, nowiggle=False, growth_at_1=None):
    """A_s / 1e-9 that gives ``cosmo.sigma8`` (sigma8 scales as sqrt(A_s)).

    ``growth_at_1`` is D(a = 1) in the matter-domination normalisation (computed from the
    growth table of ``cosmo`` when None)."""
    if growth_at_1 is None:
        growth_at_1 = _growth_today(cosmo)
    s1 = sigma_tophat(8.0, cosmo, 1.0, growth_at_1, nowiggle)
    return (cosmo.sigma8 / s1) ** 2


@functools.lru_cache(maxsize=32)
def _growth_today_key(om, ok, w0, wa):
    from .background import GrowthTable

    class _B:
        pass

    b = _B()
    b.Omega_m, b.Omega_k, b.w_0, b.w_a = om, ok, w0, wa
    b.Omega_de = 1.0 - om - ok
    nodes = np.linspace(0.0, 1.0, 257)
    return float(GrowthTable(b, nodes, nodes[1]).growth(1.0))


def _growth_today(cosmo):
    return _growth_today_key(float(cosmo.Omega_m), float(getattr(cosmo, "Omega_k", 0.0)),
                             float(getattr(cosmo, "w_0", -1.0)), float(getattr(cosmo, "w_a", 0.0)))
