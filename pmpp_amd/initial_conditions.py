"""Host-side initial conditions for the PM step: linear modes and 2LPT displacements.

Restates the reference's IC pipeline at the level the force step needs
(``src/pmpp/initial_conditions/modes.py:14-68, 399-454`` white noise and linear modes,
``src/pmpp/initial_conditions/lpt.py:339-478`` 2LPT,
``src/pmpp/cosmology/boltzmann.py:10-290`` Eisenstein-Hu transfer):

* white noise from ``numpy.random.Generator(PCG64(seed))`` (PM++'s JAX threefry stream
  cannot be reproduced without JAX, so "seed 0" means this generator's seed 0);
* the Eisenstein & Hu (1998) no-wiggle transfer function, primordial k^n_s, normalised to
  sigma_8 with a top-hat window (numerical quadrature instead of mcfit's FFTLog);
* 2LPT displacement Psi = -D1 grad(phi1) + D2 grad(phi2), D2 = -3/7 D1^2 (growth-suppression
  factors are ignored at this order, as the second-order growth approximation does).

The FFTs run on the HIP library (``pmpp_amd.numerics``); only elementwise work uses torch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .numerics import irfftn, rfftn


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


def transfer_eh_nowiggle(k_hmpc, c: LCDM):
    """Eisenstein & Hu (1998) eqs. 29-31; k in h/Mpc."""
    k = np.asarray(k_hmpc, dtype=np.float64)
    om, ob, h = c.Omega_m, c.Omega_b, c.h
    omh2, obh2 = om * h * h, ob * h * h
    fb = ob / om
    theta = c.T_cmb / 2.7
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1 + 10 * obh2 ** 0.75)  # Mpc
    alpha = 1 - 0.328 * np.log(431 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb * fb
    kmpc = k * h
    gamma_eff = om * h * (alpha + (1 - alpha) / (1 + (0.43 * kmpc * s) ** 4))
    q = k * theta * theta / np.where(gamma_eff > 0, gamma_eff, 1)
    L0 = np.log(2 * math.e + 1.8 * q)
    C0 = 14.2 + 731 / (1 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)


def linear_power(k_hmpc, c: LCDM):
    """P(k) in (Mpc/h)^3 at a=1, sigma_8-normalised."""
    k = np.asarray(k_hmpc, dtype=np.float64)
    kk = np.logspace(-5, 3, 8192)
    pk_shape = kk ** c.n_s * transfer_eh_nowiggle(kk, c) ** 2
    x = kk * 8.0
    w = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    integrand = kk ** 3 * pk_shape * w * w / (2 * np.pi ** 2)
    sig2 = np.trapezoid(integrand, np.log(kk))
    amp = c.sigma8 ** 2 / sig2
    out = np.zeros_like(k)
    pos = k > 0
    out[pos] = amp * k[pos] ** c.n_s * transfer_eh_nowiggle(k[pos], c) ** 2
    return out


def white_noise(seed, shape):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(size=shape, dtype=np.float32)


def _kgrid(shape, box, device):
    ks = []
    for a, n in enumerate(shape):
        f = np.fft.rfftfreq(n) if a == 2 else np.fft.fftfreq(n)
        ks.append(torch.as_tensor(f * (2 * np.pi * n / box[a]), dtype=torch.float32, device=device))
    kx = ks[0][:, None, None]
    ky = ks[1][None, :, None]
    kz = ks[2][None, None, :]
    return kx, ky, kz


def lpt_displacement(seed, ptcl_grid_shape, box_size, cosmo: LCDM, device="cuda", order=2, a=1.0):
    """(N,3) float32 2LPT displacement (Mpc/h) on the Lagrangian grid at scale factor ``a``.

    D1 is the linear growth normalised to 1 at a=1 (matter + Lambda, Heath integral).
    """
    shape = tuple(ptcl_grid_shape)
    box = tuple(box_size) if isinstance(box_size, (tuple, list)) else (float(box_size),) * 3
    vol = float(np.prod(box))
    ncell = int(np.prod(shape))
    noise = torch.from_numpy(white_noise(seed, shape)).to(device)
    wk = rfftn(noise)
    del noise
    kx, ky, kz = _kgrid(shape, box, device)
    k2 = kx * kx + ky * ky + kz * kz
    kmag = torch.sqrt(k2).cpu().numpy()
    pk = torch.as_tensor(linear_power(kmag, cosmo), dtype=torch.float32, device=device)
    d1 = growth_factor(a, cosmo)
    # <|delta_k|^2> = P(k) N^2 / V for delta_k = sum_x delta(x) e^{-ikx}; <|w_k|^2> = N
    delta_k = wk * torch.sqrt(pk * (ncell / vol)) * d1
    delta_k[0, 0, 0] = 0
    del wk, pk
    inv_k2 = torch.where(k2 > 0, 1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)), torch.zeros_like(k2))
    kv = (kx, ky, kz)
    disp = torch.empty(shape + (3,), dtype=torch.float32, device=device)
    # Psi1 = -grad phi1, phi1_k = -delta_k / k^2  ->  Psi1_k = i k delta_k / k^2
    for i in range(3):
        disp[..., i] = irfftn((1j * kv[i] * inv_k2) * delta_k, s=shape)
    if order >= 2:
        # phi1_ij = k_i k_j delta_k / k^2 ; source of phi2 = sum_{i<j} phi_ii phi_jj - phi_ij^2
        phi = {}
        for i in range(3):
            for j in range(i, 3):
                phi[(i, j)] = irfftn((kv[i] * kv[j] * inv_k2) * delta_k, s=shape)
        src = (phi[(0, 0)] * phi[(1, 1)] + phi[(0, 0)] * phi[(2, 2)] + phi[(1, 1)] * phi[(2, 2)]
               - phi[(0, 1)] ** 2 - phi[(0, 2)] ** 2 - phi[(1, 2)] ** 2)
        del phi
        src_k = rfftn(src)
        del src
        # grad phi2 with nabla^2 phi2 = src: phi2_k = -src_k / k^2, grad -> i k phi2_k
        d2 = -3.0 / 7.0
        for i in range(3):
            disp[..., i] += d2 * irfftn((-1j * kv[i] * inv_k2) * src_k, s=shape)
    return disp.reshape(-1, 3)


def growth_factor(a, cosmo: LCDM):
    """Linear growth D(a)/D(1) for flat LCDM (Heath 1977 integral)."""

    def d_unnorm(aa):
        om, ol = cosmo.Omega_m, 1 - cosmo.Omega_m
        e = lambda x: np.sqrt(om / x ** 3 + ol)
        xs = np.linspace(1e-6, aa, 20001)
        integ = np.trapezoid(1.0 / (xs * e(xs)) ** 3, xs)
        return 2.5 * om * e(aa) * integ

    return float(d_unnorm(a) / d_unnorm(1.0))


def particles_2lpt(conf, cosmo: LCDM, seed=0, jitter_seed=None, device="cuda", a=1.0):
    """Particles at q + U(-1/2,1/2) l_p (optional) + Psi_2LPT(q), wrapped (pmid/disp form)."""
    from .nbody.particles import Particles

    grid = Particles.gen_grid(conf, device=device)
    q = grid.pos(dtype=torch.float32, wrap=False)
    pos = q + lpt_displacement(seed, conf.ptcl_grid_shape, conf.box_size, cosmo, device=device, a=a)
    if jitter_seed is not None:
        rng = np.random.Generator(np.random.PCG64(jitter_seed))
        u = rng.uniform(-0.5, 0.5, size=(conf.ptcl_num, 3)).astype(np.float32)
        pos = pos + torch.from_numpy(u).to(device) * np.float32(conf.ptcl_spacing)
    box = torch.tensor(conf.box_size, dtype=torch.float32, device=device)
    pos = torch.remainder(pos, box)
    return Particles.from_pos(conf, pos)
