"""Float64 NumPy restatement of PM++'s initial-condition pipeline on the particle grid.

Test infrastructure only: the product never imports this module.  Follows (read as text,
never imported):

* ``src/pmpp/initial_conditions/modes.py:14-64``    -- ``white_noise``: real standard-normal
  field, orthonormal rfftn.
* ``src/pmpp/initial_conditions/modes.py:399-454``  -- ``linear_modes``: delta(k) =
  sqrt(V P_lin(k)) omega(k).
* ``src/pmpp/cosmology/boltzmann.py:428-482``       -- ``linear_power`` with the transfer
  function interpolated from its table (``boltzmann.py:140-175``).
* ``src/pmpp/initial_conditions/lpt.py:17-43``      -- ``_strain`` (Nyquist k zeroed for
  i != j), ``:239-302`` ``_L``, ``:339-449`` ``lpt``.
* ``src/pmpp/nbody/gravity.py:263-287, 355-391``    -- ``laplace_transposed``, ``neg_grad``.

The noise generator is numpy's PCG64 (PM++'s JAX threefry stream cannot run here), which the
product uses too, so the draws are identical.  Pinned in ``tests/test_oracle_ic.py`` by the
Zel'dovich displacement of a plane wave, the closed-form 2LPT displacement of two
orthogonal plane waves, and the shell-averaged power of a noise realisation.
"""

from __future__ import annotations

import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)


def kvec_spacing(shape, spacing):
    """Angular wavenumbers of the particle grid (rfft on the last axis), float64."""
    ks = []
    for a, n in enumerate(shape):
        f = np.fft.rfftfreq(n) if a == len(shape) - 1 else np.fft.fftfreq(n)
        ks.append(f * (2 * np.pi / spacing))
    return ks


def _grid(ks):
    return ks[0][:, None, None], ks[1][None, :, None], ks[2][None, None, :]


def white_noise(seed, shape):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(size=tuple(shape), dtype=np.float32).astype(np.float64)


def ortho_rfftn(x):
    return np.fft.rfftn(x) / np.sqrt(x.size)


def linear_power(k, table_k, table_T, coef, k_pivot, n_s):
    """coef (k/k_pivot)^n_s T(k)^2, T interpolated from its table; P(0) = 0."""
    k = np.asarray(k, dtype=np.float64)
    T = np.interp(k, table_k, table_T)
    prim = np.where(k > 0, (np.where(k > 0, k, 1.0) / k_pivot) ** n_s, 0.0)
    return coef * prim * T * T


def linear_modes(w_k, shape, spacing, table_k, table_T, coef, k_pivot, n_s):
    kx, ky, kz = _grid(kvec_spacing(shape, spacing))
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    vol = float(np.prod(shape)) * spacing ** 3
    return w_k * np.sqrt(linear_power(k, table_k, table_T, coef, k_pivot, n_s) * vol)


def _nyquist_zero(k, spacing):
    nyq = np.pi / spacing
    return np.where(np.abs(np.abs(k) - nyq) <= nyq * F32_EPS, 0.0, k)


def laplace(ks, src):
    kx, ky, kz = _grid(ks)
    k2 = kx * kx + ky * ky + kz * kz
    return np.where(k2 != 0, -src / np.where(k2 == 0, 1.0, k2), 0.0)


def strain(ks, i, j, pot, shape, spacing):
    ki, kj = _grid(ks)[i], _grid(ks)[j]
    if i != j:
        ki, kj = _nyquist_zero(ki, spacing), _nyquist_zero(kj, spacing)
    return np.fft.irfftn(-ki * kj * pot, s=tuple(shape), axes=(0, 1, 2))


def second_order_source(ks, pot, shape, spacing):
    """sum_{i<j} phi_ii phi_jj - phi_ij^2 (lpt.py:239-302 with one potential)."""
    d = [strain(ks, i, i, pot, shape, spacing) for i in range(3)]
    L = d[0] * d[1] + d[0] * d[2] + d[1] * d[2]
    for i in range(2):
        for j in range(i + 1, 3):
            s = strain(ks, i, j, pot, shape, spacing)
            L = L - s * s
    return L


def neg_grad(ks, i, pot, shape, spacing):
    ki = _nyquist_zero(_grid(ks)[i], spacing)
    return np.fft.irfftn(-1j * ki * pot, s=tuple(shape), axes=(0, 1, 2))


def lpt(modes, shape, spacing, growth, order=2):
    """(disp, vel) increments (N, 3) in grid order from ``modes`` ([L^3] normalisation).

    ``growth`` = [(D_1, a^2 H dD_1/dlna), (D_2, a^2 H dD_2/dlna)]."""
    shape = tuple(shape)
    ks = kvec_spacing(shape, spacing)
    src = np.asarray(modes, dtype=np.complex128) / spacing ** 3
    pots = []
    if order > 0:
        pots.append(laplace(ks, src))
    if order > 1:
        pots.append(laplace(ks, np.fft.rfftn(second_order_source(ks, pots[0], shape, spacing))))
    n = int(np.prod(shape))
    disp = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    for m, pot in enumerate(pots):
        D, V = growth[m]
        for i in range(3):
            g = neg_grad(ks, i, pot, shape, spacing).reshape(-1)
            disp[:, i] += D * g
            vel[:, i] += V * g
    return disp, vel
