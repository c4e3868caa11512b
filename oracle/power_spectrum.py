"""Float64 NumPy restatement of PM++'s integer-shell power spectrum (the P(k) metric).

Follows ``src/pmpp/analysis/power_spectrum.py`` (read as text, never imported):

* ``:46-71``   -- shell statistics: shell = floor(|n| + 1e-6), Hermitian multiplicity
  (kz = 0 and the even-N Nyquist plane weigh 1, all others 2), shell-mean |k|.
* ``:108-143`` -- power |delta_k|^2 with the CIC window deconvolution sinc(n/N)^(-4).
* ``:251-256`` -- normalisation P = sum / nmodes * cell^3 / N_m, shells from 1 upward.
* ``:348-374`` -- ``density_to_pk``: delta = rho / mean(rho) - 1.

Pinned in ``tests/test_oracle_power_spectrum.py`` by the exact recovery of a single
Fourier mode of known amplitude (all power in one shell with the closed-form value) and
by Parseval's identity sum_k |delta_k|^2 = N_m sum_x delta_x^2 when no window is removed.
"""

from __future__ import annotations

import math

import numpy as np

_MAS_POWER = {None: 0, "NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}


def mode_numbers(shape):
    mx = np.fft.fftfreq(shape[0]) * shape[0]
    my = np.fft.fftfreq(shape[1]) * shape[1]
    mz = np.fft.rfftfreq(shape[2]) * shape[2]
    return mx, my, mz


def shell_statistics(shape, cell_size):
    """(k_shell[1:], nmodes[1:]) as in power_spectrum.py:46-71."""
    box = float(shape[0]) * float(cell_size)
    kf = 2 * math.pi / box
    mx, my, mz = mode_numbers(shape)
    X, Y, Z = np.meshgrid(mx, my, mz, indexing="ij")
    mag = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    shell = np.floor(mag + 1e-6).astype(np.int64)
    mult = np.full(shell.shape, 2.0)
    mult[..., 0] = 1
    if shape[2] % 2 == 0:
        mult[..., -1] = 1
    ns = int(shell.max()) + 1
    nmodes = np.bincount(shell.ravel(), weights=mult.ravel(), minlength=ns)
    ksum = np.bincount(shell.ravel(), weights=(mag * mult).ravel(), minlength=ns)
    k = ksum / np.maximum(nmodes, 1) * kf
    return k[1:], nmodes[1:]


def delta_to_pk(delta, cell_size, mas="CIC"):
    delta = np.asarray(delta, dtype=np.float64)
    shape = delta.shape
    spec = np.fft.rfftn(delta)
    mx, my, mz = mode_numbers(shape)
    msq = mx[:, None, None] ** 2 + my[None, :, None] ** 2 + mz[None, None, :] ** 2
    num_shells = math.floor(math.sqrt(3) * (shape[0] // 2)) + 1
    shell = np.clip(np.floor(np.sqrt(msq) + 1e-9).astype(np.int64), 0, num_shells - 1)
    mult = np.full(spec.shape, 2.0)
    mult[..., 0] = 1
    if shape[2] % 2 == 0:
        mult[..., -1] = 1
    power = spec.real ** 2 + spec.imag ** 2
    p = _MAS_POWER[mas.upper() if mas else None]
    if p:
        w = (np.sinc(mx / shape[0])[:, None, None] * np.sinc(my / shape[1])[None, :, None]
             * np.sinc(mz / shape[2])[None, None, :])
        power = power * w ** (-2 * p)
    psum = np.bincount(shell.ravel(), weights=(power * mult).ravel(), minlength=num_shells)
    k, nmodes = shell_statistics(shape, cell_size)
    n = min(len(k), num_shells - 1)
    pk = psum[1:1 + n] / nmodes[:n] * (cell_size ** 3 / np.prod(shape))
    return k[:n], pk, nmodes[:n]


def power_spectrum(density, cell_size, mas="CIC"):
    """``density_to_pk``: delta = rho/mean - 1, then the shell-averaged P(k)."""
    density = np.asarray(density, dtype=np.float64)
    delta = density / density.mean() - 1.0
    return delta_to_pk(delta, cell_size, mas)
