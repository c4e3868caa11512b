"""Timed CPU baseline: the reference's single-device force step restated in float32 NumPy/SciPy.

Test/measurement infrastructure only (``bench.py``'s ``cpu_baseline`` leg).  It is the same
operator sequence as ``src/pmpp/nbody/gravity.py:786-814`` (CIC scatter, 1.5 Omega_m (rho-1),
rfftn, particle-Nyquist mask, -1/k^2, -i k_i, irfftn, CIC gather), with the CIC written as
``numpy.bincount`` over the 8 corner indices and the FFTs as ``scipy.fft`` with ``workers``
threads.  The reference's own CPU path (JAX/XLA) cannot run here: jax is not installed.
"""

from __future__ import annotations

import numpy as np
import scipy.fft

from .spectral import green_multiplier, neg_k, particle_nyquist_masks


def _cic_f32(pmid, disp, cell, mesh_shape):
    n = np.asarray(mesh_shape, dtype=np.int64)
    t = disp / np.float32(cell)
    lf = np.floor(t)
    f = (t - lf).astype(np.float32)
    base = np.mod(pmid.astype(np.int64) + lf.astype(np.int64), n)
    return base, f


def _corners(base, f, mesh_shape):
    n = np.asarray(mesh_shape, dtype=np.int64)
    for c in range(8):
        b = np.array([(c >> a) & 1 for a in range(3)])
        idx = np.mod(base + b, n)
        flat = (idx[:, 0] * n[1] + idx[:, 1]) * n[2] + idx[:, 2]
        w = np.ones(base.shape[0], dtype=np.float32)
        for a in range(3):
            w = w * (f[:, a] if b[a] else (np.float32(1) - f[:, a]))
        yield flat, w


def force_step(pmid, disp, omega_m, mesh_shape, cell, ptcl_grid_shape, ptcl_spacing, workers=1):
    """One float32 PM force evaluation on the CPU; returns acc (N,3) float32."""
    mesh_shape = tuple(mesh_shape)
    nm = int(np.prod(mesh_shape))
    val = np.float32(nm / np.prod(ptcl_grid_shape))
    base, f = _cic_f32(pmid, disp, cell, mesh_shape)
    rho = np.zeros(nm, dtype=np.float64)
    for flat, w in _corners(base, f, mesh_shape):
        rho += np.bincount(flat, weights=w * val, minlength=nm)
    src = ((rho.astype(np.float32) - 1) * np.float32(1.5 * omega_m)).reshape(mesh_shape)
    src_hat = scipy.fft.rfftn(src, workers=workers)
    masks = particle_nyquist_masks(mesh_shape, ptcl_grid_shape, ptcl_spacing, cell)
    pot = src_hat * green_multiplier(mesh_shape, cell, masks).astype(np.float32)
    nk = neg_k(mesh_shape, cell)
    shapes = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    acc = np.zeros((pmid.shape[0], 3), dtype=np.float32)
    for i in range(3):
        mult = (1j * nk[i].reshape(shapes[i])).astype(np.complex64)
        F = scipy.fft.irfftn(mult * pot, s=mesh_shape, workers=workers).astype(np.float32).ravel()
        for flat, w in _corners(base, f, mesh_shape):
            acc[:, i] += F[flat] * w
    return acc
