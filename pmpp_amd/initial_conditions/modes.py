"""White noise and linear modes on the particle grid (PM++ ``initial_conditions/modes.py``).

``white_noise`` draws a standard-normal real field on the host from
``numpy.random.Generator(PCG64(seed))`` -- PM++ draws from JAX's threefry stream, which cannot
be reproduced without JAX, so seed s here is this generator's seed s -- and returns its
orthonormal rFFT (``modes.py:14-64``).  ``linear_modes`` multiplies by sqrt(P_lin(k) V)
(``modes.py:399-454``) on the GPU (``pm_ic_linear_modes``): the transfer function is linearly
interpolated from the host table at ``transfer_k`` and P_lin formed per mode in float32, as
the reference evaluates it.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from .._torch import ensure_init, ptr, require_device, stream_handle
from ..cosmology import boltzmann
from ..nbody.integrator import growth_table
from ..numerics import irfftn, rfftn


class KGrid:
    """Per-axis angular wavenumbers of a grid (float32 device tables, rfft last axis)."""

    def __init__(self, shape, spacing, device):
        from ..core.configuration import _fftfreq_angular

        self.shape = tuple(int(s) for s in shape)
        self.tabs = [torch.from_numpy(_fftfreq_angular(n, spacing, a == 2).astype(np.float32)).to(device)
                     for a, n in enumerate(self.shape)]
        self.ptrs = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in self.tabs])
        nyq = np.float32(np.pi / spacing)
        self.nyquist = (ctypes.c_float * 3)(*([float(nyq)] * 3))
        self.tol = (ctypes.c_float * 3)(*([float(nyq * np.finfo(np.float32).eps)] * 3))

    def op(self, spec, op, i=0, j=0, scale=1.0, out=None):
        """symbol(k) * spec: 0 scale, 1 Laplace inverse, 2 strain -k_i k_j, 3 -i k_i."""
        if out is None:
            out = torch.empty_like(spec)
        _lib.call("pm_ic_spectral_op", ptr(spec), ptr(out), _lib.int3(self.shape), self.ptrs, int(op), int(i), int(j),
                  float(scale), self.nyquist, self.tol, stream_handle(spec.device))
        return out


SCALE, LAPLACE, STRAIN, NEG_GRAD = 0, 1, 2, 3


def white_noise(seed, conf, real=False, device="cuda"):
    """Standard-normal noise on the particle grid: real (float32) or its orthonormal rFFT."""
    shape = tuple(conf.ptcl_grid_shape)
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = torch.from_numpy(rng.standard_normal(size=shape, dtype=np.float32)).to(device)
    if real:
        return noise
    ensure_init(noise.device)
    kg = KGrid(shape, conf.ptcl_spacing, noise.device)
    return kg.op(rfftn(noise), SCALE, scale=1.0 / math.sqrt(math.prod(shape)))


def linear_modes(modes, cosmo, conf, a=None, real=False):
    """Linear overdensity modes delta(k) = sqrt(V P_lin(k, a)) omega(k) (PM++'s [L^3]
    normalisation); ``a = None`` leaves out the growth factor.  Real ``modes`` are treated as
    white noise in real space."""
    require_device(modes, "modes")
    dev = modes.device
    ensure_init(dev)
    shape = tuple(conf.ptcl_grid_shape)
    kg = KGrid(shape, conf.ptcl_spacing, dev)
    if not torch.is_complex(modes):
        modes = kg.op(rfftn(modes), SCALE, scale=1.0 / math.sqrt(math.prod(shape)))
    else:
        modes = modes.to(torch.complex64).contiguous().clone()
    if tuple(modes.shape) != (shape[0], shape[1], shape[2] // 2 + 1):
        raise ValueError("modes do not match the particle grid")
    D = 1.0
    if a is not None:
        D = float(growth_table(cosmo, conf).growth(float(a)))
    tk, tT = boltzmann.transfer_table(cosmo)
    tk_d = torch.from_numpy(tk.astype(np.float32)).to(dev)
    tT_d = torch.from_numpy(tT.astype(np.float32)).to(dev)
    coef = boltzmann.power_coef(cosmo, cosmo.A_s_1e9) * D * D
    _lib.call("pm_ic_linear_modes", ptr(modes), _lib.int3(shape), kg.ptrs, ptr(tk_d), ptr(tT_d), len(tk), float(coef),
              float(boltzmann.k_pivot(cosmo)), float(cosmo.n_s), float(math.sqrt(conf.box_vol)),
              stream_handle(dev))
    if real:
        return irfftn(modes, s=shape)
    return modes
