"""3D real FFTs and the fused spectral force multiplier on the HIP library.

``rfftn``/``irfftn`` replace the ``jnp.fft.rfftn``/``irfftn`` calls of the force path
(``src/pmpp/nbody/gravity.py:436, 528``; conventions of ``src/pmpp/numerics/fft.py:47-155``:
unnormalised forward, 1/N inverse).  Shapes: power-of-two axes in [2, 2048].
"""

from __future__ import annotations

import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, stream_handle, workspace


def rfftn(x: torch.Tensor) -> torch.Tensor:
    if x.ndim != 3:
        raise NotImplementedError("HIP rfftn supports 3-D real meshes")
    xin = as_contig(x, torch.float32, "x")
    ensure_init(x.device)
    shape = _lib.int3(x.shape)
    out = torch.empty((x.shape[0], x.shape[1], x.shape[2] // 2 + 1), dtype=torch.complex64, device=x.device)
    _lib.call("pm_rfft3d", ptr(xin), ptr(out), shape, None, 0, stream_handle(x.device))
    return out


def irfftn(xk: torch.Tensor, s=None) -> torch.Tensor:
    if xk.ndim != 3:
        raise NotImplementedError("HIP irfftn supports 3-D spectra")
    if s is None:
        s = (xk.shape[0], xk.shape[1], 2 * (xk.shape[2] - 1))
    if tuple(xk.shape) != (s[0], s[1], s[2] // 2 + 1):
        raise ValueError(f"spectrum shape {tuple(xk.shape)} does not match real shape {tuple(s)}")
    scratch = as_contig(xk, torch.complex64, "xk").clone()  # the inverse uses its input as scratch
    ensure_init(xk.device)
    out = torch.empty(tuple(s), dtype=torch.float32, device=xk.device)
    _lib.call("pm_irfft3d", ptr(scratch), ptr(out), _lib.int3(s), None, 0, stream_handle(xk.device))
    return out


def spectral_force_multiply(src_hat: torch.Tensor, desc, axis: int, adjoint: bool = False) -> torch.Tensor:
    """mask * (-1/k^2) * (-i k_axis) * src_hat (axis=-1: potential only); see pm_spectral_force_multiply."""
    src = as_contig(src_hat, torch.complex64, "src_hat")
    ensure_init(src.device)
    nbytes = _lib.size_query("pm_spectral_workspace", desc)
    ws = workspace(nbytes, src.device)
    out = torch.empty_like(src)
    _lib.call("pm_spectral_force_multiply", ptr(src), ptr(out), desc, int(axis), int(bool(adjoint)), ptr(ws),
              nbytes, stream_handle(src.device))
    return out
