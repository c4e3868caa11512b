"""Small torch helpers: device pointers, the current HIP stream and workspaces.

PyTorch is plumbing here: it owns device memory (caching allocator) and the stream; all
arithmetic of the PM step runs in ``libpmpp_amd.so``.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib


def ptr(t):
    """Device pointer of a tensor as ctypes.c_void_p (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(
            f"{name} is on {t.device}; the pmpp_amd operators run only on a GPU (there is no CPU fallback)"
        )
    return t


def as_contig(t: torch.Tensor, dtype, name: str) -> torch.Tensor:
    require_device(t, name)
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


_initialised = set()


def ensure_init(device) -> None:
    """Per-device constant tables (twiddles); done once before first use."""
    idx = torch.device(device).index
    if idx is None:
        idx = torch.cuda.current_device()
    if idx in _initialised:
        return
    with torch.cuda.device(idx):
        _lib.call("pm_init")
    _initialised.add(idx)
