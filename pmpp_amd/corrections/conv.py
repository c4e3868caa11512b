"""Thin wrappers of the BF16-MFMA periodic 3x3x3 convolution kernels (``pm_conv3d_*``).

Activations are channels-last torch tensors (X, Y, Z, C); inputs to a layer are bf16 with 8 or
16 channels, weights bf16 (27, 16, C_in) with tap = (dx+1)*9 + (dy+1)*3 + (dz+1), biases
float32 (16,).  Mesh axes must be multiples of 8.  The layers are those of the reference's
periodic Conv3D stack (corrections/mesh_cnn.py:54-105), with float32 accumulation.
"""

from __future__ import annotations

import torch

from .. import _lib
from .._torch import ensure_init, ptr, require_device, stream_handle, workspace


def _shape(x):
    if x.ndim != 4:
        raise ValueError("conv3d takes channels-last (X, Y, Z, C) meshes")
    return tuple(int(s) for s in x.shape[:3])


def conv3d(x, w, b, res=None, gelu=False, want_pre=True, want_act=True, pre_channels=16):
    """(pre, act): pre = conv(x) + b (+ res) float32 (its first ``pre_channels`` channels: 16 or
    1), act = gelu(pre) (or pre) bf16."""
    require_device(x, "x")
    if x.dtype != torch.bfloat16 or w.dtype != torch.bfloat16:
        raise RuntimeError("conv3d takes bf16 activations and weights")
    ensure_init(x.device)
    shape = _shape(x)
    cin = int(x.shape[3])
    if pre_channels not in (1, 16):
        raise ValueError("pre_channels must be 1 or 16")
    pre = torch.empty(shape + (pre_channels,), dtype=torch.float32, device=x.device) if want_pre else None
    act = torch.empty(shape + (16,), dtype=torch.bfloat16, device=x.device) if want_act else None
    _lib.call("pm_conv3d_fwd", ptr(x.contiguous()), cin, ptr(w.contiguous()), ptr(b.contiguous()),
              ptr(res.contiguous()) if res is not None else None, int(gelu), ptr(pre), int(pre_channels), ptr(act),
              _lib.int3(shape), stream_handle(x.device))
    return pre, act


def conv3d_bwd_data(dz, w, add=None, gact=None, channels=16, dtype=torch.float32):
    """dx = conv^T(dz; w) (+ add), times gelu'(gact) if given (16 -> 16 layers); ``channels=4``
    returns only the first four channels of dx (an (X, Y, Z, 4) mesh).  dz (and add) may be
    float32 or bf16; ``dtype`` is dx's (bf16 only from a bf16 dz)."""
    shape = _shape(dz)
    if dz.dtype not in (torch.float32, torch.bfloat16):
        raise ValueError("dz must be float32 or bf16")
    if add is not None and add.dtype != dz.dtype:
        raise ValueError("add must have dz's dtype")
    bf16_io = (1 if dz.dtype == torch.bfloat16 else 0) | (2 if dtype == torch.bfloat16 else 0)
    if bf16_io == 2:
        raise ValueError("a bf16 dx needs a bf16 dz")
    dx = torch.empty(shape + (channels,), dtype=dtype, device=dz.device)
    _lib.call("pm_conv3d_bwd_data", ptr(dz.contiguous()), ptr(w.contiguous()),
              ptr(add.contiguous()) if add is not None else None, ptr(gact.contiguous()) if gact is not None else None,
              ptr(dx), int(channels), bf16_io, _lib.int3(shape), stream_handle(dz.device))
    return dx


def conv3d_bwd_data_out1(dz0, w, gact=None, dtype=torch.float32):
    """dx (X, Y, Z, 16) of a one-output-channel layer (weights [27][16][16], channel 0) from its
    one-channel gradient dz0 (X, Y, Z), times gelu'(gact) if given; ``dtype`` float32 or bf16."""
    if dz0.ndim != 3 or dz0.dtype != torch.float32:
        raise ValueError("dz0 must be a float32 (X, Y, Z) mesh")
    shape = tuple(dz0.shape)
    dx = torch.empty(shape + (16,), dtype=dtype, device=dz0.device)
    _lib.call("pm_conv3d_bwd_data_out1", ptr(dz0.contiguous()), ptr(w.contiguous()),
              ptr(gact.contiguous()) if gact is not None else None, ptr(dx), int(dtype == torch.bfloat16),
              _lib.int3(shape), stream_handle(dz0.device))
    return dx


def conv3d_bwd_weight(dz, x, out=None):
    """(dW (27, 16, C_in) float32, db (16,) float32) of a layer with input x.

    ``out`` (optional) is a contiguous float32 tensor of 27*16*C_in + 16 elements that receives
    [dW, db] in place (a slice of a flat gradient vector)."""
    shape = _shape(dz)
    cin = int(x.shape[3])
    nbytes = _lib.size_query("pm_conv3d_bwd_weight_workspace", _lib.int3(shape), cin)
    ws = workspace(nbytes, dz.device)
    if out is None:
        out = torch.empty(27 * 16 * cin + 16, dtype=torch.float32, device=dz.device)
    elif out.numel() != 27 * 16 * cin + 16 or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError("out must be a contiguous float32 tensor of 27*16*C_in + 16 elements")
    _lib.call("pm_conv3d_bwd_weight", ptr(dz.contiguous()), ptr(x.contiguous()), cin, _lib.int3(shape), ptr(out),
              ptr(ws), nbytes, stream_handle(dz.device))
    return out[:-16].view(27, 16, cin), out[-16:]
