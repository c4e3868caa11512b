"""BF16-MFMA periodic conv3d kernels vs a plain PyTorch float32 reference (CPU).

Inputs are bf16-representable, so products are exact and only float32 summation order
differs: tolerance 1e-4 relative to the output scale.
"""

import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

from pmpp_amd.corrections import conv3d, conv3d_bwd_data, conv3d_bwd_weight
from pmpp_amd.corrections.conv import conv3d_bwd_data_out1

pytestmark = pytest.mark.gpu


def _ref_conv(x, w, b):
    """x (X,Y,Z,C) float32, w (27,16,C) -> (X,Y,Z,16) periodic conv, float32 on the CPU."""
    xt = x.permute(3, 0, 1, 2)[None]
    wt = w.reshape(3, 3, 3, 16, -1).permute(3, 4, 0, 1, 2)  # (co, ci, kx, ky, kz), tap = (dx+1)*9+...
    xp = Fn.pad(xt, (1, 1, 1, 1, 1, 1), mode="circular")
    y = Fn.conv3d(xp, wt)  # cross-correlation: out[v] = sum w[off] x[v + off]
    return y[0].permute(1, 2, 3, 0) + b


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("cin,shape,gelu", [(8, (8, 8, 8), False), (16, (16, 8, 24), True), (8, (32, 16, 16), True)])
def test_forward_matches_torch(device, cin, shape, gelu):
    g = torch.Generator().manual_seed(cin + sum(shape))
    x = _bf(torch.randn(shape + (cin,), generator=g))
    w = _bf(torch.randn(27, 16, cin, generator=g) * 0.2)
    b = torch.randn(16, generator=g)
    res = _bf(torch.randn(shape + (16,), generator=g))
    pre, act = conv3d(x.to(device, torch.bfloat16), w.to(device, torch.bfloat16), b.to(device),
                      res=res.to(device, torch.bfloat16), gelu=gelu)
    ref = _ref_conv(x, w, b) + res
    scale = ref.abs().max().item()
    np.testing.assert_allclose(pre.cpu().numpy(), ref.numpy(), atol=1e-4 * scale)
    r_act = Fn.gelu(ref, approximate="tanh") if gelu else ref
    np.testing.assert_allclose(act.float().cpu().numpy(), _bf(r_act).numpy(), atol=1e-2 * scale)
    # a single stored channel is channel 0 of the full output, bit for bit
    pre1, _ = conv3d(x.to(device, torch.bfloat16), w.to(device, torch.bfloat16), b.to(device),
                     res=res.to(device, torch.bfloat16), gelu=gelu, want_act=False, pre_channels=1)
    assert pre1.shape == shape + (1,)
    np.testing.assert_array_equal(pre1[..., 0].cpu().numpy(), pre[..., 0].cpu().numpy())


@pytest.mark.parametrize("shape", [(128, 64, 64), (160, 64, 40)])
@pytest.mark.parametrize("cin,res", [(16, True), (8, False)])
def test_forward_persistent_grid_walks_every_block(device, cin, res, shape):
    """Meshes of 1024 blocks of 4 x 8 x 16 (z a multiple of 16) and of 800 blocks of 8^3 (z
    not): more blocks than workgroups resident on the device, so each workgroup of the
    persistent grid takes several blocks (and the grid a ragged tail), in both block shapes."""
    g = torch.Generator().manual_seed(11 + cin)
    x = _bf(torch.randn(shape + (cin,), generator=g))
    w = _bf(torch.randn(27, 16, cin, generator=g) * 0.2)
    b = torch.randn(16, generator=g)
    r = _bf(torch.randn(shape + (16,), generator=g)) if res else None
    pre, act = conv3d(x.to(device, torch.bfloat16), w.to(device, torch.bfloat16), b.to(device),
                      res=r.to(device, torch.bfloat16) if res else None, gelu=True)
    ref = _ref_conv(x, w, b) + (r if res else 0)
    scale = ref.abs().max().item()
    np.testing.assert_allclose(pre.cpu().numpy(), ref.numpy(), atol=1e-4 * scale)
    np.testing.assert_allclose(act.float().cpu().numpy(), _bf(Fn.gelu(ref, approximate="tanh")).numpy(),
                               atol=1e-2 * scale)


@pytest.mark.parametrize("shape", [(16, 16, 8), (8, 16, 32), (24, 8, 48)])
def test_backward_data_and_weight_match_autograd(device, shape):
    # (16, 16, 8) walks 8^3 blocks, the others 4 x 8 x 16 blocks (x a multiple of 4, z of 16)
    g = torch.Generator().manual_seed(3)
    x = _bf(torch.randn(shape + (16,), generator=g))
    w = _bf(torch.randn(27, 16, 16, generator=g) * 0.2)
    dz = _bf(torch.randn(shape + (16,), generator=g))
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    br = torch.zeros(16, requires_grad=True)
    y = _ref_conv(xr, wr, br)
    y.backward(dz)
    dx = conv3d_bwd_data(dz.to(device), w.to(device, torch.bfloat16))
    np.testing.assert_allclose(dx.cpu().numpy(), xr.grad.numpy(), atol=1e-4 * xr.grad.abs().max().item())
    dw, db = conv3d_bwd_weight(dz.to(device), x.to(device, torch.bfloat16))
    np.testing.assert_allclose(dw.cpu().numpy(), wr.grad.numpy(), atol=1e-4 * wr.grad.abs().max().item())
    np.testing.assert_allclose(db.cpu().numpy(), br.grad.numpy(), atol=1e-4 * br.grad.abs().max().item())
    # the fused gelu' and residual terms of the data gradient
    gact = torch.randn(shape + (16,), generator=g)
    add = torch.randn(shape + (16,), generator=g)
    dx2 = conv3d_bwd_data(dz.to(device), w.to(device, torch.bfloat16), add=add.to(device), gact=gact.to(device))
    t = torch.tanh(0.7978845608028654 * (gact + 0.044715 * gact ** 3))
    gg = 0.5 * (1 + t) + 0.5 * gact * (1 - t * t) * 0.7978845608028654 * (1 + 3 * 0.044715 * gact ** 2)
    np.testing.assert_allclose(dx2.cpu().numpy(), ((xr.grad + add) * gg).numpy(), rtol=1e-4, atol=1e-4)
    # only the first four channels (the input path's): the same values
    dx4 = conv3d_bwd_data(dz.to(device), w.to(device, torch.bfloat16), add=add.to(device), gact=gact.to(device),
                          channels=4)
    assert tuple(dx4.shape) == shape + (4,)
    np.testing.assert_array_equal(dx4.cpu().numpy(), dx2[..., :4].cpu().numpy())
    # bf16 dz and add (dz is bf16-valued already, as the matrix cores read it): the same sums,
    # with dx float32 or rounded to bf16
    addb = _bf(add)
    wd = w.to(device, torch.bfloat16)
    ref32 = conv3d_bwd_data(dz.to(device), wd, add=addb.to(device), gact=gact.to(device))
    got32 = conv3d_bwd_data(dz.to(device, torch.bfloat16), wd, add=addb.to(device, torch.bfloat16),
                            gact=gact.to(device))
    np.testing.assert_array_equal(got32.cpu().numpy(), ref32.cpu().numpy())
    got16 = conv3d_bwd_data(dz.to(device, torch.bfloat16), wd, add=addb.to(device, torch.bfloat16),
                            gact=gact.to(device), dtype=torch.bfloat16)
    assert got16.dtype == torch.bfloat16
    np.testing.assert_array_equal(got16.float().cpu().numpy(), ref32.to(torch.bfloat16).float().cpu().numpy())


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 24, 32)])
def test_one_channel_layer_data_gradient(device, shape):
    """pm_conv3d_bwd_data_out1 (the data gradient of the one-output-channel layer from its
    one-channel gradient) against autograd of the float32 conv with the same bf16 operands, and
    against the 16 -> 16 data gradient of a dz whose channel 0 is that gradient."""
    g = torch.Generator().manual_seed(6)
    w = _bf(torch.randn(27, 16, 16, generator=g) * 0.2)
    dz0 = _bf(torch.randn(shape, generator=g))
    gact = torch.randn(shape + (16,), generator=g)
    xr = torch.zeros(shape + (16,), requires_grad=True)
    w1 = torch.zeros(27, 16, 16)
    w1[:, 0, :] = w[:, 0, :]
    y = _ref_conv(xr, w1, torch.zeros(16))
    dz16 = torch.zeros(shape + (16,))
    dz16[..., 0] = dz0
    y.backward(dz16)
    t = torch.tanh(0.7978845608028654 * (gact + 0.044715 * gact ** 3))
    gg = 0.5 * (1 + t) + 0.5 * gact * (1 - t * t) * 0.7978845608028654 * (1 + 3 * 0.044715 * gact ** 2)
    got = conv3d_bwd_data_out1(dz0.to(device), w.to(device, torch.bfloat16), gact=gact.to(device)).cpu()
    ref = xr.grad * gg
    np.testing.assert_allclose(got.numpy(), ref.numpy(), rtol=1e-4, atol=1e-5 * ref.abs().max().item())
    wide = conv3d_bwd_data(dz16.to(device), w.to(device, torch.bfloat16), gact=gact.to(device)).cpu()
    np.testing.assert_allclose(got.numpy(), wide.numpy(), rtol=1e-5, atol=1e-6 * wide.abs().max().item())
    got16 = conv3d_bwd_data_out1(dz0.to(device), w.to(device, torch.bfloat16), gact=gact.to(device),
                                 dtype=torch.bfloat16).cpu()
    np.testing.assert_array_equal(got16.float().numpy(), got.to(torch.bfloat16).float().numpy())


def test_weight_gradient_cin8(device):
    shape = (8, 16, 16)
    g = torch.Generator().manual_seed(4)
    x = _bf(torch.randn(shape + (8,), generator=g))
    w = _bf(torch.randn(27, 16, 8, generator=g) * 0.2)
    dz = _bf(torch.randn(shape + (16,), generator=g))
    wr = w.clone().requires_grad_(True)
    _ref_conv(x, wr, torch.zeros(16)).backward(dz)
    dw, db = conv3d_bwd_weight(dz.to(device), x.to(device, torch.bfloat16))
    np.testing.assert_allclose(dw.cpu().numpy(), wr.grad.numpy(), atol=1e-4 * wr.grad.abs().max().item())


@pytest.mark.parametrize("shape", [(8, 8, 16), (24, 16, 32), (32, 16, 48), (256, 32, 64), (256, 64, 128)])
def test_one_channel_output_layer_matches_torch(device, shape):
    """The one-output-channel layer (taps on the matrix rows, k_conv_out1: meshes that divide
    into 8 x 8 x 16 blocks) against channel 0 of the float32 convolution, on meshes of one
    block along y and z (every y and z halo wraps) and of several blocks along each axis, up to
    meshes of more blocks than the persistent grid has workgroups."""
    g = torch.Generator().manual_seed(sum(shape))
    x = _bf(torch.randn(shape + (16,), generator=g))
    w = _bf(torch.randn(27, 16, 16, generator=g) * 0.2)
    b = torch.randn(16, generator=g)
    pre1, act = conv3d(x.to(device, torch.bfloat16), w.to(device, torch.bfloat16), b.to(device), want_act=False,
                       pre_channels=1)
    assert act is None and pre1.shape == shape + (1,)
    ref = _ref_conv(x, w, b)[..., 0]
    scale = ref.abs().max().item()
    np.testing.assert_allclose(pre1[..., 0].cpu().numpy(), ref.numpy(), atol=1e-4 * scale)
