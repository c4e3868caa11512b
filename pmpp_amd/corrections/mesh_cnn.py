"""Mesh-CNN correction of the PM force on the HIP library (PM++ ``corrections/mesh_cnn.py``).

The network is the reference's periodic residual CNN (``mesh_cnn.py:54-105``): a 3^3 stem
convolution with GELU, ``depth - 1`` residual blocks (conv, GELU, conv, + input, GELU), a
zero-initialised 3^3 output convolution bounded by ``max_residual * tanh``, and a
conditioning-only "direct gain" ``max_residual * tanh(Linear([a, Omega_m, sigma8]))``.  With
``block="plain"`` each block is one GELU convolution instead (no residual), so ``depth=2``
is the three-layer network stem -> hidden -> output.  The MI355X build differs in three
deliberate ways:

* hidden width is 16 channels (one MFMA tile of output channels) and activations are bf16
  with float32 accumulation (``pm_conv3d_*``);
* the input channels are [S / sigma, a, Omega_m, sigma8] padded to 8; with ``potential=True``
  the reference's potential channel is added, [S / sigma, phi / sigma_phi, a, Omega_m, sigma8]
  (``mesh_cnn.py:20-27, 76-84``), phi being the force step's own filtered potential of S
  (``gravity_potential``, one extra banded r2c/c2r pair per force);
* the residual is applied to the source of the Poisson solve, not to the potential:
  ``S' = (1 + g) S + max_residual tanh(o) sigma`` enters ``pm_gravity_corrected`` as the
  density residual ``rho_add = (S' - S) / (1.5 Omega_m)``.  The reference's ``g phi`` term is
  exactly ``g S`` through the linear Poisson solve; the convolutional term is learned in
  source units instead of potential units.

Parameters live in one flat float32 device vector (with a bf16 mirror the matrix cores read),
laid out layer by layer as ``[W (27*16*C_in), b (16)]`` at 64-element-aligned offsets, then the
four gain parameters.  The backward pass (``residual_backward``) returns the gradient in the
same layout for the optimizer (``training.py``).
"""

from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import torch

from .. import _lib
from .._torch import ensure_init, ptr, require_device, stream_handle, workspace
from .conv import conv3d, conv3d_bwd_data, conv3d_bwd_data_out1, conv3d_bwd_weight

CHANNELS = 16
IN_CHANNELS = 8
LIVE_IN_CHANNELS = 4  # S / sigma, a, Omega_m, sigma8 (5 with the potential channel)
TAPS = 27


def _align(n: int, a: int = 64) -> int:
    return (n + a - 1) // a * a


@dataclasses.dataclass(frozen=True)
class Layer:
    name: str
    cin: int
    offset: int  # first weight element; the bias follows the 27*16*cin weights

    @property
    def nweights(self) -> int:
        return TAPS * CHANNELS * self.cin

    @property
    def size(self) -> int:
        return self.nweights + CHANNELS


BLOCKS = ("residual", "plain")
SYMMETRIES = ("none", "cubic")

# taps of a 3^3 kernel grouped by their number of nonzero offsets (centre, faces, edges,
# corners): the orbits of the cube's 48 rotations and reflections acting on the taps
_TAP_ORBITS = [[t for t in range(27) if sum(((t // 9) % 3 != 1, (t // 3) % 3 != 1, t % 3 != 1)) == o]
               for o in range(4)]


def project_cubic(vec: torch.Tensor, layers) -> torch.Tensor:
    """Projects the convolution kernels stored in the flat parameter-layout vector ``vec`` (in
    place) onto the kernels that are invariant under the symmetries of the cube: each
    (output, input) channel pair's 27 taps are replaced by their mean over each orbit.  A
    network of such kernels with pointwise activations commutes with the cube's rotations and
    reflections, as the PM force's window does; biases and the gain stay as they are."""
    for layer in layers:
        w = vec[layer.offset:layer.offset + layer.nweights].view(TAPS, -1)
        for orbit in _TAP_ORBITS:
            idx = torch.tensor(orbit, device=vec.device)
            w[idx] = w[idx].mean(0, keepdim=True)
    return vec


def param_layout(depth: int, block: str = "residual"):
    """(conv layers, gain offset, total size) of the flat parameter vector."""
    if depth < 1:
        raise ValueError("depth must be >= 1")
    if block not in BLOCKS:
        raise ValueError(f"block must be one of {BLOCKS}")
    layers: List[Layer] = []
    off = 0

    def add(name, cin):
        nonlocal off
        layers.append(Layer(name, cin, off))
        off = _align(off + TAPS * CHANNELS * cin + CHANNELS)

    add("stem", IN_CHANNELS)
    for b in range(depth - 1):
        add(f"block_{b}_conv1", CHANNELS)
        if block == "residual":
            add(f"block_{b}_conv2", CHANNELS)
    add("out", CHANNELS)
    gain = off
    return layers, gain, _align(gain + 4)


@dataclasses.dataclass
class MeshCNNPotentialCorrection:
    """Trainable mesh correction (``mesh_cnn.py:136-160``): parameters plus hyperparameters."""

    params: torch.Tensor
    depth: int = 2
    max_residual: float = 0.1
    output_init_scale: float = 1e-2
    sigma8_value: float = 0.8
    params_bf16: Optional[torch.Tensor] = None
    block: str = "residual"
    symmetry: str = "none"
    potential: bool = False

    def __post_init__(self):
        require_device(self.params, "params")
        if self.symmetry not in SYMMETRIES:
            raise ValueError(f"symmetry must be one of {SYMMETRIES}")
        layers, gain, total = param_layout(self.depth, self.block)
        if self.params.dtype != torch.float32 or self.params.numel() != total or not self.params.is_contiguous():
            raise ValueError(f"params must be a contiguous float32 vector of {total} elements")
        self.layers = layers
        self.gain_offset = gain
        if self.params_bf16 is None:
            self.params_bf16 = torch.empty(total, dtype=torch.bfloat16, device=self.params.device)
            self.sync_bf16()

    def sync_bf16(self):
        """Refresh the bf16 mirror after the float32 parameters changed outside the optimizer."""
        ensure_init(self.params.device)
        _lib.call("pm_to_bf16", ptr(self.params), self.params.numel(), ptr(self.params_bf16),
                  stream_handle(self.params.device))

    # views into the flat vectors
    def weight(self, layer: Layer) -> torch.Tensor:
        return self.params_bf16[layer.offset:layer.offset + layer.nweights]

    def bias(self, layer: Layer) -> torch.Tensor:
        return self.params[layer.offset + layer.nweights:layer.offset + layer.size]

    @property
    def gain_params(self) -> torch.Tensor:
        return self.params[self.gain_offset:self.gain_offset + 4]


def _truncated_lecun(gen, fan_in, n):
    # VarianceScaling(1, "fan_in", "truncated_normal"): a normal truncated at two standard
    # deviations, rescaled so that the truncated distribution has variance 1 / fan_in
    std = math.sqrt(1.0 / fan_in) / 0.87962566103423978
    t = torch.empty(n, dtype=torch.float32)
    torch.nn.init.trunc_normal_(t, mean=0.0, std=std, a=-2 * std, b=2 * std, generator=gen)
    return t


def initial_params(seed=0, depth=2, output_init_scale=1e-2, block="residual", potential=False) -> torch.Tensor:
    """The flat float32 initial parameters on the host (see ``init_mesh_cnn_potential_correction``)."""
    layers, gain, total = param_layout(depth, block)
    stem_live = LIVE_IN_CHANNELS + (1 if potential else 0)
    gen = torch.Generator().manual_seed(int(seed))
    host = torch.zeros(total, dtype=torch.float32)
    for layer in layers:
        if layer.name == "out":
            continue
        live = stem_live if layer.name == "stem" else CHANNELS
        w = _truncated_lecun(gen, TAPS * live, TAPS * CHANNELS * layer.cin).view(TAPS, CHANNELS, layer.cin)
        w[:, :, live:] = 0.0
        host[layer.offset:layer.offset + layer.nweights] = w.reshape(-1)
    host[gain:gain + 3] = torch.randn(3, generator=gen) * output_init_scale
    return host


def init_mesh_cnn_potential_correction(seed=0, depth=2, max_residual=0.1, output_init_scale=1e-2, sigma8_value=0.8,
                                       device="cuda", block="residual", symmetry="none", potential=False):
    """A freshly initialised correction (``mesh_cnn.py:163-214``): LeCun truncated-normal
    convolutions with zero biases, a zero output layer and a small random gain layer, so the
    initial correction is exactly zero.  ``symmetry="cubic"`` projects the kernels onto the
    cube-symmetric ones (``project_cubic``), rescaled to the unprojected initial variance, and
    training keeps them there.  ``potential=True`` adds the potential input channel."""
    host = initial_params(seed, depth, output_init_scale, block, potential)
    if symmetry == "cubic":
        layers, _, _ = param_layout(depth, block)
        for layer in layers:
            w = host[layer.offset:layer.offset + layer.nweights].clone()
            std0 = float(w.std())
            project_cubic(host, [layer])
            ws = host[layer.offset:layer.offset + layer.nweights]
            if std0 > 0 and float(ws.std()) > 0:
                ws.mul_(std0 / float(ws.std()))
    return MeshCNNPotentialCorrection(params=host.to(device), depth=depth, max_residual=max_residual,
                                      output_init_scale=output_init_scale, sigma8_value=sigma8_value, block=block,
                                      symmetry=symmetry, potential=potential)


def with_potential_channel(model: MeshCNNPotentialCorrection) -> MeshCNNPotentialCorrection:
    """The same network reading the potential channel as well: the stem's weights of the
    constant channels (a, Omega_m, sigma8) move from input channels 1..3 to 2..4 and the
    potential's (channel 1) start at zero, so it computes what ``model`` does until those are
    trained (the starting point of a fine-tuning with the potential)."""
    if model.potential:
        return model
    stem = model.layers[0]
    p = model.params.detach().clone()
    w = p[stem.offset:stem.offset + stem.nweights].view(TAPS, CHANNELS, IN_CHANNELS)
    w2 = torch.zeros_like(w)
    w2[:, :, 0] = w[:, :, 0]
    w2[:, :, 2:2 + LIVE_IN_CHANNELS - 1] = w[:, :, 1:LIVE_IN_CHANNELS]
    p[stem.offset:stem.offset + stem.nweights] = w2.reshape(-1)
    return MeshCNNPotentialCorrection(params=p, depth=model.depth, max_residual=model.max_residual,
                                      output_init_scale=model.output_init_scale, sigma8_value=model.sigma8_value,
                                      block=model.block, symmetry=model.symmetry, potential=True)


@dataclasses.dataclass
class Tape:
    """What the backward pass needs from a forward evaluation."""

    rho: torch.Tensor
    stats: torch.Tensor
    x0: torch.Tensor
    pres: list  # float32 pre-activations of stem and block layers
    acts: list  # bf16 activations
    out: torch.Tensor
    a: float
    omega_m: float
    phi: Optional[torch.Tensor] = None  # the potential channel's mesh (None: no such channel)
    conf: object = None  # the configuration the potential was computed with


def _workspace(dev):
    return workspace(_lib.size_query("pm_correction_workspace"), dev)


@dataclasses.dataclass
class OutputTape:
    """The forward's statistics and output-layer pre-activation, without the activations
    (what ``network_input_vjp`` needs from the whole mesh)."""

    stats: torch.Tensor
    out: torch.Tensor
    phi: Optional[torch.Tensor] = None
    conf: object = None


def _potential(model, rho, om, conf, phi):
    """The potential channel's mesh for ``rho``: ``phi`` when given (a slab of the whole mesh's),
    else the filtered potential of the whole mesh ``rho`` (``gravity_potential``)."""
    if not model.potential:
        return None
    if phi is not None:
        if tuple(phi.shape) != tuple(rho.shape) or phi.dtype != torch.float32:
            raise ValueError("phi must be a float32 mesh shaped like rho")
        return phi.contiguous()
    if conf is None:
        raise ValueError("a correction with the potential channel needs conf (or phi) to evaluate")
    from ..nbody.gravity import gravity_potential

    return gravity_potential(rho, om, conf)


def evaluate_residual(model: MeshCNNPotentialCorrection, rho, a, omega_m, keep=False, stats=None, keep_out=False,
                      conf=None, phi=None, output=None):
    """Density residual ``rho_add`` of the correction for the density mesh ``rho`` (X, Y, Z).

    Returns ``(rho_add, tape)``; ``tape`` is None unless ``keep`` (then it holds the
    activations for ``residual_backward``) or ``keep_out`` (then an ``OutputTape``).  ``stats``
    (the [sigma, gain, tanh, sigma_phi] device floats of a forward on the whole mesh) are used
    instead of this mesh's own: ``rho`` is then a slab of that mesh.  A model with the potential
    channel needs ``conf`` (the force step's configuration, for the potential of ``rho``) or
    ``phi`` (the potential mesh, or the slab of it that matches ``rho``).  With ``keep`` and
    ``output`` (this mesh's output-layer pre-activation, e.g. the slab of a whole-mesh
    ``OutputTape.out``) the output layer and the activation only it reads are not computed:
    the tape then serves the inputs' cotangent (``residual_backward(weights=False)``) and
    ``rho_add`` is None."""
    require_device(rho, "rho")
    if rho.dtype != torch.float32 or rho.ndim != 3:
        raise ValueError("rho must be a float32 (X, Y, Z) mesh")
    dev = rho.device
    ensure_init(dev)
    rho = rho.contiguous()
    shape = tuple(rho.shape)
    n = rho.numel()
    a, om = float(a), float(omega_m)
    phi = _potential(model, rho, om, conf, phi)
    ws = _workspace(dev)
    given = stats is not None
    if not given:
        stats = torch.empty(4, dtype=torch.float32, device=dev)
    s = stream_handle(dev)
    s8 = float(model.sigma8_value)
    layers = model.layers
    pres, acts = [], []
    gain = None if given else ptr(model.gain_params)
    fused = (not keep and model.block == "plain" and model.depth >= 2 and shape[0] % 4 == 0 and shape[2] % 16 == 0)
    skip_out = keep and output is not None  # the last activation feeds only the output layer
    if keep:
        # the backward's stem weight gradient needs the network input x0
        x0 = torch.empty(shape + (IN_CHANNELS,), dtype=torch.bfloat16, device=dev)
        _lib.call("pm_correction_features", ptr(rho), ptr(phi), n, om, a, s8, float(model.max_residual), gain,
                  ptr(x0), ptr(stats), ptr(ws), ws.numel(), s)
        pre, x = conv3d(x0, model.weight(layers[0]), model.bias(layers[0]), gelu=True, want_pre=True,
                        want_act=not (skip_out and model.depth == 1))
    else:
        # inference: the stem reads the density directly (constant input channels folded
        # into its bias), so x0 is never written
        x0, pre = None, None
        if not given:
            _lib.call("pm_correction_features", ptr(rho), ptr(phi), n, om, a, s8, float(model.max_residual), gain,
                      None, ptr(stats), ptr(ws), ws.numel(), s)
        x = torch.empty(shape + (CHANNELS,), dtype=torch.bfloat16, device=dev)
        if fused:
            # stem and first (plain) hidden layer in one kernel: the stem's activation stays in LDS
            _lib.call("pm_conv3d_stem_hidden", ptr(rho), ptr(phi), _lib.int3(shape), om, a, s8, ptr(stats),
                      ptr(model.weight(layers[0])), ptr(model.bias(layers[0])), ptr(model.weight(layers[1])),
                      ptr(model.bias(layers[1])), ptr(x), s)
        else:
            _lib.call("pm_conv3d_stem", ptr(rho), ptr(phi), _lib.int3(shape), om, a, s8, ptr(stats),
                      ptr(model.weight(layers[0])), ptr(model.bias(layers[0])), ptr(x), s)
    pres.append(pre)
    acts.append(x)
    for i in range(1 if fused else 0, model.depth - 1):
        last_act = not (skip_out and i == model.depth - 2)
        if model.block == "plain":
            l1 = layers[1 + i]
            p1, x = conv3d(x, model.weight(l1), model.bias(l1), gelu=True, want_pre=keep, want_act=last_act)
            pres.append(p1)
            acts.append(x)
            continue
        l1, l2 = layers[1 + 2 * i], layers[2 + 2 * i]
        p1, h = conv3d(x, model.weight(l1), model.bias(l1), gelu=True, want_pre=keep)
        p2, x = conv3d(h, model.weight(l2), model.bias(l2), res=x, gelu=True, want_pre=keep, want_act=last_act)
        pres += [p1, p2]
        acts += [h, x]
    if skip_out:
        out = output.contiguous()
        if out.dtype != torch.float32 or tuple(out.shape[:3]) != shape:
            raise ValueError("output must be a float32 mesh shaped like rho")
        return None, Tape(rho, stats, x0, pres, acts, out, a, om, phi, conf)
    rho_add = torch.empty_like(rho)
    if not keep and not keep_out:
        # inference: the output layer writes rho_add directly (its pre-activation is not kept)
        _lib.call("pm_conv3d_out_residual", ptr(x), ptr(model.weight(layers[-1])), ptr(model.bias(layers[-1])),
                  ptr(rho), _lib.int3(shape), om, float(model.max_residual), ptr(stats), ptr(rho_add), s)
        return rho_add, None
    out, _ = conv3d(x, model.weight(layers[-1]), model.bias(layers[-1]), gelu=False, want_pre=True, want_act=False,
                    pre_channels=1)
    _lib.call("pm_correction_residual", ptr(out), ptr(rho), n, om, float(model.max_residual), ptr(stats),
              ptr(rho_add), s)
    if not keep:
        return rho_add, OutputTape(stats, out, phi, conf)
    return rho_add, Tape(rho, stats, x0, pres, acts, out, a, om, phi, conf)


def _stem_weight_16(model):
    """The stem's bf16 weights [27][16][8] zero-padded to 16 input channels (for its data
    gradient, which the 16 -> 16 kernel computes)."""
    w = model.weight(model.layers[0]).view(TAPS, CHANNELS, IN_CHANNELS)
    out = torch.zeros((TAPS, CHANNELS, CHANNELS), dtype=torch.bfloat16, device=w.device)
    out[:, :, :IN_CHANNELS] = w
    return out


def residual_backward(model: MeshCNNPotentialCorrection, tape: Tape, rho_cot, grads=None, inputs=False,
                      weights=True, input_grad=False):
    """Gradient of ``<rho_cot, rho_add>`` with respect to the parameters (flat float32, the
    parameter layout); written into ``grads`` when given.  With ``inputs`` also the cotangents
    through the network's inputs: returns ``(grads, rho_bar, omega_bar)`` with ``rho_bar`` the
    density cotangent (mesh) and ``omega_bar`` a device float (``pm_correction_input_bwd``).
    With ``input_grad`` it returns only the gradient of the network input x0's first four
    channels (a float32 (X, Y, Z, 4) mesh, what the input path reads); ``weights=False`` skips
    the weight gradients (the output layer's gradient then stays one channel wide:
    ``pm_conv3d_bwd_data_out1``)."""
    dev = tape.rho.device
    n = tape.rho.numel()
    shape = tuple(tape.rho.shape)
    if grads is None:
        grads = torch.zeros_like(model.params) if weights else torch.zeros(model.params.numel() - model.gain_offset,
                                                                            dtype=torch.float32, device=dev)
    rc = rho_cot.contiguous()
    if rc.dtype != torch.float32 or tuple(rc.shape) != shape:
        raise ValueError("rho_cot must be a float32 mesh shaped like rho")
    ws = _workspace(dev)
    # the output layer's weight gradient reads its 16-channel dz; without it one channel suffices
    dz_ch = CHANNELS if weights else 1
    dz = torch.empty(shape + ((CHANNELS,) if weights else ()), dtype=torch.float32, device=dev)
    g = model.gain_offset if weights else 0
    _lib.call("pm_correction_residual_bwd", ptr(rc), ptr(tape.out), ptr(tape.rho), n, tape.omega_m, tape.a,
              float(model.sigma8_value), float(model.max_residual), ptr(tape.stats), ptr(dz), dz_ch,
              ptr(grads[g:g + 4]), ptr(ws), ws.numel(), stream_handle(dev))
    layers = model.layers

    def wgrad(layer, dzl, x):
        if weights:
            conv3d_bwd_weight(dzl, x, out=grads[layer.offset:layer.offset + layer.size])

    # output layer: input acts[-1] = gelu(pres[-1])
    wgrad(layers[-1], dz, tape.acts[-1])
    # without the weight gradients, the pre-activation gradients only feed further data
    # gradients, whose matrix cores read them as bf16: they are stored as bf16
    gdt = torch.float32 if weights else torch.bfloat16
    if weights:
        dz = conv3d_bwd_data(dz, model.weight(layers[-1]), gact=tape.pres[-1])
    else:
        dz = conv3d_bwd_data_out1(dz, model.weight(layers[-1]), gact=tape.pres[-1], dtype=gdt)
    for i in reversed(range(model.depth - 1)):
        if model.block == "plain":
            # dz is the gradient of block i's pre-activation; its input is acts[i] = gelu(pres[i])
            l1 = layers[1 + i]
            wgrad(l1, dz, tape.acts[i])
            dz = conv3d_bwd_data(dz, model.weight(l1), gact=tape.pres[i], dtype=gdt)
            continue
        l1, l2 = layers[1 + 2 * i], layers[2 + 2 * i]
        x_in, pre_in = tape.acts[2 * i], tape.pres[2 * i]
        h, pre1 = tape.acts[2 * i + 1], tape.pres[2 * i + 1]
        wgrad(l2, dz, h)
        dz1 = conv3d_bwd_data(dz, model.weight(l2), gact=pre1, dtype=gdt)
        wgrad(l1, dz1, x_in)
        dz = conv3d_bwd_data(dz1, model.weight(l1), add=dz, gact=pre_in, dtype=gdt)
    wgrad(layers[0], dz, tape.x0)
    if not inputs and not input_grad:
        return grads
    dx0 = conv3d_bwd_data(dz, _stem_weight_16(model), channels=4)
    if input_grad:
        return dx0
    rho_bar, omega_bar = _input_bwd(model, tape.rho, rc, tape.out, tape.stats, dx0, tape.omega_m, tape.phi,
                                    tape.conf)
    return grads, rho_bar, omega_bar


def _input_bwd(model, rho, rho_cot, out, stats, dx0c, omega_m, phi=None, conf=None):
    """``pm_correction_input_bwd``; with the potential channel, the potential cotangent's path to
    the density (1.5 Omega_m times the symmetric potential operator) is added to rho_bar."""
    dev = rho.device
    ws = _workspace(dev)
    rho_bar = torch.empty_like(rho)
    omega_bar = torch.zeros((), dtype=torch.float32, device=dev)
    phi_bar = None if phi is None else torch.empty_like(rho)
    _lib.call("pm_correction_input_bwd", ptr(rho_cot), ptr(out), ptr(rho), ptr(phi), ptr(dx0c), rho.numel(),
              float(omega_m), float(model.max_residual), ptr(stats), ptr(model.gain_params), ptr(rho_bar),
              ptr(phi_bar), ptr(omega_bar), ptr(ws), ws.numel(), stream_handle(dev))
    if phi is not None:
        from ..nbody.gravity import gravity_potential

        gravity_potential(phi_bar, omega_m, conf, out=rho_bar, shift=0.0, add=rho_bar)
    return rho_bar, omega_bar


# cells of the extended slab (slab + halo) that one network reverse pass may hold: its tape and
# gradients take about 350 bytes per cell (44 GiB at this size)
SLAB_CELLS = 1 << 27


def network_input_vjp(model: MeshCNNPotentialCorrection, rho, rho_cot, a, omega_m, tape: OutputTape, slab=None):
    """``(rho_bar, omega_bar)``: the cotangents of ``<rho_cot, rho_add(rho, Omega_m)>`` through
    the network's inputs (``pm_correction_input_bwd``), with the network's reverse pass run over
    x-slabs so that its activations and gradients never exist for the whole mesh at once.

    Each slab of ``slab`` x planes (a multiple of 8; default: the largest whose extended slab
    holds at most ``SLAB_CELLS`` cells, or the whole mesh when that fits) is extended by a halo
    of H planes on both sides (wrapped periodically), H = 2 x the number of conv layers rounded
    up to a multiple of 8; the network's forward up to the output layer (with the whole
    mesh's statistics from ``tape``; the output layer's pre-activation is the slab of the whole
    mesh's, ``tape.out``) and its data backward run on the extended slab as if it were periodic.
    Each conv reaches one plane further, so after the at most 2 x layers convolutions of the
    forward and the backward only the outer H planes hold wrapped values and the slab's own
    planes equal the whole-mesh computation.  The input gradient of those planes is kept (its first 4
    channels, all the input path reads); the global terms (through sigma, the gain and Omega_m)
    follow on the whole mesh."""
    require_device(rho, "rho")
    dev = rho.device
    shape = tuple(rho.shape)
    n0, plane = shape[0], shape[1] * shape[2]
    om = float(omega_m)
    rc = rho_cot.contiguous()
    halo = -(-2 * len(model.layers) // 8) * 8
    if slab is None:
        slab = n0 if n0 * plane <= SLAB_CELLS else max(8, (SLAB_CELLS // plane - 2 * halo) // 8 * 8)
    if slab % 8 or slab < 8:
        raise ValueError("slab must be a positive multiple of 8 planes")
    dx0c = torch.empty(shape + (4,), dtype=torch.float32, device=dev)
    phi = tape.phi
    if slab >= n0:
        _, full = evaluate_residual(model, rho, a, om, keep=True, stats=tape.stats, phi=phi, output=tape.out)
        dx0c.copy_(residual_backward(model, full, rc, weights=False, input_grad=True))
        del full
    else:
        for x0 in range(0, n0, slab):
            s = min(slab, n0 - x0)
            lo, hi = x0 - halo, x0 + s + halo
            if lo >= 0 and hi <= n0:  # the extended slab does not wrap: contiguous views, read only
                def ext_of(t, lo=lo, hi=hi):
                    return t[lo:hi]
            else:
                idx = torch.arange(lo, hi, device=dev).remainder_(n0)

                def ext_of(t, idx=idx):
                    return t.index_select(0, idx)
            rho_e = ext_of(rho)
            phi_e = None if phi is None else ext_of(phi)
            _, ext = evaluate_residual(model, rho_e, a, om, keep=True, stats=tape.stats, phi=phi_e,
                                       output=ext_of(tape.out))
            dx0_e = residual_backward(model, ext, ext_of(rc), weights=False, input_grad=True)
            del ext
            dx0c[x0:x0 + s] = dx0_e[halo:halo + s]
            del dx0_e, rho_e, phi_e
    return _input_bwd(model, rho, rc, tape.out, tape.stats, dx0c, om, phi, tape.conf)


TRAINED_DIR = __import__("pathlib").Path(__file__).resolve().parent / "weights"
TRAINED_DEFAULT = "mesh_cnn_2x_from_4x.npz"
TRAINED_DENSITY_ONLY = "mesh_cnn_2x_from_4x_density_only.npz"


def save_correction(model: MeshCNNPotentialCorrection, path, **meta) -> None:
    """Writes the float32 parameters and hyperparameters to an ``.npz`` file (plain arrays, no
    pickled objects); ``meta`` adds numeric or string fields (the training recipe)."""
    import numpy as np

    fields = {"params": model.params.detach().cpu().numpy().astype(np.float32),
              "depth": np.int64(model.depth), "max_residual": np.float64(model.max_residual),
              "output_init_scale": np.float64(model.output_init_scale),
              "sigma8_value": np.float64(model.sigma8_value), "block": np.str_(model.block),
              "symmetry": np.str_(model.symmetry), "potential": np.bool_(model.potential)}
    for k, v in meta.items():
        fields[f"meta_{k}"] = np.asarray(v)
    with open(path, "wb") as f:
        np.savez(f, **fields)


def load_correction(path=None, device="cuda") -> MeshCNNPotentialCorrection:
    """A correction saved by ``save_correction`` (default: the weights shipped with the package,
    trained on coarse (2x) / fine (4x particle-grid) force pairs and whole runs at 1.95 Mpc/h
    particle spacing, reading the potential channel; ``TRAINED_DENSITY_ONLY`` in the same folder
    is the network without it).  Loaded with ``numpy.load(allow_pickle=False)``."""
    import numpy as np

    path = TRAINED_DIR / TRAINED_DEFAULT if path is None else path
    with np.load(path, allow_pickle=False) as z:
        params = torch.from_numpy(np.ascontiguousarray(z["params"], dtype=np.float32)).to(device)
        return MeshCNNPotentialCorrection(params=params, depth=int(z["depth"]), max_residual=float(z["max_residual"]),
                                          output_init_scale=float(z["output_init_scale"]),
                                          sigma8_value=float(z["sigma8_value"]), block=str(z["block"]),
                                          symmetry=str(z["symmetry"]) if "symmetry" in z.files else "none",
                                          potential=bool(z["potential"]) if "potential" in z.files else False)


def correction_meta(path=None) -> dict:
    """The ``meta_*`` fields of a saved correction (its training recipe), as host values."""
    import numpy as np

    path = TRAINED_DIR / TRAINED_DEFAULT if path is None else path
    with np.load(path, allow_pickle=False) as z:
        return {k[5:]: z[k].item() if z[k].ndim == 0 else z[k].tolist() for k in z.files if k.startswith("meta_")}
