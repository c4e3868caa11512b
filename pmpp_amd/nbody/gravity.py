"""PM gravity on the HIP library (drop-in for ``pmpp.nbody.gravity``).

``gravity(a, ptcl, cosmo, conf, correction=None)`` follows ``src/pmpp/nbody/gravity.py:786-814``
on the single-device streamed branch (:597-606, :506-543): CIC density of ``N_m/N_p`` per valid
particle, source 1.5 Omega_m (rho - 1), rfftn, particle-Nyquist mask, -1/k^2, -i k_i, irfftn and
CIC gather per component.  The whole step is one call into ``pm_gravity``; its reverse is
``pm_gravity_vjp`` (the VJP walked by ``force_adj``, ``nbody/integrator.py:620-686``), exposed
through ``torch.autograd`` with cotangents for ``disp`` and ``Omega_m``.
"""

from __future__ import annotations

import weakref

import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, stream_handle, workspace


def gravity_desc(conf, omega_m) -> _lib.GravityDesc:
    d = _lib.GravityDesc()
    for a in range(3):
        d.mesh_shape[a] = int(conf.mesh_shape[a])
        d.brick[a] = int(conf.brick[a])
    d.cell_size = float(conf.cell_size)
    d.ptcl_spacing = float(conf.ptcl_spacing)
    d.mask_active = int(conf.mask_active)
    d.omega_m = float(torch.as_tensor(omega_m, dtype=torch.float32))  # cast as gravity.py:630
    d.particle_value = float(conf.mesh_size / conf.ptcl_num)
    return d


def _check_particles(pm, dp):
    n = pm.shape[0]
    if pm.ndim != 2 or pm.shape[1] != 3 or tuple(dp.shape) != (n, 3):
        raise ValueError(f"pmid {tuple(pm.shape)} and disp {tuple(dp.shape)} must both be (N, 3)")
    return n


# the uint8 validity mask of an unused_index tensor, kept while that tensor lives and is not
# modified in place (its _version), so that a run whose particles carry the same mask through
# every step converts it once instead of at every force
_VALID_MASKS = {}  # id(unused_index) -> (weak reference to it, its _version, the mask)


def _valid_u8(ptcl):
    ui = ptcl.unused_index
    if ui is None:
        return None
    if ui.shape != (ptcl.pmid.shape[0],):
        raise ValueError(f"unused_index has shape {tuple(ui.shape)}, expected ({ptcl.pmid.shape[0]},)")
    key = id(ui)
    hit = _VALID_MASKS.get(key)
    if hit is not None and hit[0]() is ui and hit[1] == ui._version:
        return hit[2]
    valid = (~ui.detach()).to(torch.uint8).contiguous()
    _VALID_MASKS[key] = (weakref.ref(ui, lambda _r, k=key: _VALID_MASKS.pop(k, None)), ui._version, valid)
    return valid


class _GravityFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, omega_m, pmid, valid, conf):
        dev = disp.device
        ensure_init(dev)
        d = gravity_desc(conf, omega_m.detach())
        pm = as_contig(pmid, torch.int16, "pmid")
        dp = as_contig(disp.detach(), torch.float32, "disp")
        n = _check_particles(pm, dp)
        nbytes = _lib.size_query("pm_gravity_workspace", d, n)
        ws = workspace(nbytes, dev)
        acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
        _lib.call("pm_gravity", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(acc), ptr(ws), nbytes, stream_handle(dev))
        ctx.save_for_backward(dp, omega_m, pm, valid)
        ctx.conf = conf
        return acc

    @staticmethod
    def backward(ctx, acc_cot):
        dp, omega_m, pm, valid = ctx.saved_tensors
        conf = ctx.conf
        dev = dp.device
        d = gravity_desc(conf, omega_m.detach())
        n = _check_particles(pm, dp)
        nbytes = _lib.size_query("pm_gravity_vjp_workspace", d, n)
        ws = workspace(nbytes, dev)
        ac = as_contig(acc_cot, torch.float32, "acc_cot")
        disp_cot = torch.empty((n, 3), dtype=torch.float32, device=dev)
        omega_cot = torch.zeros((), dtype=torch.float32, device=dev)
        _lib.call("pm_gravity_vjp", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(ac), ptr(disp_cot), ptr(omega_cot),
                  ptr(ws), nbytes, stream_handle(dev))
        return disp_cot, omega_cot.to(omega_m.dtype).reshape(omega_m.shape), None, None, None


class _CorrectedGravityFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, omega_m, ptcl, conf, correction, a):
        from ..corrections.training import corrected_gravity

        p = ptcl.replace(disp=disp.detach())
        ctx.state = (p, conf, correction, a, omega_m)
        return corrected_gravity(a, p, float(omega_m.detach()), conf, correction)

    @staticmethod
    def backward(ctx, acc_cot):
        from ..corrections.training import corrected_gravity_vjp

        p, conf, correction, a, omega_m = ctx.state
        d_disp, d_om = corrected_gravity_vjp(a, p, float(omega_m.detach()), conf, correction, acc_cot)
        om = torch.as_tensor(d_om, dtype=omega_m.dtype, device=omega_m.device).reshape(omega_m.shape)
        return d_disp, om, None, None, None, None


def gravity(a, ptcl, cosmo, conf, correction=None):
    """Gravitational accelerations of particles in [H_0^2] (gravity.py:786-814).

    ``cosmo`` needs an ``Omega_m`` (float or tensor; a tensor with ``requires_grad`` gets a
    cotangent).  ``a`` is unused by the uncorrected force, as in the reference.  A
    ``MeshCNNPotentialCorrection`` adds its density residual before the Poisson solve
    (gravity.py:635-640 with corrections/core.py:286-334); its reverse runs through the network
    and its density input (``corrected_gravity_vjp``).
    """
    if correction is not None:
        omega = cosmo.Omega_m if hasattr(cosmo, "Omega_m") else cosmo
        if not isinstance(omega, torch.Tensor):
            omega = torch.tensor(float(omega), dtype=torch.float32, device=ptcl.disp.device)
        return _CorrectedGravityFn.apply(ptcl.disp, omega, ptcl, conf, correction, float(a))
    del a
    omega = cosmo.Omega_m if hasattr(cosmo, "Omega_m") else cosmo
    if not isinstance(omega, torch.Tensor) and not (torch.is_grad_enabled() and ptcl.disp.requires_grad):
        # nothing to differentiate: a plain call, with Omega_m passed by value (a device tensor
        # made from it would be read back to the host, a synchronisation per force)
        return _gravity_plain(ptcl, float(omega), conf)
    if not isinstance(omega, torch.Tensor):
        omega = torch.tensor(float(omega), dtype=torch.float32, device=ptcl.disp.device)
    return _GravityFn.apply(ptcl.disp, omega, ptcl.pmid.detach(), _valid_u8(ptcl), conf)


def _gravity_plain(ptcl, omega, conf):
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, omega)
    pm = as_contig(ptcl.pmid.detach(), torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    valid = _valid_u8(ptcl)
    _lib.call("pm_gravity", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(acc), ptr(ws), nbytes, stream_handle(dev))
    return acc


class KeptForce:
    """A workspace holding the forward meshes of one force evaluation (``pm_gravity_keep``)
    for the VJP at the same particles (``pm_gravity_vjp_resume``): the reverse sweep recomputes
    the force at the reconstructed positions and pulls the next cotangent back there
    (integrator.py:781-837).  ``key`` names the particle arrays (and their versions) the
    meshes belong to; a different key means the workspace holds nothing usable."""

    def __init__(self):
        self.ws = None
        self.nbytes = 0
        self.key = None

    @staticmethod
    def key_of(ptcl, omega):
        return (ptcl.pmid.data_ptr(), ptcl.pmid._version, ptcl.disp.data_ptr(), ptcl.disp._version,
                None if ptcl.unused_index is None else (ptcl.unused_index.data_ptr(), ptcl.unused_index._version),
                len(ptcl), float(omega))


def _vjp_call(ptcl, cosmo, conf, acc_cot, kept=None):
    omega = cosmo.Omega_m if hasattr(cosmo, "Omega_m") else cosmo
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, float(omega))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_vjp_workspace", d, n)
    resume = (kept is not None and kept.ws is not None and kept.nbytes == nbytes
              and kept.key == KeptForce.key_of(ptcl, omega))
    if kept is not None:  # the sweep's VJPs all run in its one kept workspace
        kept.key = None  # the VJP's cotangent meshes overwrite part of the workspace
        if kept.ws is None or kept.ws.numel() < nbytes or kept.ws.device != dev:
            kept.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        kept.nbytes = nbytes
        ws = kept.ws
    else:
        ws = workspace(nbytes, dev)
    ac = as_contig(acc_cot.detach(), torch.float32, "acc_cot")
    disp_cot = torch.empty((n, 3), dtype=torch.float32, device=dev)
    omega_cot = torch.zeros((), dtype=torch.float32, device=dev)
    valid = _valid_u8(ptcl)
    _lib.call("pm_gravity_vjp_resume" if resume else "pm_gravity_vjp", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(ac),
              ptr(disp_cot), ptr(omega_cot), ptr(ws), nbytes, stream_handle(dev))
    return disp_cot, omega_cot


def gravity_vjp(ptcl, cosmo, conf, acc_cot, kept=None, device_omega=False):
    """(disp cotangent (N,3) tensor, Omega_m cotangent) of ``gravity`` for ``acc_cot``.

    The reverse of the force step without autograd bookkeeping (``pm_gravity_vjp``), as the
    N-body reverse sweep calls it once per drift stage (integrator.py:620-686).  With a
    ``KeptForce`` that ``gravity_keep`` filled for these same particles, the forward meshes
    are taken from it (``pm_gravity_vjp_resume``) instead of being recomputed.  The Omega_m
    cotangent is a float, or with ``device_omega`` a 0-d device tensor (no host round trip).
    """
    disp_cot, omega_cot = _vjp_call(ptcl, cosmo, conf, acc_cot, kept)
    return disp_cot, (omega_cot if device_omega else float(omega_cot.item()))


def gravity_keep(ptcl, cosmo, conf, kept):
    """``gravity`` (accelerations (N,3), no autograd) that leaves its forward meshes in
    ``kept`` for a following ``gravity_vjp`` at the same particles (``pm_gravity_keep``)."""
    omega = cosmo.Omega_m if hasattr(cosmo, "Omega_m") else cosmo
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, float(omega))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_vjp_workspace", d, n)
    kept.key = None  # the call overwrites the workspace: it names nothing until it has succeeded
    if kept.ws is None or kept.ws.numel() < nbytes or kept.ws.device != dev:
        kept.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    kept.nbytes = nbytes
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    valid = _valid_u8(ptcl)
    _lib.call("pm_gravity_keep", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(acc), ptr(kept.ws), nbytes,
              stream_handle(dev))
    kept.key = KeptForce.key_of(ptcl, omega)
    return acc


def gravity_potential(dens, omega_m, conf, out=None, shift=-1.0, scale=None, add=None):
    """Real-space potential of a density mesh on the force step's spectral band:
    ``irfftn(mask * (-1/k^2) * rfftn(1.5 Omega_m (dens - 1)))``, the potential the reference's
    corrected force builds (``nbody/gravity.py:419-430``) and reads back in real space for the
    mesh CNN's potential channel (``corrections/core.py:286-334``, ``mesh_cnn.py:76-84``).

    ``shift`` and ``scale`` (default -1 and 1.5 Omega_m) set the source ``scale * (dens + shift)``;
    the operator is symmetric, so ``shift=0`` applies it to a cotangent.  ``add`` (a mesh, may be
    ``out``) is added to the result.  One ``pm_gravity_potential`` call."""
    dev = dens.device
    ensure_init(dev)
    d = gravity_desc(conf, omega_m)
    x = as_contig(dens, torch.float32, "dens")
    shape = tuple(int(v) for v in conf.mesh_shape)
    if tuple(x.shape) != shape:
        raise ValueError(f"mesh {tuple(x.shape)} does not match the configuration's {shape}")
    if out is None:
        out = torch.empty_like(x)
    if tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 mesh of the configuration's shape")
    if add is not None and (tuple(add.shape) != shape or add.dtype != torch.float32 or not add.is_contiguous()):
        raise ValueError("add must be a contiguous float32 mesh of the configuration's shape")
    sc = 1.5 * d.omega_m if scale is None else float(scale)
    nbytes = _lib.size_query("pm_potential_workspace", d)
    ws = workspace(nbytes, dev)
    _lib.call("pm_gravity_potential", ptr(x), float(shift), sc, ptr(add), d, ptr(out), ptr(ws), nbytes,
              stream_handle(dev))
    return out
