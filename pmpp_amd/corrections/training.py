"""Training of the mesh-CNN force correction on coarse/fine force pairs.

The target of a particle snapshot is its PM force on a mesh ``fine`` times finer per axis (the
same particles, ``Configuration(mesh_shape=fine)``); the trained correction makes the
coarse-mesh force match it.  One training step is entirely on the GPU:

  rho      = pm_gravity_density               (the coarse step's own CIC density)
  rho_add  = mesh CNN(rho)                    (evaluate_residual, BF16 MFMA convolutions)
  acc      = pm_gravity_corrected(rho_add)
  loss     = |acc - target|^2 / |target|^2    (pm_force_loss, with d loss / d acc)
  rho_cot  = pm_gravity_corrected_vjp         (the force step's adjoint at fixed particles)
  grads    = residual_backward(rho_cot)
  params  <- pm_adamax_step                   (clip_by_global_norm + adamax, skipped when
                                               the gradient is not finite)

The optimizer mirrors ``build_correction_optimizer`` (``corrections/common.py:136-174``).
"""

from __future__ import annotations

import dataclasses

import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, stream_handle, workspace
from ..nbody.gravity import _check_particles, _valid_u8, gravity_desc
from .mesh_cnn import MeshCNNPotentialCorrection, evaluate_residual, residual_backward


@dataclasses.dataclass
class AdamaxOptimizer:
    learning_rate: float
    gradient_clip_norm: float = 100.0
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1e-8

    def init(self, params: torch.Tensor) -> "OptState":
        dev = params.device
        return OptState(torch.zeros_like(params), torch.zeros_like(params),
                        torch.zeros(1, dtype=torch.int32, device=dev))

    def step(self, model: MeshCNNPotentialCorrection, grads: torch.Tensor, state: "OptState") -> None:
        dev = model.params.device
        if grads.dtype != torch.float32 or grads.numel() != model.params.numel() or not grads.is_contiguous():
            raise ValueError("grads must be a contiguous float32 vector shaped like the parameters")
        for t in (state.mu, state.nu):
            if t.numel() != model.params.numel():
                raise ValueError("optimizer state does not match the parameters")
        ws = workspace(_lib.size_query("pm_correction_workspace"), dev)
        clip = float(self.gradient_clip_norm) if self.gradient_clip_norm else 0.0
        _lib.call("pm_adamax_step", ptr(model.params), ptr(grads), ptr(state.mu), ptr(state.nu), ptr(state.count),
                  model.params.numel(), float(self.learning_rate), self.b1, self.b2, self.eps, clip,
                  ptr(model.params_bf16), ptr(ws), ws.numel(), stream_handle(dev))


@dataclasses.dataclass
class OptState:
    mu: torch.Tensor
    nu: torch.Tensor
    count: torch.Tensor  # int32 on the device: applied steps


def build_correction_optimizer(learning_rate, gradient_clip_norm=100.0, optimizer_name="adamax",
                               apply_if_finite_steps=100):
    """Global-norm clipping then Adamax; non-finite steps are skipped on the device.

    ``apply_if_finite_steps`` is accepted for the reference's signature: the device-side
    skip has no limit on consecutive non-finite steps."""
    del apply_if_finite_steps
    if optimizer_name != "adamax":
        raise NotImplementedError(f"optimizer {optimizer_name!r}: only 'adamax' runs on the HIP library")
    return AdamaxOptimizer(float(learning_rate), gradient_clip_norm)


def _omega(cosmo):
    om = cosmo.Omega_m if hasattr(cosmo, "Omega_m") else cosmo
    return float(om)


def density(ptcl, cosmo, conf):
    """The force step's CIC density mesh of ``ptcl`` (``pm_gravity_density``)."""
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, _omega(cosmo))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    rho = torch.empty(tuple(int(s) for s in conf.mesh_shape), dtype=torch.float32, device=dev)
    _lib.call("pm_gravity_density", ptr(pm), ptr(dp), ptr(_valid_u8(ptcl)), n, d, ptr(rho), ptr(ws), nbytes,
              stream_handle(dev))
    return rho


def _residual_mesh(rho_add, conf):
    if rho_add is None:
        return None
    shape = tuple(int(s) for s in conf.mesh_shape)
    if rho_add.dtype != torch.float32 or tuple(rho_add.shape) != shape:
        raise ValueError(f"rho_add must be a float32 mesh of shape {shape}")
    return as_contig(rho_add, torch.float32, "rho_add")


def gravity_with_residual(ptcl, cosmo, conf, rho_add):
    """Accelerations with the density residual ``rho_add`` (``pm_gravity_corrected``)."""
    dev = ptcl.disp.device
    d = gravity_desc(conf, _omega(cosmo))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    ra = _residual_mesh(rho_add, conf)
    _lib.call("pm_gravity_corrected", ptr(pm), ptr(dp), ptr(_valid_u8(ptcl)), n, d, ptr(ra), ptr(acc), ptr(ws),
              nbytes, stream_handle(dev))
    return acc


def gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot):
    """(disp_cot, omega_cot (device float), rho_cot) of ``gravity_with_residual``."""
    dev = ptcl.disp.device
    d = gravity_desc(conf, _omega(cosmo))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    nbytes = _lib.size_query("pm_gravity_vjp_workspace", d, n)
    ws = workspace(nbytes, dev)
    ac = as_contig(acc_cot.detach(), torch.float32, "acc_cot")
    disp_cot = torch.empty((n, 3), dtype=torch.float32, device=dev)
    omega_cot = torch.zeros((), dtype=torch.float32, device=dev)
    rho_cot = torch.empty(tuple(int(s) for s in conf.mesh_shape), dtype=torch.float32, device=dev)
    ra = _residual_mesh(rho_add, conf)
    _lib.call("pm_gravity_corrected_vjp", ptr(pm), ptr(dp), ptr(_valid_u8(ptcl)), n, d, ptr(ra), ptr(ac),
              ptr(disp_cot), ptr(omega_cot), ptr(rho_cot), ptr(ws), nbytes, stream_handle(dev))
    return disp_cot, omega_cot, rho_cot


def corrected_gravity(a, ptcl, cosmo, conf, correction: MeshCNNPotentialCorrection):
    """Coarse-mesh accelerations with the learned correction (no gradient).

    One force step: the density (``pm_gravity_density``), the network on it, then the same
    step resumed with the residual (``pm_gravity_corrected_resume``), which reuses the
    bucketed particles and the density the first call left in the workspace (held here
    across the network's own allocations)."""
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, _omega(cosmo))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    valid = _valid_u8(ptcl)
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    s = stream_handle(dev)
    rho = torch.empty(tuple(int(v) for v in conf.mesh_shape), dtype=torch.float32, device=dev)
    _lib.call("pm_gravity_density", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(rho), ptr(ws), nbytes, s)
    rho_add, _ = evaluate_residual(correction, rho, a, _omega(cosmo))
    del rho
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    _lib.call("pm_gravity_corrected_resume", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(rho_add), ptr(acc), ptr(ws),
              nbytes, s)
    return acc


def corrected_gravity_vjp(a, ptcl, cosmo, conf, correction: MeshCNNPotentialCorrection, acc_cot):
    """(disp_cot, omega_cot float) of ``corrected_gravity``: the force step's adjoint at the
    corrected source (``pm_gravity_corrected_vjp``) plus the path through the network's inputs
    (``pm_correction_input_bwd``), whose density cotangent reaches the particles through the
    CIC scatter adjoint."""
    from ..cic.kernels import hip_scatter_bwd

    om = _omega(cosmo)
    rho = density(ptcl, cosmo, conf)
    rho_add, tape = evaluate_residual(correction, rho, a, om, keep=True)
    disp_cot, omega_cot, rho_cot = gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot)
    _, rho_bar, omega_bar = residual_backward(correction, tape, rho_cot, inputs=True)
    valid = _valid_u8(ptcl)
    dc, _ = hip_scatter_bwd(ptcl.pmid, ptcl.disp.detach(), conf.mesh_size / conf.ptcl_num, rho_bar, offset=0.0,
                            particle_cell_size=conf.cell_size, cell_size=None, global_shape=tuple(conf.mesh_shape),
                            valid_mask=None if valid is None else valid.bool())
    return disp_cot + dc, float(omega_cot.item()) + float(omega_bar.item())


def force_loss(acc, target, want_grad=True):
    """(loss, acc_cot): loss a device float64 tensor [|acc - t|^2 / |t|^2, 1 / |t|^2]."""
    dev = acc.device
    ws = workspace(_lib.size_query("pm_correction_workspace"), dev)
    a = as_contig(acc.detach(), torch.float32, "acc")
    t = as_contig(target, torch.float32, "target")
    if a.ndim != 2 or a.shape[1] != 3 or t.shape != a.shape:
        raise ValueError("acc and target must both be (N, 3)")
    loss = torch.empty(2, dtype=torch.float64, device=dev)
    cot = torch.empty_like(a) if want_grad else None
    _lib.call("pm_force_loss", ptr(a), ptr(t), a.shape[0], ptr(loss), ptr(cot), ptr(ws), ws.numel(),
              stream_handle(dev))
    return loss, cot


def train_step(model, optimizer, state, ptcl, a, cosmo, conf, target, grads=None):
    """One optimizer step on one snapshot; returns the loss before the step (device float64)."""
    rho = density(ptcl, cosmo, conf)
    rho_add, tape = evaluate_residual(model, rho, a, _omega(cosmo), keep=True)
    acc = gravity_with_residual(ptcl, cosmo, conf, rho_add)
    loss, acc_cot = force_loss(acc, target)
    _, _, rho_cot = gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot)
    grads = residual_backward(model, tape, rho_cot, grads)
    optimizer.step(model, grads, state)
    return loss[0]


def fine_run_pairs(cosmo, conf, conf_fine, seeds, every=1, device="cuda"):
    """Training pairs generated by the method: snapshots of fine-mesh N-body runs (one per
    seed, 2LPT initial conditions at ``conf_fine.a_start``), each re-anchored on the coarse
    mesh, with the fine-mesh force at that state as the target.  Returns
    ``[(a, coarse_particles, target_acc)]``; ``every`` keeps every n-th step."""
    from ..initial_conditions import linear_modes, lpt, white_noise
    from ..nbody.particles import Particles
    from ..nbody.solver import nbody

    pairs = []
    for seed in seeds:
        modes = linear_modes(white_noise(seed, conf_fine, device=device), cosmo, conf_fine)
        state = lpt(modes, cosmo, conf_fine)
        steps = []

        def observe(a, p):
            steps.append((a, p))

        nbody(state, cosmo, conf_fine, observer=observe)
        for i, (a, p) in enumerate(steps):
            if i % every:
                continue
            coarse = Particles.from_pos(conf, p.pos(dtype=torch.float64), vel=p.vel)
            pairs.append((a, coarse, p.acc.clone()))
    return pairs


def compare_runs(model, cosmo, conf, conf_fine, seed, device="cuda"):
    """The judge of the correction: from one seed's initial conditions, the final density of
    the plain coarse run, the corrected coarse run and the fine run, each measured on the fine
    mesh.  Returns a dict with ``k``, ``pk_plain``, ``pk_corrected``, ``pk_fine`` and the
    cross-correlations ``r_plain``, ``r_corrected`` against the fine run (torch, device)."""
    from ..analysis import density_to_cross_correlation, density_to_pk
    from ..cic import scatter
    from ..initial_conditions import linear_modes, lpt, white_noise
    from ..nbody.particles import Particles
    from ..nbody.solver import nbody

    modes = linear_modes(white_noise(seed, conf, device=device), cosmo, conf)
    plain = nbody(lpt(modes, cosmo, conf), cosmo, conf, observer=lambda a, p: None)
    corr = nbody(lpt(modes, cosmo, conf), cosmo, conf, correction=model)
    fine = nbody(lpt(modes, cosmo, conf_fine), cosmo, conf_fine, observer=lambda a, p: None)

    def dens(p):
        q = Particles.from_pos(conf_fine, p.pos(dtype=torch.float64))
        return scatter(q, conf_fine)

    d_plain, d_corr, d_fine = dens(plain), dens(corr), dens(fine)
    k, pk_fine, _ = density_to_pk(d_fine, conf_fine)
    _, r_plain, _, pk_plain, _, _ = density_to_cross_correlation(d_plain, d_fine, conf_fine)
    _, r_corr, _, pk_corr, _, _ = density_to_cross_correlation(d_corr, d_fine, conf_fine)
    return {"k": k, "pk_plain": pk_plain, "pk_corrected": pk_corr, "pk_fine": pk_fine, "r_plain": r_plain,
            "r_corrected": r_corr}
