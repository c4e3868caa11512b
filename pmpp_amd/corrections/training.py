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

import math

import dataclasses

import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, stream_handle, workspace
from ..nbody.gravity import _check_particles, _valid_u8, gravity_desc
from .mesh_cnn import (MeshCNNPotentialCorrection, evaluate_residual, network_input_vjp, project_cubic,
                       residual_backward)


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
    ensure_init(dev)
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
    ensure_init(dev)
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


def density_cot_add(ptcl, cosmo, conf, dens_cot, disp_cot):
    """disp_cot += the CIC scatter adjoint's displacement term for the density cotangent
    ``dens_cot`` (``pm_gravity_density_cot``: the particles binned by brick, the tiled column
    pass); invalid particles' rows are left as they are.  Returns ``disp_cot``."""
    dev = ptcl.disp.device
    ensure_init(dev)
    d = gravity_desc(conf, _omega(cosmo))
    pm = as_contig(ptcl.pmid, torch.int16, "pmid")
    dp = as_contig(ptcl.disp.detach(), torch.float32, "disp")
    n = _check_particles(pm, dp)
    dc = as_contig(dens_cot, torch.float32, "dens_cot")
    if tuple(dc.shape) != tuple(int(v) for v in conf.mesh_shape):
        raise ValueError("dens_cot must be a mesh of the configuration's shape")
    if disp_cot.dtype != torch.float32 or tuple(disp_cot.shape) != (n, 3) or not disp_cot.is_contiguous():
        raise ValueError("disp_cot must be a contiguous float32 (N, 3) tensor")
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    _lib.call("pm_gravity_density_cot", ptr(pm), ptr(dp), ptr(_valid_u8(ptcl)), n, d, ptr(dc), ptr(disp_cot),
              ptr(ws), nbytes, stream_handle(dev))
    return disp_cot


class KeptCorrection:
    """The density, the network's residual and its output tape of one corrected force
    evaluation (``corrected_gravity(..., kept=)``), for ``corrected_gravity_vjp`` at the same
    particles, scale factor and Omega_m: the reverse sweep recomputes the force at the
    reconstructed positions and pulls the next cotangent back there (integrator.py:781-837),
    so the VJP need not run the density and the network's forward again.  ``key`` names the
    particle arrays, a and Omega_m; None means nothing usable is held."""

    def __init__(self):
        self.key = None
        self.rho = self.rho_add = self.tape = None

    @staticmethod
    def key_of(ptcl, a, omega):
        from ..nbody.gravity import KeptForce

        return KeptForce.key_of(ptcl, omega) + (float(a),)

    def clear(self):
        self.key = None
        self.rho = self.rho_add = self.tape = None


def corrected_gravity(a, ptcl, cosmo, conf, correction: MeshCNNPotentialCorrection, kept: KeptCorrection = None):
    """Coarse-mesh accelerations with the learned correction (no gradient).

    One force step: the density (``pm_gravity_density``), the network on it, then the same
    step resumed with the residual (``pm_gravity_corrected_resume``), which reuses the
    bucketed particles and the density the first call left in the workspace (held here
    across the network's own allocations).  With ``kept`` the density, the residual and the
    network's output tape are kept there for a following ``corrected_gravity_vjp``."""
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
    shape = tuple(int(v) for v in conf.mesh_shape)
    if kept is None:  # read in place: the resume overwrites it only after the network is done with it
        off = _lib.size_query("pm_gravity_density_offset", d, n)
        rho = ws[off:off + 4 * math.prod(shape)].view(torch.float32).view(shape)
    else:  # kept for the VJP after the resume: a buffer of its own
        rho = torch.empty(shape, dtype=torch.float32, device=dev)
    _lib.call("pm_gravity_density", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(rho), ptr(ws), nbytes, s)
    if kept is None:
        rho_add, _ = evaluate_residual(correction, rho, a, _omega(cosmo), conf=conf)
        del rho
    else:
        kept.clear()
        rho_add, tape = evaluate_residual(correction, rho, a, _omega(cosmo), keep_out=True, conf=conf)
        kept.rho, kept.rho_add, kept.tape = rho, rho_add, tape
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    _lib.call("pm_gravity_corrected_resume", ptr(pm), ptr(dp), ptr(valid), n, d, ptr(rho_add), ptr(acc), ptr(ws),
              nbytes, s)
    if kept is not None:
        kept.key = KeptCorrection.key_of(ptcl, a, _omega(cosmo))
    return acc


def corrected_gravity_vjp(a, ptcl, cosmo, conf, correction: MeshCNNPotentialCorrection, acc_cot, slab=None,
                          param_grads=None, kept: KeptCorrection = None):
    """(disp_cot, omega_cot float) of ``corrected_gravity``: the force step's adjoint at the
    corrected source (``pm_gravity_corrected_vjp``) plus the path through the network's inputs
    (``pm_correction_input_bwd``), whose density cotangent reaches the particles through the
    CIC scatter adjoint (``density_cot_add``).  The network's reverse runs over x-slabs
    (``network_input_vjp``); with ``param_grads`` it runs on the whole mesh and also adds the
    parameters' gradient.  A
    ``kept`` evaluation of these same particles (``corrected_gravity(..., kept=)``) supplies the
    density, the residual and the network's output tape instead of recomputing them."""
    om = _omega(cosmo)
    if param_grads is not None:
        rho = density(ptcl, cosmo, conf)
        rho_add, full = evaluate_residual(correction, rho, a, om, keep=True, conf=conf)
        disp_cot, omega_cot, rho_cot = gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot)
        del rho_add
        g, rho_bar, omega_bar = residual_backward(correction, full, rho_cot, inputs=True)
        del full
        param_grads += g
    else:
        if kept is not None and kept.key is not None and kept.key == KeptCorrection.key_of(ptcl, a, om):
            rho, rho_add, tape = kept.rho, kept.rho_add, kept.tape
            kept.clear()
        else:
            rho = density(ptcl, cosmo, conf)
            rho_add, tape = evaluate_residual(correction, rho, a, om, keep_out=True, conf=conf)
        disp_cot, omega_cot, rho_cot = gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot)
        del rho_add
        rho_bar, omega_bar = network_input_vjp(correction, rho, rho_cot, a, om, tape, slab=slab)
    density_cot_add(ptcl, cosmo, conf, rho_bar, disp_cot)
    return disp_cot, float(omega_cot.item()) + float(omega_bar.item())


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
    """One optimizer step on one snapshot; returns the loss before the step (device float64).
    A model with ``symmetry="cubic"`` has its gradient projected onto the cube-symmetric
    kernels (``project_cubic``) before the step, so its weights stay in that subspace."""
    rho = density(ptcl, cosmo, conf)
    rho_add, tape = evaluate_residual(model, rho, a, _omega(cosmo), keep=True, conf=conf)
    acc = gravity_with_residual(ptcl, cosmo, conf, rho_add)
    loss, acc_cot = force_loss(acc, target)
    _, _, rho_cot = gravity_with_residual_vjp(ptcl, cosmo, conf, rho_add, acc_cot)
    grads = residual_backward(model, tape, rho_cot, grads)
    if model.symmetry == "cubic":
        project_cubic(grads, model.layers)
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


def compare_runs(model, cosmo, conf, conf_fine, seed, device="cuda", measure_conf=None):
    """The judge of the correction: from one seed's initial conditions, the final density of
    the plain coarse run, the corrected coarse run and the fine run, each measured on the mesh
    of ``measure_conf`` (default: the fine mesh).  Returns a dict with ``k``, ``pk_plain``,
    ``pk_corrected``, ``pk_fine`` and the cross-correlations ``r_plain``, ``r_corrected``
    against the fine run (torch, device)."""
    from ..analysis import density_to_cross_correlation, density_to_pk
    from ..cic import scatter
    from ..initial_conditions import linear_modes, lpt, white_noise
    from ..nbody.particles import Particles
    from ..nbody.solver import nbody

    modes = linear_modes(white_noise(seed, conf, device=device), cosmo, conf)
    plain = nbody(lpt(modes, cosmo, conf), cosmo, conf, observer=lambda a, p: None)
    corr = nbody(lpt(modes, cosmo, conf), cosmo, conf, correction=model)
    fine = nbody(lpt(modes, cosmo, conf_fine), cosmo, conf_fine, observer=lambda a, p: None)

    mc = conf_fine if measure_conf is None else measure_conf

    def dens(p):
        q = Particles.from_pos(mc, p.pos(dtype=torch.float64))
        return scatter(q, mc)

    d_plain, d_corr, d_fine = dens(plain), dens(corr), dens(fine)
    del plain, corr, fine
    k, pk_fine, _ = density_to_pk(d_fine, mc)
    _, r_plain, _, pk_plain, _, _ = density_to_cross_correlation(d_plain, d_fine, mc)
    _, r_corr, _, pk_corr, _, _ = density_to_cross_correlation(d_corr, d_fine, mc)
    return {"k": k, "pk_plain": pk_plain, "pk_corrected": pk_corr, "pk_fine": pk_fine, "r_plain": r_plain,
            "r_corrected": r_corr}


def mean_force_loss(model, pairs, cosmo, conf):
    """Mean relative squared force error |acc - t|^2 / |t|^2 over ``pairs`` of the coarse force
    with the correction ``model`` (None: the plain force); a host float."""
    from ..nbody.gravity import gravity

    tot = None
    for a, p, tgt in pairs:
        acc = gravity(a, p, cosmo.Omega_m, conf, correction=model)
        loss = force_loss(acc, tgt, want_grad=False)[0][0]
        tot = loss if tot is None else tot + loss
    return float(tot) / len(pairs)


def train_correction(pairs, cosmo, conf, *, iters, learning_rate, max_residual, depth=2, block="plain", seed=0,
                     gradient_clip_norm=100.0, symmetry="none", device="cuda", progress=None, progress_every=500,
                     potential=False):
    """Trains a freshly initialised correction on ``pairs`` ([(a, coarse particles, target acc)],
    ``fine_run_pairs``): ``iters`` Adamax steps (``build_correction_optimizer``), one pair per
    step drawn by numpy's PCG64(``seed``).  ``progress(it, mean loss of the last steps)`` is
    called every ``progress_every`` steps.  Returns ``(model, losses)``, the per-step training
    losses as a host numpy array (read once at the end, so the loop never waits on the GPU)."""
    import numpy as np

    from .mesh_cnn import init_mesh_cnn_potential_correction

    model = init_mesh_cnn_potential_correction(seed=seed, depth=depth, max_residual=max_residual, device=device,
                                               block=block, symmetry=symmetry, potential=potential)
    opt = build_correction_optimizer(learning_rate, gradient_clip_norm)
    state = opt.init(model.params)
    rng = np.random.default_rng(seed)
    grads = torch.zeros_like(model.params)
    losses = torch.empty(iters, dtype=torch.float64, device=model.params.device)
    for it in range(iters):
        a, p, tgt = pairs[int(rng.integers(len(pairs)))]
        losses[it] = train_step(model, opt, state, p, a, cosmo, conf, tgt, grads=grads)
        if progress is not None and (it + 1) % progress_every == 0:
            progress(it + 1, float(losses[it + 1 - progress_every:it + 1].mean()))
    return model, losses.cpu().numpy()


def pk_errors(res, kmax=1.0):
    """Summary of ``compare_runs``: over the shells 0 < k <= ``kmax`` (h/Mpc), the mean and the
    maximum of |P / P_fine - 1| and the mean cross-correlation with the fine run, for the plain
    and the corrected coarse runs (host floats)."""
    k = res["k"].cpu().numpy()
    sel = (k > 0) & (k <= kmax)
    fine = res["pk_fine"].cpu().numpy()[sel]
    out = {"kmax": kmax, "shells": int(sel.sum())}
    for name in ("plain", "corrected"):
        e = abs(res[f"pk_{name}"].cpu().numpy()[sel] / fine - 1)
        out[name] = {"pk_err_mean": float(e.mean()), "pk_err_max": float(e.max()),
                     "r_mean": float(res[f"r_{name}"].cpu().numpy()[sel].mean())}
    return out


class RunTarget:
    """One seed's end-to-end training target: the coarse run's initial state and the final
    density of the fine run from the same white noise, measured on the coarse mesh."""

    def __init__(self, seed, cosmo, conf, conf_fine, device="cuda"):
        from ..cic import scatter
        from ..initial_conditions import linear_modes, lpt, white_noise
        from ..nbody.particles import Particles
        from ..nbody.solver import nbody

        modes = linear_modes(white_noise(seed, conf, device=device), cosmo, conf)
        self.seed = seed
        self.p0 = lpt(modes, cosmo, conf)
        fine = nbody(lpt(modes, cosmo, conf_fine), cosmo, conf_fine, observer=lambda a, p: None)
        self.dens_fine = scatter(Particles.from_pos(conf, fine.pos(dtype=torch.float64)), conf)


def run_loss(model, target: RunTarget, cosmo, conf, kmax=1.0, r_weight=1.0, grads=None):
    """End-to-end loss of one corrected run and (with ``grads``) its parameter gradient: the
    corrected coarse run from ``target.p0`` to the last scale factor, its final CIC density on
    the coarse mesh, and over the shells 0 < k <= ``kmax``
        L = mean (log P(k) - log P_fine(k))^2 + r_weight mean (1 - r(k))
    with r the cross-correlation with the fine run.  The gradient runs through the P(k) and r
    adjoints, the scatter adjoint and the N-body reverse sweep with the network's parameter
    gradient (``nbody_adj(param_grads=...)``), added to ``grads``.  Returns the host float L."""
    import numpy as np

    from ..analysis import (density_to_cross_correlation, density_to_cross_correlation_vjp, density_to_pk,
                            density_to_pk_vjp)
    from ..cic import scatter
    from ..nbody.solver import nbody, nbody_adj

    with torch.no_grad():
        out = nbody(target.p0, cosmo, conf, correction=model, observer=lambda a, p: None)
    disp = out.disp.detach().clone().requires_grad_(True)
    dens = scatter(out.replace(disp=disp), conf)
    k, pk, _ = density_to_pk(dens.detach(), conf)
    _, pkf, _ = density_to_pk(target.dens_fine, conf)
    _, r, _, _, _, _ = density_to_cross_correlation(dens.detach(), target.dens_fine, conf)
    k, pk, pkf, r = (x.double().cpu().numpy() for x in (k, pk, pkf, r))
    sel = (k > 0) & (k <= kmax)
    ns = int(sel.sum())
    d = np.log(pk[sel]) - np.log(pkf[sel])
    loss = float((d ** 2).mean() + r_weight * (1.0 - r[sel]).mean())
    if grads is None:
        return loss
    pk_cot = np.zeros(len(pk))
    pk_cot[sel] = 2.0 * d / (ns * pk[sel])
    r_cot = np.zeros(len(pk))
    r_cot[sel] = -r_weight / ns
    dens_cot = density_to_pk_vjp(dens.detach(), conf, pk_cot)
    if r_weight:
        dens_cot += density_to_cross_correlation_vjp(dens.detach(), target.dens_fine, conf, r_cot)
    dens.backward(dens_cot)
    cot = out.replace(disp=disp.grad, vel=torch.zeros_like(out.vel), acc=torch.zeros_like(out.acc))
    nbody_adj(out, cot, cosmo, conf, correction=model, param_grads=grads)
    return loss


def train_correction_runs(model, targets, cosmo, conf, *, iters, learning_rate, kmax=1.0, r_weight=1.0, seed=0,
                          gradient_clip_norm=100.0, validate=None, validate_every=50, progress=None):
    """Fine-tunes ``model`` end to end on whole corrected runs (``run_loss`` over ``targets``,
    one drawn per step by numpy's PCG64(``seed``)) with Adamax.  Every ``validate_every`` steps
    ``validate(model)`` returns a score (lower is better); the parameters of the best-scoring
    checkpoint (the starting point included) are restored at the end.  Returns
    ``(losses, [(step, score), ...])``."""
    import numpy as np

    opt = build_correction_optimizer(learning_rate, gradient_clip_norm)
    state = opt.init(model.params)
    rng = np.random.default_rng(seed)
    grads = torch.zeros_like(model.params)
    losses, scores = [], []
    best = None
    if validate is not None:
        s0 = validate(model)
        scores.append((0, s0))
        best = (s0, model.params.detach().clone())
    for it in range(iters):
        t = targets[int(rng.integers(len(targets)))]
        grads.zero_()
        losses.append(run_loss(model, t, cosmo, conf, kmax=kmax, r_weight=r_weight, grads=grads))
        if model.symmetry == "cubic":
            project_cubic(grads, model.layers)
        opt.step(model, grads, state)
        if progress is not None:
            progress(it + 1, losses[-1])
        if validate is not None and (it + 1) % validate_every == 0:
            sc = validate(model)
            scores.append((it + 1, sc))
            if sc < best[0]:
                best = (sc, model.params.detach().clone())
    if best is not None:
        model.params.copy_(best[1])
        model.sync_bf16()
    return np.asarray(losses), scores
