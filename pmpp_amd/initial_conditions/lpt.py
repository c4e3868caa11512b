"""Lagrangian perturbation theory initial conditions (PM++ ``initial_conditions/lpt.py:339-449``).

``lpt(modes, cosmo, conf)`` returns particles on the Lagrangian grid at ``conf.a_start`` with
displacements and canonical velocities to ``conf.lpt_order`` (0, 1 or 2):

  phi_1 = laplace^-1(modes / V_cell)
  L     = sum_{i<j} phi_1,ii phi_1,jj - phi_1,ij^2          (the 2LPT source, ``_L`` :239-302)
  phi_2 = laplace^-1(rfftn(L))
  disp += D_m(a) grad_m,   vel += a^2 H(a) dD_m/dln a grad_m,   grad_m = irfftn(-i k phi_m)

with the strain and gradient symbols zeroing the Nyquist wavenumber where the reference does
(``_strain`` :17-43 for i != j, ``neg_grad`` gravity.py:355-391).  Every spectral product is
one ``pm_ic_spectral_op`` launch and every transform the library's own rFFT; the growth
factors are the host float64 table's, cast to float32 as the reference casts them.
Particles come from ``Particles.gen_grid``, so particle n samples grid point n (the
single-device branch of ``_sample_lpt_field`` :55-81).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib
from .._torch import ensure_init, ptr, require_device, stream_handle
from ..cosmology.background import E2
from ..nbody.integrator import growth_table
from ..nbody.particles import Particles
from ..numerics import irfftn, rfftn
from .modes import KGrid, LAPLACE, NEG_GRAD, SCALE, STRAIN


def _acc_product(acc, a, b, coef, init):
    _lib.call("pm_ic_product_acc", ptr(acc), ptr(a), ptr(b), acc.numel(), float(coef), int(init),
              stream_handle(acc.device))


def second_order_source(kg, pot):
    """L = sum_{i<j} phi_ii phi_jj - phi_ij^2 of a potential spectrum (lpt.py:239-302, one
    potential, diagonal strains cached), accumulated in the reference's order."""
    shape = kg.shape
    scratch = torch.empty_like(pot)
    diag = []
    for i in range(3):
        kg.op(pot, STRAIN, i, i, out=scratch)
        diag.append(irfftn(scratch, s=shape))
    L = torch.empty(shape, dtype=torch.float32, device=pot.device)
    first = True
    for i in range(3):
        for j in range(2, i, -1):
            _acc_product(L, diag[i], diag[j], 1.0, first)
            first = False
    del diag
    for i in range(2):
        for j in range(i + 1, 3):
            kg.op(pot, STRAIN, i, j, out=scratch)
            t = irfftn(scratch, s=shape)
            _acc_product(L, t, t, -1.0, False)
    return L


def lpt(modes, cosmo, conf):
    """Particles at ``conf.a_start`` with LPT displacements and velocities."""
    if conf.dim != 3:
        raise ValueError(f"dim={conf.dim} not supported")
    if conf.lpt_order not in (0, 1, 2, 3):
        raise ValueError(f"lpt_order={conf.lpt_order} not supported")
    require_device(modes, "modes")
    dev = modes.device
    ensure_init(dev)
    shape = tuple(conf.ptcl_grid_shape)
    if tuple(modes.shape) != (shape[0], shape[1], shape[2] // 2 + 1):
        raise ValueError("modes do not match the particle grid")
    if conf.lpt_order > 2:
        raise NotImplementedError("third-order LPT is not implemented (nor in the reference)")
    kg = KGrid(shape, conf.ptcl_spacing, dev)
    cell_vol = conf.ptcl_spacing ** 3
    pots = []
    if conf.lpt_order > 0:
        src = kg.op(modes.to(torch.complex64).contiguous(), SCALE, scale=1.0 / cell_vol)
        pots.append(kg.op(src, LAPLACE))
        del src
    if conf.lpt_order > 1:
        L = second_order_source(kg, pots[0])
        src2 = rfftn(L)
        del L
        pots.append(kg.op(src2, LAPLACE))
        del src2
    ptcl = Particles.gen_grid(conf, vel=True, device=dev)
    disp = ptcl.disp.contiguous().clone()
    vel = ptcl.vel.contiguous().clone()
    a = float(conf.a_start)
    tab = growth_table(cosmo, conf)
    scratch = torch.empty_like(pots[0]) if pots else None
    n = len(ptcl)
    for order, pot in enumerate(pots, start=1):
        D = np.float32(tab.growth(a, order))
        a2HDp = np.float32(a ** 2 * math.sqrt(float(E2(a, tab.cosmo))) * float(tab.growth(a, order, 1)))
        for i in range(3):
            kg.op(pot, NEG_GRAD, i, out=scratch)
            g = irfftn(scratch, s=shape)
            _lib.call("pm_ic_lpt_accumulate", ptr(disp), ptr(vel), ptr(g), n, i, float(D), float(a2HDp),
                      stream_handle(dev))
    return ptcl.replace(disp=disp, vel=vel, acc=None)


def particles_2lpt(conf, cosmo, seed=0, jitter_seed=None, device="cuda", a=None):
    """2LPT particles from seeded white noise at ``a`` (default ``conf.a_start``), re-anchored
    to their nearest mesh cell, with an optional uniform lattice jitter of +-l_p/2 per axis (a
    synthetic workload that breaks the perfect lattice)."""
    from .modes import linear_modes, white_noise

    lconf = conf if a is None else conf.replace(a_start=float(a))
    modes = linear_modes(white_noise(seed, lconf, device=device), cosmo, lconf)
    ptcl = lpt(modes, cosmo, lconf)
    pos = ptcl.pos(dtype=torch.float64, wrap=False)
    if jitter_seed is not None:
        rng = np.random.Generator(np.random.PCG64(jitter_seed))
        u = rng.uniform(-0.5, 0.5, size=(conf.ptcl_num, 3))
        pos = pos + torch.from_numpy(u).to(pos.device) * conf.ptcl_spacing
    box = torch.tensor(conf.box_size, dtype=torch.float64, device=pos.device)
    pos = torch.remainder(pos, box)
    return Particles.from_pos(conf, pos, vel=ptcl.vel)


def _axpy(acc, a, coef, init=False):
    """acc = (init ? 0 : acc) + coef a on float (or complex, as float pairs) tensors."""
    av = torch.view_as_real(a) if torch.is_complex(a) else a
    cv = torch.view_as_real(acc) if torch.is_complex(acc) else acc
    _lib.call("pm_ic_product_acc", ptr(cv), ptr(av), None, cv.numel(), float(coef), int(init),
              stream_handle(acc.device))


def lpt_vjp(noise, cosmo, conf, disp_cot, vel_cot=None):
    """Cotangent of the real white-noise field for ``lpt(linear_modes(noise))``.

    ``disp_cot`` / ``vel_cot`` are (N, 3) cotangents of the LPT displacements and velocities in
    grid order.  Every map is a real convolution with a Hermitian symbol (the noise filter
    sqrt(V P) / V_cell, the inverse Laplacian, the strains, -i k), so its transpose is the
    convolution with the conjugate symbol; the 2LPT source L = sum_{i<j} s_ii s_jj - s_ij^2 pulls
    a cotangent lam back to sum_i T_ii(lam sum_{j != i} s_jj) - 2 sum_{i<j} T_ij(lam s_ij) on the
    first-order potential (the reverse of lpt.py:239-302, which the reference takes by
    autodiff)."""
    from .modes import linear_modes

    require_device(noise, "noise")
    dev = noise.device
    shape = tuple(conf.ptcl_grid_shape)
    if tuple(noise.shape) != shape or noise.dtype != torch.float32:
        raise ValueError("noise must be the real float32 field on the particle grid")
    if conf.lpt_order > 2:
        raise NotImplementedError("third-order LPT is not implemented (nor in the reference)")
    kg = KGrid(shape, conf.ptcl_spacing, dev)
    cell_vol = conf.ptcl_spacing ** 3
    n = math.prod(shape)
    dc = disp_cot.detach().to(torch.float32).contiguous()
    vc = None if vel_cot is None else vel_cot.detach().to(torch.float32).contiguous()
    a = float(conf.a_start)
    tab = growth_table(cosmo, conf)
    nzh = shape[2] // 2 + 1
    grad1 = torch.zeros((shape[0], shape[1], nzh), dtype=torch.complex64, device=dev)
    if conf.lpt_order == 0:
        return torch.zeros(shape, dtype=torch.float32, device=dev)
    # phi_1 spectrum (recomputed)
    modes = linear_modes(noise, cosmo, conf)
    phi1 = kg.op(kg.op(modes, SCALE, scale=1.0 / cell_vol), LAPLACE)
    scratch = torch.empty_like(phi1)
    c = torch.empty(shape, dtype=torch.float32, device=dev)

    def pot_cot(order):
        """sum_i K_i^T c_i with c_i = D disp_cot_i + V vel_cot_i, K_i^T = +i k_i = -(-i k_i)."""
        D = float(np.float32(tab.growth(a, order)))
        V = float(np.float32(a ** 2 * math.sqrt(float(E2(a, tab.cosmo))) * float(tab.growth(a, order, 1))))
        out = torch.zeros_like(phi1)
        for i in range(3):
            _axpy(c.view(-1), dc[:, i].contiguous(), D, init=True)
            if vc is not None:
                _axpy(c.view(-1), vc[:, i].contiguous(), V)
            kg.op(rfftn(c), NEG_GRAD, i, out=scratch)
            _axpy(out, scratch, -1.0)
        return out

    _axpy(grad1, pot_cot(1), 1.0)
    if conf.lpt_order > 1:
        lam = irfftn(kg.op(pot_cot(2), LAPLACE), s=shape)  # the inverse Laplacian is symmetric
        strains = {}
        for i in range(3):
            for j in range(i, 3):
                kg.op(phi1, STRAIN, i, j, out=scratch)
                strains[(i, j)] = irfftn(scratch, s=shape)
        t = torch.empty(shape, dtype=torch.float32, device=dev)
        for i in range(3):
            others = [j for j in range(3) if j != i]
            _acc_product(t, lam, strains[(others[0], others[0])], 1.0, True)
            _acc_product(t, lam, strains[(others[1], others[1])], 1.0, False)
            kg.op(rfftn(t), STRAIN, i, i, out=scratch)
            _axpy(grad1, scratch, 1.0)
        for i in range(2):
            for j in range(i + 1, 3):
                _acc_product(t, lam, strains[(i, j)], 1.0, True)
                kg.op(rfftn(t), STRAIN, i, j, out=scratch)
                _axpy(grad1, scratch, -2.0)
    # back through phi_1 = Lap^-1 (filter(ortho rfftn(noise)) / V_cell): all symbols real
    g = kg.op(grad1, LAPLACE)
    g = kg.op(g, SCALE, scale=1.0 / cell_vol)
    from ..cosmology import boltzmann

    tk, tT = boltzmann.transfer_table(cosmo)
    tk_d = torch.from_numpy(tk.astype(np.float32)).to(dev)
    tT_d = torch.from_numpy(tT.astype(np.float32)).to(dev)
    coef = boltzmann.power_coef(cosmo, cosmo.A_s_1e9)
    _lib.call("pm_ic_linear_modes", ptr(g), _lib.int3(shape), kg.ptrs, ptr(tk_d), ptr(tT_d), len(tk), float(coef),
              float(boltzmann.k_pivot(cosmo)), float(cosmo.n_s), float(math.sqrt(conf.box_vol)), stream_handle(dev))
    g = kg.op(g, SCALE, scale=1.0 / math.sqrt(n))
    return irfftn(g, s=shape)
