"""HIP initial conditions vs the float64 oracle (oracle/ic.py).

Tolerance: float32 FFTs and float32 spectral products against float64: 1e-5 relative (L2)
on the linear modes, 2e-5 on the LPT displacements and velocities (a chain of ~10 transforms).
"""

import math

import numpy as np
import pytest
import torch

from oracle import ic as oic
from pmpp_amd import Configuration, Particles
from pmpp_amd.cosmology import LCDM, boltzmann
from pmpp_amd.cosmology.background import E2
from pmpp_amd.initial_conditions import linear_modes, lpt, particles_2lpt, white_noise
from pmpp_amd.nbody.integrator import growth_table

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("grid,spacing", [((16, 16, 16), 4.0), ((32, 16, 8), 2.5), ((64, 64, 64), 15.625)])
def test_white_noise_and_linear_modes_match_oracle(device, grid, spacing):
    conf = Configuration(spacing, grid, mesh_shape=1)
    cos = LCDM()
    w_real = white_noise(5, conf, real=True, device=device)
    np.testing.assert_array_equal(w_real.cpu().numpy(), oic.white_noise(5, grid).astype(np.float32))
    w = white_noise(5, conf, device=device)
    w0 = oic.ortho_rfftn(oic.white_noise(5, grid))
    assert _rel(w.cpu().numpy(), w0) < 1e-6
    tk, tT = boltzmann.transfer_table(cos)
    coef = boltzmann.power_coef(cos, cos.A_s_1e9)
    for a in (None, 0.5):
        m = linear_modes(w, cos, conf, a=a).cpu().numpy()
        D = 1.0 if a is None else float(growth_table(cos, conf).growth(a))
        ref = oic.linear_modes(w0, grid, spacing, tk, tT, coef * D * D, boltzmann.k_pivot(cos), cos.n_s)
        assert _rel(m, ref) < 1e-5
        assert m[0, 0, 0] == 0
    # real white noise in gives the same modes as its spectrum
    m_real = linear_modes(w_real, cos, conf).cpu().numpy()
    assert _rel(m_real, linear_modes(w, cos, conf).cpu().numpy()) < 1e-6


@pytest.mark.parametrize("grid,spacing,ratio,order", [((16, 16, 16), 8.0, 1, 2), ((32, 16, 8), 4.0, 2, 2),
                                                      ((32, 32, 32), 2.0, 1, 1), ((64, 64, 64), 4.0, 1, 2)])
def test_lpt_matches_oracle(device, grid, spacing, ratio, order):
    conf = Configuration(spacing, grid, mesh_shape=ratio, lpt_order=order, a_start=1 / 16)
    cos = LCDM()
    modes = linear_modes(white_noise(3, conf, device=device), cos, conf)
    ptcl = lpt(modes, cos, conf)
    tab = growth_table(cos, conf)
    a = conf.a_start
    growth = []
    for m in (1, 2):
        growth.append((float(tab.growth(a, m)), a * a * math.sqrt(float(E2(a, cos))) * float(tab.growth(a, m, 1))))
    d0, v0 = oic.lpt(modes.cpu().numpy().astype(np.complex128), grid, spacing, growth, order=order)
    lattice = Particles.gen_grid(conf, device=device).disp.cpu().numpy()
    disp = ptcl.disp.cpu().numpy() - lattice
    vel = ptcl.vel.cpu().numpy()
    assert _rel(disp, d0) < 2e-5
    assert _rel(vel, v0) < 2e-5
    assert ptcl.acc is None


def test_lpt_orders_zero_and_three(device):
    # tests/test_deep_lpt_orders.py:50-66
    conf = Configuration(1.0, (4, 4, 4), mesh_shape=1, lpt_order=0, a_start=1 / 32)
    modes = torch.zeros((4, 4, 3), dtype=torch.complex64, device=device)
    p = lpt(modes, LCDM(), conf)
    assert torch.all(p.disp == 0) and torch.all(p.vel == 0)
    with pytest.raises(NotImplementedError):
        lpt(modes, LCDM(), conf.replace(lpt_order=3))


def test_2lpt_particles_are_reanchored_with_velocities(device):
    conf = Configuration(4.0, (32, 32, 32), mesh_shape=2)
    p = particles_2lpt(conf, LCDM(), seed=2, jitter_seed=4, device=device, a=1.0)
    assert p.vel is not None and p.vel.shape == (conf.ptcl_num, 3)
    # nearest-cell anchors: |disp| <= half a cell, positions inside the box
    assert float(p.disp.abs().max()) <= conf.cell_size / 2 + 1e-4
    pos = p.pos().cpu().numpy()
    assert pos.min() >= 0 and pos.max() < conf.box_size[0]
    # at a = 1 the rms 2LPT displacement of a sigma8 = 0.8 field is a few Mpc/h
    ref = lpt(linear_modes(white_noise(2, conf.replace(a_start=1.0), device=device), LCDM(), conf.replace(a_start=1.0)),
              LCDM(), conf.replace(a_start=1.0))
    rms = float((ref.disp - Particles.gen_grid(conf, device=device).disp).pow(2).sum(1).mean().sqrt())
    assert 1.0 < rms < 20.0


def test_lpt_vjp_matches_oracle_directional_derivative(device):
    # <c, lpt(w + e v) - lpt(w - e v)> / 2e (float64 oracle) == <w_cot, v> (HIP adjoint)
    from pmpp_amd.initial_conditions import lpt_vjp

    grid, spacing = (16, 16, 16), 4.0
    conf = Configuration(spacing, grid, mesh_shape=1, a_start=0.5)
    cos = LCDM()
    w = white_noise(9, conf, real=True, device=device)
    rng = np.random.default_rng(1)
    dcot = rng.normal(size=(conf.ptcl_num, 3))
    vcot = rng.normal(size=(conf.ptcl_num, 3))
    g = lpt_vjp(w, cos, conf, torch.from_numpy(dcot).float().to(device),
                torch.from_numpy(vcot).float().to(device)).cpu().numpy().astype(np.float64)
    tk, tT = boltzmann.transfer_table(cos)
    coef = boltzmann.power_coef(cos, cos.A_s_1e9)
    tab = growth_table(cos, conf)
    a = conf.a_start
    growth = [(float(tab.growth(a, m)), a * a * math.sqrt(float(E2(a, cos))) * float(tab.growth(a, m, 1)))
              for m in (1, 2)]

    def f(x):
        modes = oic.linear_modes(oic.ortho_rfftn(x), grid, spacing, tk, tT, coef, boltzmann.k_pivot(cos), cos.n_s)
        d, v = oic.lpt(modes, grid, spacing, growth, order=2)
        return np.sum(d * dcot) + np.sum(v * vcot)

    w0 = w.cpu().numpy().astype(np.float64)
    for s in range(3):
        v = np.random.default_rng(10 + s).normal(size=grid)
        eps = 1e-3
        fd = (f(w0 + eps * v) - f(w0 - eps * v)) / (2 * eps)
        assert np.sum(g * v) == pytest.approx(fd, rel=2e-5)


def test_pk_vjp_matches_oracle_finite_differences(device):
    from oracle import power_spectrum as ops
    from pmpp_amd.analysis import delta_to_pk, density_to_pk_vjp, delta_to_pk_vjp

    shape = (16, 16, 16)
    conf = Configuration(2.0, shape, mesh_shape=1)
    rng = np.random.default_rng(3)
    delta = rng.normal(size=shape) * 0.5
    cot = rng.normal(size=len(delta_to_pk(torch.from_numpy(delta.astype(np.float32)).to(device), conf)[1]))
    g = delta_to_pk_vjp(torch.from_numpy(delta.astype(np.float32)).to(device), conf, cot).cpu().numpy()
    v = rng.normal(size=shape)
    eps = 1e-4
    fd = (np.sum(cot * ops.delta_to_pk(delta + eps * v, conf.cell_size)[1])
          - np.sum(cot * ops.delta_to_pk(delta - eps * v, conf.cell_size)[1])) / (2 * eps)
    assert np.sum(g * v) == pytest.approx(fd, rel=1e-5)
    # density input: same up to the mean's constant, which a zero-mean direction does not see
    rho = 1 + delta
    gd = density_to_pk_vjp(torch.from_numpy(rho.astype(np.float32)).to(device), conf, cot).cpu().numpy()
    v0 = v - v.mean()
    fdd = (np.sum(cot * ops.power_spectrum(rho + eps * v0, conf.cell_size)[1])
           - np.sum(cot * ops.power_spectrum(rho - eps * v0, conf.cell_size)[1])) / (2 * eps)
    assert np.sum(gd * v0) == pytest.approx(fdd, rel=1e-4)


def test_pk_loss_gradient_wrt_white_noise_matches_oracle_finite_differences(device):
    # the differentiable run at test size: white noise -> linear modes -> 2LPT -> 3 KDK steps
    # -> CIC density -> MSE of log P(k) against a fixed target; the HIP reverse chain
    # (P(k) adjoint, scatter adjoint, N-body reverse sweep, LPT adjoint) against float64 central
    # differences of the oracle's forward chain on 8 random noise voxels.  The CIC density is
    # piecewise linear in the positions, so the loss has kinks where a particle crosses a cell
    # face; eps = 1e-5 (float64) keeps the difference interval clear of them.
    from oracle import cic as ocic, nbody as onb, power_spectrum as ops, spectral
    from pmpp_amd.analysis import density_to_pk, density_to_pk_vjp
    from pmpp_amd.cic import scatter
    from pmpp_amd.initial_conditions import lpt_vjp
    from pmpp_amd.nbody.solver import nbody

    grid, spacing = (16, 16, 16), 4.0
    conf = Configuration(spacing, grid, mesh_shape=2, a_start=0.25, a_nbody_maxstep=0.25)
    cos = LCDM()
    w = white_noise(1, conf, real=True, device=device)
    p0 = lpt(linear_modes(w, cos, conf), cos, conf)
    disp0 = p0.disp.clone().requires_grad_(True)
    vel0 = p0.vel.clone().requires_grad_(True)
    fin = nbody(p0.replace(disp=disp0, vel=vel0), cos, conf)
    dens = scatter(fin, conf)
    _, pk, _ = density_to_pk(dens.detach(), conf)
    pk = pk.cpu().numpy().astype(np.float64)
    nb = min(32, len(pk))
    target = 1.2 * pk[:nb] * np.linspace(0.9, 1.1, nb)
    pk_cot = np.zeros(len(pk))
    pk_cot[:nb] = 2 * (np.log(pk[:nb]) - np.log(target)) / (nb * pk[:nb])
    dens.backward(density_to_pk_vjp(dens.detach(), conf, pk_cot))
    g = lpt_vjp(w, cos, conf, disp0.grad, vel0.grad).cpu().numpy().astype(np.float64)

    tk, tT = boltzmann.transfer_table(cos)
    coef = boltzmann.power_coef(cos, cos.A_s_1e9)
    tab = growth_table(cos, conf)
    a = conf.a_start
    growth = [(float(tab.growth(a, m)), a * a * math.sqrt(float(E2(a, cos))) * float(tab.growth(a, m, 1)))
              for m in (1, 2)]
    base = Particles.gen_grid(conf, device="cpu")
    pmid = base.pmid.numpy().astype(np.int64)
    lattice = base.disp.numpy().astype(np.float64)
    masks = spectral.particle_nyquist_masks(conf.mesh_shape, grid, spacing, conf.cell_size)
    S = onb.Setup(conf.mesh_shape, conf.cell_size, conf.ptcl_num, masks, cos.Omega_m, conf.growth_a, conf.a_lpt_step)
    val = spectral.default_value(conf.ptcl_num, conf.mesh_shape, conf.ptcl_num)

    def loss(x):
        modes = oic.linear_modes(oic.ortho_rfftn(x), grid, spacing, tk, tT, coef, boltzmann.k_pivot(cos), cos.n_s)
        d, v = oic.lpt(modes, grid, spacing, growth, order=2)
        fd, _, _ = onb.nbody(S, pmid, lattice + d, v, conf.a_nbody)
        rho = ocic.scatter(pmid, fd, val, np.zeros(conf.mesh_shape), conf.cell_size, conf.mesh_shape)
        p = ops.power_spectrum(rho, conf.cell_size, "CIC")[1][:nb]
        return np.mean((np.log(p) - np.log(target)) ** 2)

    w0 = w.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(8)
    idx = [tuple(rng.integers(0, 16, size=3)) for _ in range(8)]
    fds = []
    for i in idx:
        e = np.zeros(grid)
        e[i] = 1e-5
        fds.append((loss(w0 + e) - loss(w0 - e)) / 2e-5)
    got = np.array([g[i] for i in idx])
    fds = np.array(fds)
    print("gradient vs FD:", np.c_[got, fds])
    np.testing.assert_allclose(got, fds, rtol=1e-3, atol=1e-3 * np.abs(fds).max())
