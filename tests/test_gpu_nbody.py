"""HIP KDK N-body and its reverse sweep vs the float64 oracle.

Tolerances: float32 trajectory vs float64 over 4 KDK steps, 1e-4 relative (L2) in positions
and velocities; gradients through the hand-written reverse sweep (pm_kick_adj /
pm_drift_adj / pm_gravity_vjp with state reconstruction) within 1e-3 relative of the oracle's
adjoint, which is itself pinned by finite differences (tests/test_oracle_nbody.py).
"""

import numpy as np
import pytest
import torch

from oracle import cosmology as oc
from oracle import nbody as onb
from oracle import spectral
from pmpp_amd import Configuration, Particles
from pmpp_amd.initial_conditions import LCDM
from pmpp_amd.nbody.solver import nbody, nbody_adj

pytestmark = pytest.mark.gpu


def _case(device, ng=8, ratio=2, spacing=4.0, seed=5):
    conf = Configuration(spacing, (ng,) * 3, mesh_shape=ratio, a_start=0.5, a_stop=1.0, a_nbody_maxstep=0.125)
    base = Particles.gen_grid(conf, device=device)
    rng = np.random.default_rng(seed)
    disp = (base.disp.cpu().numpy() + rng.normal(size=base.disp.shape) * 0.6 * spacing).astype(np.float32)
    vel = (rng.normal(size=base.disp.shape) * 0.3 * spacing).astype(np.float32)
    ptcl = base.replace(disp=torch.as_tensor(disp, device=device), vel=torch.as_tensor(vel, device=device))
    S = onb.Setup(conf.mesh_shape, conf.cell_size, conf.ptcl_num,
                  spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                                  conf.cell_size),
                  0.3, conf.growth_a, conf.a_lpt_step)
    return conf, ptcl, S


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_matches_oracle(device):
    conf, ptcl, S = _case(device)
    out = nbody(ptcl, LCDM(), conf)
    pm = ptcl.pmid.cpu().numpy().astype(np.int64)
    d, v, a = onb.nbody(S, pm, ptcl.disp.cpu().numpy(), ptcl.vel.cpu().numpy(), conf.a_nbody)
    assert _rel(out.disp.cpu().numpy(), d) < 1e-4
    assert _rel(out.vel.cpu().numpy(), v) < 1e-4
    assert _rel(out.acc.cpu().numpy(), a) < 1e-4


def test_reverse_sweep_reconstructs_the_initial_state(device):
    conf, ptcl, S = _case(device, seed=6)
    out = nbody(ptcl, LCDM(), conf)
    zero = torch.zeros_like(out.disp)
    pin, _, _ = nbody_adj(out, out.replace(disp=zero, vel=zero, acc=zero), LCDM(), conf)
    scale = np.abs(ptcl.disp.cpu().numpy()).max()
    np.testing.assert_allclose(pin.disp.cpu().numpy(), ptcl.disp.cpu().numpy(), atol=2e-5 * scale)
    np.testing.assert_allclose(pin.vel.cpu().numpy(), ptcl.vel.cpu().numpy(), atol=2e-5 * scale)


def test_gradients_match_oracle_adjoint(device):
    conf, ptcl, S = _case(device, seed=7)
    rng = np.random.default_rng(8)
    wd, wv, wa = (rng.normal(size=ptcl.disp.shape) for _ in range(3))
    disp = ptcl.disp.clone().requires_grad_(True)
    vel = ptcl.vel.clone().requires_grad_(True)
    om = torch.tensor(0.3, dtype=torch.float64, device=device, requires_grad=True)

    class C:
        Omega_m = om

    out = nbody(ptcl.replace(disp=disp, vel=vel), C, conf)
    t = lambda w: torch.as_tensor(w, dtype=torch.float32, device=device)  # noqa: E731
    loss = (out.disp * t(wd)).sum() + (out.vel * t(wv)).sum() + (out.acc * t(wa)).sum()
    loss.backward()
    pm = ptcl.pmid.cpu().numpy().astype(np.int64)
    d, v, a = onb.nbody(S, pm, ptcl.disp.cpu().numpy(), ptcl.vel.cpu().numpy(), conf.a_nbody)
    _, _, gd, gv, gom = onb.nbody_adj(S, pm, d, v, a, wd, wv, wa, conf.a_nbody)
    assert _rel(disp.grad.cpu().numpy(), gd) < 1e-3
    assert _rel(vel.grad.cpu().numpy(), gv) < 1e-3
    assert float(om.grad) == pytest.approx(gom, rel=1e-3)
