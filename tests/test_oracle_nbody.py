"""The float64 N-body oracle: step factors, linear growth, reversibility and its adjoint.

Pins (integrator.py:114-170, :429-837; solver.py:385-444; the growth-matched factors of
docs/source/internals/integration_and_adjoint.md:62-95):
* Einstein-de Sitter closed forms of the drift and kick factors;
* a force that is exactly linear in the displacement (acc = 1.5 Omega_m disp, the Zel'dovich
  force) grows the displacement as D(a) to round-off -- the property the growth-matched
  factors are built for;
* the reverse sweep reconstructs the initial state and its gradient matches central finite
  differences of the forward map.
"""

import numpy as np
import pytest

from oracle import cosmology as oc
from oracle import nbody, spectral


def _setup(Om=0.3, ng=4, ratio=2, spacing=4.0, a_start=0.5, a_stop=1.0, maxstep=0.125):
    nm = ng * ratio
    cell = spacing * ng / nm
    masks = spectral.particle_nyquist_masks((nm,) * 3, (ng,) * 3, spacing, cell)
    a_lpt, a_nb, ga = oc.a_grids(a_start, a_stop, maxstep, 1 / 128)
    S = nbody.Setup((nm,) * 3, cell, ng ** 3, masks, Om, ga, a_lpt[1] - a_lpt[0])
    q = np.array(np.meshgrid(*[np.arange(ng) * ratio] * 3, indexing="ij")).reshape(3, -1).T
    return S, q.astype(np.int64), a_nb


def test_factors_match_einstein_de_sitter_closed_forms():
    S, _, a_nb = _setup(Om=1.0)
    for a0, a1 in zip(a_nb[:-1], a_nb[1:]):
        am = 0.5 * (a0 + a1)
        # D = a, G_D = a^1.5, G_K = 1.5 a
        assert S.drift_factor(am, a0, a1) == pytest.approx((a1 - a0) / am ** 1.5, rel=1e-10)
        assert S.kick_factor(a0, a0, am) == pytest.approx((am ** 1.5 - a0 ** 1.5) / (1.5 * a0), rel=1e-10)


class _LinearForce(nbody.Setup):
    def force(self, pmid, disp):
        return 1.5 * self.Om * disp


@pytest.mark.parametrize("Om", [0.3, 1.0])
def test_zeldovich_force_grows_as_linear_growth(Om):
    S0, q, a_nb = _setup(Om=Om, a_start=0.1, maxstep=0.05)
    S = _LinearForce(S0.mesh_shape, S0.cell, S0.ptcl_num, S0.masks, Om, S0.growth_a, S0.a_lpt_step)
    rng = np.random.default_rng(0)
    s = rng.normal(size=q.shape)
    t = S.table()
    a0, a1 = a_nb[0], a_nb[-1]
    disp0 = oc.growth(a0, t, S.growth_a) * s
    vel0 = oc.G_D(a0, t, S.growth_a, Om) * s
    d, v, acc = nbody.nbody(S, q, disp0, vel0, a_nb)
    np.testing.assert_allclose(d, oc.growth(a1, t, S.growth_a) * s, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(v, oc.G_D(a1, t, S.growth_a, Om) * s, rtol=1e-10, atol=1e-12)


def _loss_weights(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape), rng.normal(size=shape), rng.normal(size=shape)


def test_reverse_sweep_reconstructs_and_matches_finite_differences():
    S, q, a_nb = _setup()
    rng = np.random.default_rng(3)
    disp0 = rng.normal(size=q.shape) * 0.8
    vel0 = rng.normal(size=q.shape) * 0.3
    wd, wv, wa = _loss_weights(q.shape, 4)

    def loss(dsp, vl, Om=None):
        S2 = S if Om is None else nbody.Setup(S.mesh_shape, S.cell, S.ptcl_num, S.masks, Om, S.growth_a,
                                              S.a_lpt_step)
        d, v, a = nbody.nbody(S2, q, dsp, vl, a_nb)
        return float(np.sum(wd * d) + np.sum(wv * v) + np.sum(wa * a))

    d, v, a = nbody.nbody(S, q, disp0, vel0, a_nb)
    din, vin, gd, gv, gom = nbody.nbody_adj(S, q, d, v, a, wd, wv, wa, a_nb)
    np.testing.assert_allclose(din, disp0, atol=1e-10)
    np.testing.assert_allclose(vin, vel0, atol=1e-10)
    eps = 1e-5
    for (i, c) in [(0, 0), (5, 1), (17, 2), (40, 0)]:
        e = np.zeros_like(disp0)
        e[i, c] = eps
        fd = (loss(disp0 + e, vel0) - loss(disp0 - e, vel0)) / (2 * eps)
        assert gd[i, c] == pytest.approx(fd, rel=1e-6, abs=1e-8)
        fd = (loss(disp0, vel0 + e) - loss(disp0, vel0 - e)) / (2 * eps)
        assert gv[i, c] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    h = 1e-5
    fd = (loss(disp0, vel0, S.Om + h) - loss(disp0, vel0, S.Om - h)) / (2 * h)
    assert gom == pytest.approx(fd, rel=1e-5)
