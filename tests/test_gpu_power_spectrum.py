"""HIP shell power spectra vs the float64 oracle (analysis/power_spectrum.py restated).

Tolerance: the float32 FFT and float32 per-mode power, summed in float64, against the float64
oracle: 1e-5 relative per shell (the north-star P(k) bar is 1e-4).
"""

import numpy as np
import pytest
import torch

from oracle import power_spectrum as ops
from pmpp_amd import Configuration
from pmpp_amd.analysis import (delta_to_cross_correlation, delta_to_pk, density_to_cross_correlation,
                               density_to_pk)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,mas", [((16, 16, 16), "CIC"), ((32, 16, 8), None), ((8, 8, 8), "TSC"),
                                       ((64, 64, 64), "CIC")])
def test_delta_to_pk_matches_oracle(device, shape, mas):
    conf = Configuration(1.5, shape, mesh_shape=1)
    rng = np.random.default_rng(sum(shape))
    delta = rng.normal(size=shape).astype(np.float32)
    k, pk, nm = delta_to_pk(torch.from_numpy(delta).to(device), conf, mas=mas)
    k0, pk0, nm0 = ops.delta_to_pk(delta.astype(np.float64), conf.cell_size, mas or None)
    np.testing.assert_array_equal(nm.cpu().numpy(), nm0)
    np.testing.assert_allclose(k.cpu().numpy(), k0, rtol=1e-6)
    np.testing.assert_allclose(pk.cpu().numpy(), pk0, rtol=1e-5)


def test_density_to_pk_matches_oracle(device):
    shape = (32, 32, 32)
    conf = Configuration(2.0, shape, mesh_shape=1)
    rng = np.random.default_rng(4)
    rho = (1.0 + 0.3 * rng.normal(size=shape)).astype(np.float32)
    k, pk, nm = density_to_pk(torch.from_numpy(rho).to(device), conf)
    k0, pk0, nm0 = ops.power_spectrum(rho.astype(np.float64), conf.cell_size, "CIC")
    np.testing.assert_allclose(pk.cpu().numpy(), pk0, rtol=1e-5)


def test_cross_correlation(device):
    shape = (16, 16, 16)
    conf = Configuration(1.0, shape, mesh_shape=1)
    rng = np.random.default_rng(5)
    a = rng.normal(size=shape).astype(np.float32)
    b = (0.6 * a + 0.8 * rng.normal(size=shape)).astype(np.float32)
    A, B = torch.from_numpy(a).to(device), torch.from_numpy(b).to(device)
    k, r, pab, pa, pb, nm = delta_to_cross_correlation(A, B, conf)
    _, pa0, _ = ops.delta_to_pk(a.astype(np.float64), 1.0)
    _, pb0, _ = ops.delta_to_pk(b.astype(np.float64), 1.0)
    # cross power from the polarisation identity on the oracle: P_ab = (P_{a+b} - P_a - P_b) / 2
    _, ps0, _ = ops.delta_to_pk((a + b).astype(np.float64), 1.0)
    np.testing.assert_allclose(pa.cpu().numpy(), pa0, rtol=1e-5)
    np.testing.assert_allclose(pb.cpu().numpy(), pb0, rtol=1e-5)
    np.testing.assert_allclose(pab.cpu().numpy(), (ps0 - pa0 - pb0) / 2, rtol=1e-4, atol=1e-5 * pa0.max())
    np.testing.assert_allclose(r.cpu().numpy(), (ps0 - pa0 - pb0) / 2 / np.sqrt(pa0 * pb0), rtol=1e-4, atol=1e-5)
    _, r_self, *_ = density_to_cross_correlation(A + 2.0, A + 2.0, conf)
    np.testing.assert_allclose(r_self.cpu().numpy(), 1.0, rtol=1e-6)


def test_single_mode_power_is_exact(device):
    # delta = 2 A cos(2 pi n.x / N): all power in shell |n| with P = A^2 N_m cell^3 (both +-n
    # counted once each through the Hermitian multiplicity)
    shape = (16, 16, 16)
    conf = Configuration(1.0, shape, mesh_shape=1)
    A = 0.1
    x = np.indices(shape)
    n = np.array([1, 2, 2])  # |n| = 3
    delta = (2 * A * np.cos(2 * np.pi * np.tensordot(n, x, axes=1) / 16)).astype(np.float32)
    k, pk, nm = delta_to_pk(torch.from_numpy(delta).to(device), conf, mas=None)
    pk = pk.cpu().numpy()
    nm = nm.cpu().numpy()
    s = 3 - 1
    assert pk[s] == pytest.approx(2 * A * A * 16 ** 3 / nm[s], rel=1e-5)
    assert np.abs(np.delete(pk, s)).max() < 1e-9


def test_cic_window_of_a_randomly_placed_particle_on_the_gpu(device):
    # the device shell sums against the closed form the oracle is pinned by: averaged over its
    # offset in a cell, a CIC particle's per-mode power is prod_i [1 - 2/3 sin^2(pi m_i/N)],
    # and the CIC deconvolution divides it by prod_i sinc^4(m_i/N)
    from pmpp_amd.cic import scatter
    from pmpp_amd import Particles
    from tests.test_oracle_power_spectrum import _expected_shell_mean

    n, cell = 16, 1.5
    conf = Configuration(cell, (n,) * 3, mesh_shape=1)
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    offs = np.array([[a, b, c] for a in g for b in g for c in g]) * cell
    pks = []
    for o in offs:
        p = Particles(conf, torch.tensor([[3, 5, 7]], dtype=torch.int16, device=device),
                      torch.from_numpy(o[None].astype(np.float32)).to(device))
        mesh = scatter(p, conf, val=torch.ones(1, device=device))
        pks.append(delta_to_pk(mesh, conf, mas="CIC")[1].cpu().numpy())

    def per_mode(a, b, c):
        return np.prod([(1 - 2.0 / 3.0 * np.sin(np.pi * x / n) ** 2) / np.sinc(x / n) ** 4 for x in (a, b, c)])

    pk = np.mean(pks, axis=0)
    exp = _expected_shell_mean(n, per_mode)[:len(pk)] * cell ** 3 / n ** 3
    np.testing.assert_allclose(pk, exp, rtol=1e-5)


def test_cross_correlation_vjp_matches_central_differences(device):
    """d/d(rho_a) of <r_cot, r(k)> (cross and auto spectrum adjoints) against float64 central
    differences of the cross-correlation along a random density perturbation (zero mean)."""
    from pmpp_amd import Configuration
    from pmpp_amd.analysis import density_to_cross_correlation, density_to_cross_correlation_vjp

    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2)
    rng = np.random.default_rng(5)
    shape = conf.mesh_shape
    a = 1 + 0.3 * rng.normal(size=shape)
    b = a + 0.2 * rng.normal(size=shape)
    a, b = a / a.mean(), b / b.mean()
    ta = torch.as_tensor(a.astype(np.float32), device=device)
    tb = torch.as_tensor(b.astype(np.float32), device=device)
    _, r, _, _, _, _ = density_to_cross_correlation(ta, tb, conf)
    r_cot = rng.normal(size=len(r))
    g = density_to_cross_correlation_vjp(ta, tb, conf, r_cot).double().cpu().numpy()
    v = rng.normal(size=shape)
    v -= v.mean()
    eps = 1e-3

    def f(x):
        _, rr, _, _, _, _ = density_to_cross_correlation(torch.as_tensor(x.astype(np.float32), device=device), tb, conf)
        return float((rr.double().cpu().numpy() * r_cot).sum())

    fd = (f(a + eps * v) - f(a - eps * v)) / (2 * eps)
    assert float((g * v).sum()) == pytest.approx(fd, rel=2e-2)
