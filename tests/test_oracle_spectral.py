"""Pins the spectral force oracle (oracle/spectral.py)."""

import numpy as np
import pytest

from oracle import cic, spectral


def _dft3_bruteforce(x):
    """O(N^2) DFT along each axis by explicit exponent matrices (no FFT library)."""
    out = x.astype(np.complex128)
    for ax in range(3):
        n = x.shape[ax]
        k = np.arange(n)
        F = np.exp(-2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(F, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


def test_oracle_rfftn_matches_bruteforce_dft():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(4, 6, 8))
    full = _dft3_bruteforce(x)
    np.testing.assert_allclose(spectral.rfftn(x), full[..., :5], atol=1e-11)
    np.testing.assert_allclose(np.fft.rfftn(x), full[..., :5], atol=1e-11)
    np.testing.assert_allclose(spectral.irfftn(spectral.rfftn(x), x.shape), x, atol=1e-13)


def test_chunked_cic_equals_single_pass(monkeypatch):
    # the particle chunking of the large-mesh oracle changes only the summation order
    rng = np.random.default_rng(4)
    ms, n = (8, 6, 10), 1000
    pm = rng.integers(0, 6, (n, 3))
    dp = rng.normal(size=(n, 3)) * 2.0
    val = rng.random(n)
    valid = rng.random(n) > 0.2
    ref = cic.scatter(pm, dp, val, np.zeros(ms), 0.7, ms, valid=valid)
    meshes = [rng.normal(size=ms) for _ in range(3)]
    gref = np.stack([cic.gather(pm, dp, m, 0.7, ms, valid=valid) for m in meshes], axis=-1)
    monkeypatch.setattr(spectral, "CHUNK", 97)
    np.testing.assert_allclose(spectral.scatter_chunked(pm, dp, val, ms, 0.7, valid=valid), ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(spectral.gather_chunked(pm, dp, meshes, 0.7, ms, valid=valid), gref)
    np.testing.assert_array_equal(cic.gather_many(pm, dp, meshes, 0.7, ms, valid=valid), gref)


def test_poisson_symbol_and_nyquist_support_closed_forms():
    # tests/test_deep_gravity_spectral.py:31-113
    ms, cell = (4, 4, 4), 1.25
    kx, ky, kz = spectral.kvec(ms, cell)
    np.testing.assert_allclose(kx, 2 * np.pi / (4 * cell) * np.array([0, 1, -2, -1]))
    np.testing.assert_allclose(kz, 2 * np.pi / (4 * cell) * np.array([0, 1, 2]))
    g = spectral.green_multiplier(ms, cell)
    assert g[0, 0, 0] == 0
    k2 = kx[1] ** 2
    assert g[1, 0, 0] == pytest.approx(-1 / k2)
    nk = spectral.neg_k(ms, cell)
    # Nyquist derivative zeroed on every axis, other entries -k
    assert nk[0][2] == 0 and nk[2][2] == 0
    np.testing.assert_allclose(nk[0][[0, 1, 3]], -kx[[0, 1, 3]])
    # mesh_shape=2 keeps |m| <= N/4: exactly the same support as PM++'s float32 test of
    # test_particle_nyquist_filter_and_cached_spectral_operators_have_exact_support
    masks = spectral.particle_nyquist_masks((8, 8, 8), (4, 4, 4), 1.0, 0.5)
    np.testing.assert_array_equal(masks[0], [1, 1, 1, 0, 0, 0, 1, 1])
    np.testing.assert_array_equal(masks[2], [1, 1, 1, 0, 0])
    assert spectral.particle_nyquist_masks((4, 4, 4), (4, 4, 4), 1.0, 1.0) == ()


def test_single_plane_wave_force_closed_form():
    # S = A cos(k0 . x)  =>  phi = -S/k0^2 ... a = -grad phi = -A k0 sin(k0.x) / k0^2
    ms, cell = (16, 8, 8), 0.5
    A, om = 0.01, 0.3
    x = np.arange(16) * cell
    L = 16 * cell
    k0 = 2 * np.pi * 2 / L
    src_real = A * np.cos(k0 * x)[:, None, None] * np.ones(ms)
    density = 1 + src_real / (1.5 * om)
    fx, fy, fz = spectral.force_meshes(density, om, ms, cell)
    np.testing.assert_allclose(fx, (-A * np.sin(k0 * x) / k0)[:, None, None] * np.ones(ms), atol=1e-14)
    np.testing.assert_allclose(fy, 0, atol=1e-15)
    np.testing.assert_allclose(fz, 0, atol=1e-15)


def test_potential_plane_wave_and_gradient_consistency():
    # S = A cos(k0 . x) along y  =>  phi = -A cos(k0 y) / k0^2
    ms, cell = (8, 16, 8), 0.5
    A, om = 0.02, 0.3
    y = np.arange(16) * cell
    k0 = 2 * np.pi * 3 / (16 * cell)
    density = 1 + (A * np.cos(k0 * y))[None, :, None] * np.ones(ms) / (1.5 * om)
    phi = spectral.potential(density, om, ms, cell)
    np.testing.assert_allclose(phi, (-A * np.cos(k0 * y) / k0 ** 2)[None, :, None] * np.ones(ms), atol=1e-14)
    # on a random density with particle-Nyquist masks, the force meshes are the spectral
    # gradient of the potential (F_i = irfftn(-i k_i phi_hat) with the cached -k_i)
    rng = np.random.default_rng(3)
    ms, cell = (16, 8, 16), 1.0
    masks = spectral.particle_nyquist_masks(ms, (8, 4, 8), 2.0, cell)
    dens = 1 + 0.3 * rng.normal(size=ms)
    phi = spectral.potential(dens, om, ms, cell, masks)
    forces = spectral.force_meshes(dens, om, ms, cell, masks)
    nk = spectral.neg_k(ms, cell)
    ph = spectral.rfftn(phi)
    shapes = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    for i in range(3):
        ref = spectral.irfftn(1j * nk[i].reshape(shapes[i]) * ph, ms)
        np.testing.assert_allclose(forces[i], ref, atol=1e-12 * np.abs(ref).max())
    assert abs(phi.mean()) < 1e-15  # the k = 0 mode is removed


def _lattice(ng, ratio):
    q = np.stack(np.meshgrid(*[np.arange(ng) * ratio] * 3, indexing="ij"), -1).reshape(-1, 3)
    return q


def test_uniform_lattice_has_zero_force_and_perturbation_conserves_momentum():
    # tests/test_deep_gravity_spectral.py:253-264
    ng, ratio, cell = 4, 2, 0.5
    ms = (8, 8, 8)
    q = _lattice(ng, ratio)
    masks = spectral.particle_nyquist_masks(ms, (ng,) * 3, 1.0, cell)
    acc = spectral.gravity(q, np.zeros(q.shape), 0.3, ms, cell, ng ** 3, masks)
    np.testing.assert_allclose(acc, 0, atol=1e-15)
    disp = np.zeros(q.shape)
    disp[0, 0] = 1e-3
    acc = spectral.gravity(q, disp, 0.3, ms, cell, ng ** 3, masks)
    assert np.linalg.norm(acc) > 1e-7
    assert np.linalg.norm(acc.sum(0)) < 1e-12


def test_force_is_linear_in_omega_m_and_vjp_matches_finite_differences():
    rng = np.random.default_rng(5)
    ng, cell, ms = 4, 0.5, (8, 8, 8)
    q = _lattice(ng, 2)
    disp = rng.uniform(-0.4, 0.4, q.shape)
    masks = spectral.particle_nyquist_masks(ms, (ng,) * 3, 1.0, cell)
    a1 = spectral.gravity(q, disp, 0.3, ms, cell, ng ** 3, masks)
    a2 = spectral.gravity(q, disp, 0.6, ms, cell, ng ** 3, masks)
    np.testing.assert_allclose(a2, 2 * a1, rtol=1e-12, atol=1e-16)
    ac = rng.normal(size=a1.shape)
    dc, oc = spectral.gravity_vjp(q, disp, 0.3, ac, ms, cell, ng ** 3, masks)
    assert oc == pytest.approx(np.sum(a1 * ac) / 0.3, rel=1e-10)
    eps = 1e-6
    for idx in [(0, 0), (5, 1), (17, 2), (40, 0)]:
        dp = disp.copy(); dp[idx] += eps
        dm = disp.copy(); dm[idx] -= eps
        fd = np.sum((spectral.gravity(q, dp, 0.3, ms, cell, ng ** 3, masks)
                     - spectral.gravity(q, dm, 0.3, ms, cell, ng ** 3, masks)) * ac) / (2 * eps)
        assert fd == pytest.approx(dc[idx], rel=1e-6, abs=1e-10)


def test_density_residual_plane_wave_and_its_cotangent():
    # particles on a uniform lattice exert no force, so the force of an added density
    # rho_add = A cos(k0 x) / (1.5 Omega_m) is the plane-wave closed form at the particle sites
    ng, ratio, cell, om, A = 4, 2, 0.5, 0.3, 0.01
    ms = (8, 8, 8)
    q = _lattice(ng, ratio)
    x = np.arange(8) * cell
    k0 = 2 * np.pi / (8 * cell)
    rho_add = (A * np.cos(k0 * x) / (1.5 * om))[:, None, None] * np.ones(ms)
    acc = spectral.gravity(q, np.zeros(q.shape), om, ms, cell, ng ** 3, rho_add=rho_add)
    xp = q[:, 0] * cell  # pmid counts mesh cells here (mesh_shape = ratio)
    np.testing.assert_allclose(acc[:, 0], -A * np.sin(k0 * xp) / k0, atol=1e-15)
    np.testing.assert_allclose(acc[:, 1:], 0, atol=1e-15)
    # the force is affine in rho_add: <acc_cot, J v> == <rho_cot, v> for the density cotangent
    rng = np.random.default_rng(9)
    disp = rng.uniform(-0.4, 0.4, q.shape)
    masks = spectral.particle_nyquist_masks(ms, (ng,) * 3, 1.0, cell)
    base = rng.normal(scale=0.1, size=ms)
    v = rng.normal(size=ms)
    ac = rng.normal(size=q.shape)
    f0 = spectral.gravity(q, disp, om, ms, cell, ng ** 3, masks, rho_add=base)
    f1 = spectral.gravity(q, disp, om, ms, cell, ng ** 3, masks, rho_add=base + v)
    dc, oc, rc = spectral.gravity_vjp(q, disp, om, ac, ms, cell, ng ** 3, masks, rho_add=base)
    assert np.sum(ac * (f1 - f0)) == pytest.approx(np.sum(rc * v), rel=1e-10)
    eps = 1e-6
    for idx in [(3, 0), (21, 2)]:  # the gather derivative now sees the residual's force too
        dp = disp.copy(); dp[idx] += eps
        dm = disp.copy(); dm[idx] -= eps
        fd = np.sum((spectral.gravity(q, dp, om, ms, cell, ng ** 3, masks, rho_add=base)
                     - spectral.gravity(q, dm, om, ms, cell, ng ** 3, masks, rho_add=base)) * ac) / (2 * eps)
        assert fd == pytest.approx(dc[idx], rel=1e-6, abs=1e-10)
    assert oc == pytest.approx(np.sum(f0 * ac) / om, rel=1e-10)
