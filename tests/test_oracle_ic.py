"""Pins the initial-condition oracle (oracle/ic.py) and the host power spectrum
(pmpp_amd/cosmology/boltzmann.py) by closed forms and published limits."""

import math

import numpy as np
import pytest

from oracle import ic


def _plane_setup(n=16, spacing=2.0):
    x = np.arange(n) * spacing
    L = n * spacing
    return x, L


def test_zeldovich_displacement_of_a_plane_wave():
    # delta = A cos(k0 x): phi_1 = -A cos(k0 x)/k0^2, Psi_1 = -grad phi_1 = -A sin(k0 x)/k0 x^
    n, sp = 16, 2.0
    x, L = _plane_setup(n, sp)
    k0, A = 2 * np.pi * 3 / L, 0.05
    delta = A * np.cos(k0 * x)[:, None, None] * np.ones((n, n, n))
    modes = np.fft.rfftn(delta) * sp ** 3  # the [L^3] normalisation
    D1, V1 = 0.37, 0.81
    disp, vel = ic.lpt(modes, (n, n, n), sp, [(D1, V1), (0.0, 0.0)], order=1)
    psi = (-A * np.sin(k0 * x) / k0)[:, None, None] * np.ones((n, n, n))
    np.testing.assert_allclose(disp[:, 0], D1 * psi.reshape(-1), atol=1e-15)
    np.testing.assert_allclose(vel[:, 0], V1 * psi.reshape(-1), atol=1e-15)
    np.testing.assert_allclose(disp[:, 1:], 0, atol=1e-15)
    # one plane wave has no second-order source
    ks = ic.kvec_spacing((n, n, n), sp)
    pot = ic.laplace(ks, modes / sp ** 3)
    assert np.abs(ic.second_order_source(ks, pot, (n, n, n), sp)).max() < 1e-18


def test_second_order_displacement_of_two_orthogonal_waves():
    # delta = A cos(kx x) + B cos(ky y): phi_xx = A cos(kx x), phi_yy = B cos(ky y), so
    # L = A B cos(kx x) cos(ky y), phi_2 = -L / (kx^2 + ky^2) and
    # Psi_2 = -grad phi_2 = -A B (kx sin(kx x) cos(ky y), ky cos(kx x) sin(ky y), 0) / K^2
    n, sp = 16, 1.5
    x, L = _plane_setup(n, sp)
    kx, ky = 2 * np.pi * 1 / L, 2 * np.pi * 2 / L
    A, B = 0.03, -0.02
    X, Y = np.meshgrid(x, x, indexing="ij")
    delta = (A * np.cos(kx * X) + B * np.cos(ky * Y))[:, :, None] * np.ones((1, 1, n))
    modes = np.fft.rfftn(delta) * sp ** 3
    D2, V2 = 0.6, -1.3
    disp, vel = ic.lpt(modes, (n, n, n), sp, [(0.0, 0.0), (D2, V2)], order=2)
    K2 = kx * kx + ky * ky
    px = (-A * B * kx * np.sin(kx * X) * np.cos(ky * Y) / K2)[:, :, None] * np.ones((1, 1, n))
    py = (-A * B * ky * np.cos(kx * X) * np.sin(ky * Y) / K2)[:, :, None] * np.ones((1, 1, n))
    np.testing.assert_allclose(disp[:, 0], D2 * px.reshape(-1), atol=1e-15)
    np.testing.assert_allclose(disp[:, 1], D2 * py.reshape(-1), atol=1e-15)
    np.testing.assert_allclose(vel[:, 1], V2 * py.reshape(-1), atol=1e-15)
    np.testing.assert_allclose(disp[:, 2], 0, atol=1e-15)


def test_linear_modes_carry_the_linear_power_on_average():
    # <|delta_k|^2> = V P(k) for orthonormal white noise: the shell-averaged sample power of a
    # 64^3 realisation is P(k) within its Gaussian scatter (shells with > 2000 modes)
    from pmpp_amd.cosmology import LCDM, boltzmann

    n, sp = 64, 4.0
    cos = LCDM()
    tk, tT = boltzmann.transfer_table(cos)
    coef = boltzmann.power_coef(cos, cos.A_s_1e9)
    w = ic.ortho_rfftn(ic.white_noise(7, (n, n, n)))
    modes = ic.linear_modes(w, (n, n, n), sp, tk, tT, coef, boltzmann.k_pivot(cos), cos.n_s)
    ks = ic.kvec_spacing((n, n, n), sp)
    kx, ky, kz = ks[0][:, None, None], ks[1][None, :, None], ks[2][None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    vol = (n * sp) ** 3
    pk_sample = np.abs(modes) ** 2 / vol
    shell = np.floor(k / (2 * np.pi / (n * sp))).astype(int)
    checked = 0
    for s in range(8, 30):
        sel = (shell == s) & (kz > 0)
        m = sel.sum()
        if m < 2000:
            continue
        ratio = pk_sample[sel].mean() / ic.linear_power(k[sel], tk, tT, coef, boltzmann.k_pivot(cos), cos.n_s).mean()
        assert abs(ratio - 1) < 5 / math.sqrt(m), (s, ratio)
        checked += 1
    assert checked > 5
    assert modes[0, 0, 0] == 0


def test_transfer_fit_limits_and_closed_forms():
    # tests/test_deep_boltzmann_contracts.py:19-30
    from pmpp_amd.cosmology import LCDM, boltzmann

    cos = LCDM()
    k = np.concatenate(([0.0], np.geomspace(1e-4, 0.5, 64)))
    full = boltzmann.transfer_fit(k, cos)
    smooth = boltzmann.transfer_fit(k, cos, nowiggle=True)
    assert np.all(np.isfinite(smooth))
    assert smooth[0] == 1.0 and full[0] == pytest.approx(1.0, abs=1e-15)
    np.testing.assert_allclose(smooth[1], full[1], rtol=2e-3)
    assert np.all(np.diff(smooth) <= 1e-12)
    # the wiggles oscillate around the no-wiggle fit (EH98 sec. 4)
    kk = np.geomspace(0.02, 0.3, 400)
    r = boltzmann.transfer_fit(kk, cos) / boltzmann.transfer_fit(kk, cos, nowiggle=True)
    assert np.abs(r - 1).max() < 0.1 and np.sum(np.diff(np.sign(r - np.mean(r))) != 0) >= 4
    # no baryons: EH98 eq. 29 with alpha = beta = 1 in closed form
    nob = LCDM(Omega_b=0.0)
    h = nob.h
    w_m = nob.Omega_m * h * h
    k_eq = 7.46e-2 * w_m / (nob.T_cmb / 2.7) ** 2
    q = kk * h / (13.41 * k_eq)
    Lq = np.log(math.e + 1.8 * q)
    expected = Lq / (Lq + (14.2 + 386 / (1 + 69.9 * q ** 1.08)) * q * q)
    np.testing.assert_allclose(boltzmann.transfer_fit(kk, nob), expected, rtol=1e-14)


def test_sigma8_normalisation_round_trip_and_scalings():
    from pmpp_amd.cosmology import LCDM, boltzmann

    for cos in (LCDM(), LCDM(sigma8=0.7, Omega_m=0.28, h=0.68)):
        A = cos.A_s_1e9
        D1 = boltzmann._growth_today(cos)
        assert boltzmann.sigma_tophat(8.0, cos, A, D1) == pytest.approx(cos.sigma8, rel=1e-12)
        assert 1.0 < A < 3.0  # the familiar A_s ~ 2e-9
        k = np.geomspace(1e-3, 1.0, 7)
        p1 = boltzmann.linear_power(k, cos, A)
        np.testing.assert_allclose(boltzmann.linear_power(k, cos, 2 * A), 2 * p1, rtol=1e-14)
        np.testing.assert_allclose(boltzmann.linear_power(k, cos, A, D=0.5), 0.25 * p1, rtol=1e-14)
    # in the matter-dominated era the growth normalisation gives D(a) -> a
    assert boltzmann._growth_today(LCDM(Omega_m=1.0)) == pytest.approx(1.0, rel=1e-9)
