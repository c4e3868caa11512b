"""Pins the P(k) oracle (oracle/power_spectrum.py)."""

import numpy as np
import pytest

from oracle import power_spectrum as ps


def test_single_mode_power_lands_in_its_shell():
    n, cell = 16, 2.0
    x = np.arange(n)
    A = 0.1
    delta = A * np.cos(2 * np.pi * 3 * x / n)[:, None, None] * np.ones((n, n, n))
    k, pk, nm = ps.delta_to_pk(delta, cell, mas=None)
    # |delta_k|^2 = (A N^3 / 2)^2 at the two modes (+-3,0,0); both have multiplicity 2 in the
    # half spectrum? no: kz = 0 plane -> multiplicity 1, so shell 3 sums 2 * (A N^3/2)^2.
    expected_sum = 2 * (A * n ** 3 / 2) ** 2
    assert pk[2] == pytest.approx(expected_sum / nm[2] * cell ** 3 / n ** 3, rel=1e-12)
    others = np.delete(pk, 2)
    assert np.all(np.abs(others) < 1e-20)


def test_parseval_without_window():
    rng = np.random.default_rng(0)
    n, cell = 8, 1.0
    delta = rng.normal(size=(n, n, n))
    k, pk, nm = ps.delta_to_pk(delta, cell, mas=None)
    dk = np.fft.rfftn(delta)
    mult = np.full(dk.shape, 2.0)
    mult[..., 0] = 1
    mult[..., -1] = 1
    total = np.sum(np.abs(dk) ** 2 * mult) - np.abs(dk[0, 0, 0]) ** 2
    # shells beyond the largest listed one are clipped into it, so the listed sum is complete
    assert np.sum(pk * nm) / (cell ** 3 / n ** 3) == pytest.approx(total, rel=1e-10)
    assert total == pytest.approx(n ** 3 * np.sum(delta ** 2) - np.abs(dk[0, 0, 0]) ** 2, rel=1e-10)


def _single_particle_average(n, cell, mas):
    """P(k) averaged over CIC-assigned single particles at the 2-point Gauss-Legendre offsets
    of a cell per axis (exact for the per-axis quadratic |DFT|^2 of a CIC particle)."""
    from oracle import cic

    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    pks = []
    for fx in g:
        for fy in g:
            for fz in g:
                mesh = cic.scatter(np.array([[3, 5, 7]]), np.array([[fx, fy, fz]]) * cell, 1.0, np.zeros((n,) * 3),
                                   cell, (n,) * 3)
                k, pk, nm = ps.delta_to_pk(mesh, cell, mas=mas)
                pks.append(pk)
    return np.mean(pks, axis=0), nm


def _expected_shell_mean(n, per_mode):
    """Brute-force shell means (multiplicity-weighted) of a per-mode value over the rfft half
    spectrum: shells floor(|m| + 1e-6) from 1 upward."""
    m = np.fft.fftfreq(n) * n
    mz = np.fft.rfftfreq(n) * n
    sums, wts = {}, {}
    for i, a in enumerate(m):
        for j, b in enumerate(m):
            for l, c in enumerate(mz):
                s = int(np.floor(np.sqrt(a * a + b * b + c * c) + 1e-6))
                w = 1.0 if (l == 0 or l == len(mz) - 1) else 2.0
                sums[s] = sums.get(s, 0.0) + w * per_mode(a, b, c)
                wts[s] = wts.get(s, 0.0) + w
    return np.array([sums[s] / wts[s] for s in sorted(sums) if s >= 1])


@pytest.mark.parametrize("mas", [None, "CIC"])
def test_cic_window_of_a_randomly_placed_particle(mas):
    # Averaged over its position in a cell, one CIC-assigned particle has per-mode power
    # prod_i [1 - 2/3 sin^2(pi m_i / N)] (the CIC alias sum; Jing 2005, ApJ 620, 559, eq. 20).
    # The raw spectrum must be that product, and deconvolving the CIC window (mas="CIC") must
    # divide it by prod_i sinc^4(m_i / N).  A wrong window exponent or sinc convention fails.
    n, cell = 16, 1.5
    pk, nm = _single_particle_average(n, cell, mas)

    def alias(a, b, c):
        return np.prod([1 - 2.0 / 3.0 * np.sin(np.pi * x / n) ** 2 for x in (a, b, c)])

    if mas is None:
        per_mode = alias
    else:
        def per_mode(a, b, c):
            return alias(a, b, c) / np.prod([np.sinc(x / n) ** 4 for x in (a, b, c)])

    exp = _expected_shell_mean(n, per_mode)[:len(pk)] * cell ** 3 / n ** 3
    np.testing.assert_allclose(pk, exp, rtol=1e-12)
    if mas == "CIC":
        # the deconvolved shot noise rises above the raw one at high k (the window is < 1)
        raw, _ = _single_particle_average(n, cell, None)
        assert np.all(pk[1:] > raw[1:])
