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
