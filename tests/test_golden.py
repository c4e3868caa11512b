"""The float64 oracle against the brute-force golden fixtures of ``tests/golden/``.

The fixtures come from ``tests/golden/make_golden.py`` (per-corner loops, O(N^2) DFT sums,
Python's stable ``sorted``), an implementation independent of the oracle's vectorised one.
"""

import os

import numpy as np
import pytest

from oracle import cic, sort, spectral

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))


def test_cic_forward_matches_golden():
    g = load("cic")
    pm, dp = g["pmid"], g["disp"]
    okw = dict(cell_size=1.0, mesh_shape=(4, 4, 4))
    np.testing.assert_allclose(cic.gather(pm, dp, g["mesh"], **okw), g["gathered"], rtol=1e-12, atol=1e-12)
    s = cic.scatter(pm, dp, g["val"], np.zeros((4, 4, 4)), **okw)
    np.testing.assert_allclose(s, g["scattered"], rtol=1e-12, atol=1e-12)
    assert np.isclose(s.sum(), g["val"].astype(np.float64).sum())  # mass conservation


def test_cic_explicit_cell_matches_golden():
    g = load("cic")
    pm, dp = g["pmid"], g["disp"]
    okw = dict(cell_size=1.0, mesh_shape=(4, 4, 4), offset=g["offset8"], mesh_cell_size=0.5)
    np.testing.assert_allclose(cic.gather(pm, dp, g["mesh8"], **okw), g["gathered8"], rtol=1e-12, atol=1e-12)
    s = cic.scatter(pm, dp, g["val"], np.zeros((8, 8, 8)), **okw)
    np.testing.assert_allclose(s, g["scattered8"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("with_valid", [True, False])
def test_sort_matches_golden(with_valid):
    g = load("sort")
    ms, cell, brick = tuple(g["mesh_shape"]), float(g["cell"]), tuple(g["brick"])
    valid = g["valid"] if with_valid else None
    keys, sentinel = sort.cell_keys(g["pmid"], g["disp"], cell, ms, brick)
    np.testing.assert_array_equal(keys, g["keys"])
    assert sentinel == len(g["bins"]) - 2
    perm, offs = sort.sort_cells(g["pmid"], g["disp"], cell, ms, brick, valid)
    np.testing.assert_array_equal(perm, g["perm"] if with_valid else g["perm_all"])
    if with_valid:
        np.testing.assert_array_equal(offs[: len(g["bins"])], g["bins"])


def test_fft_matches_golden():
    g = load("fft")
    np.testing.assert_allclose(np.fft.rfftn(g["x"].astype(np.float64)), g["X"], rtol=1e-10, atol=1e-10)


def test_gravity_matches_golden():
    g = load("gravity")
    ng, nm, sp = int(g["ptcl_grid"]), int(g["mesh"]), float(g["ptcl_spacing"])
    cell = sp * ng / nm
    ms, pg = (nm,) * 3, (ng,) * 3
    masks = spectral.particle_nyquist_masks(ms, pg, sp, cell)
    assert len(masks) == 3  # the filter is active at mesh_shape=2
    acc = spectral.gravity(g["pmid"], g["disp"], float(g["omega_m"]), ms, cell, ng ** 3, masks)
    np.testing.assert_allclose(acc, g["acc"], rtol=1e-9, atol=1e-9 * np.abs(g["acc"]).max())
