"""Pins the CIC oracle (oracle/cic.py) independently of its own formulas."""

import numpy as np
import pytest

from oracle import cic


def _trilinear_bruteforce(pos, mesh_shape, cell):
    """Direct geometric CIC: every mesh node within one cell gets prod(1 - |x - x_node|/cell)."""
    mesh = np.zeros(mesh_shape)
    L = np.asarray(mesh_shape) * cell
    for p in pos:
        for ix in range(mesh_shape[0]):
            for iy in range(mesh_shape[1]):
                for iz in range(mesh_shape[2]):
                    node = np.array([ix, iy, iz]) * cell
                    d = p - node
                    d = d - np.round(d / L) * L  # periodic minimum image
                    w = np.prod(np.clip(1 - np.abs(d) / cell, 0, None))
                    mesh[ix, iy, iz] += w
    return mesh


def test_enmesh_partition_of_unity_and_zero_gradient_sum():
    # tests/test_deep_scatter_fallback.py:75-95
    rng = np.random.default_rng(0)
    pmid = rng.integers(-3, 9, (200, 3))
    disp = rng.uniform(-2.5, 2.5, (200, 3))
    for mesh_cell in (None, 0.25):  # L = 8 * 0.5 = 4 is a multiple of a2, as in the reference test
        _, frac, grad = cic.enmesh(pmid, disp, 0.5, (8, 8, 8), offset=[0.1, -0.2, 0.05],
                                   mesh_cell_size=mesh_cell, grad=True)
        np.testing.assert_allclose(frac.sum(1), 1, atol=2e-15)
        np.testing.assert_allclose(grad.sum(1), 0, atol=2e-15)


def test_scatter_matches_geometric_bruteforce():
    rng = np.random.default_rng(1)
    ms, cell = (5, 4, 6), 0.7
    pmid = rng.integers(0, 4, (7, 3))
    disp = rng.uniform(-1.3, 1.3, (7, 3))
    pos = pmid * cell + disp
    got = cic.scatter(pmid, disp, 1.0, np.zeros(ms), cell, ms)
    np.testing.assert_allclose(got, _trilinear_bruteforce(pos, ms, cell), atol=1e-13)


def test_uniform_lattice_density_is_exactly_one():
    # tests/test_deep_scatter_fallback.py:57-62
    n = 4
    q = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    d = cic.scatter(q, np.zeros(q.shape), 1.0, np.zeros((n, n, n)), 1.0, (n, n, n))
    np.testing.assert_array_equal(d, 1.0)
    assert d.sum() == n ** 3


def test_adjoint_dot_product_identities():
    rng = np.random.default_rng(2)
    ms = (6, 5, 4, 2)
    pmid = rng.integers(0, 6, (40, 3))
    disp = rng.uniform(-1, 1, (40, 3))
    v = rng.normal(size=(40, 2))
    M = rng.normal(size=ms)
    lhs = np.sum(cic.scatter(pmid, disp, v, np.zeros(ms), 0.5, ms[:3]) * M)
    rhs = np.sum(v * cic.gather(pmid, disp, M, 0.5, ms[:3]))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # gather adjoint mesh part is the scatter
    vc = rng.normal(size=(40, 2))
    _, mc = cic.gather_adj(pmid, disp, M, vc, 0.5, ms[:3])
    np.testing.assert_allclose(mc, cic.scatter(pmid, disp, vc, np.zeros(ms), 0.5, ms[:3]), atol=1e-12)
    # scatter adjoint value part is the gather
    _, vcot = cic.scatter_adj(pmid, disp, v, M, 0.5, ms[:3])
    np.testing.assert_allclose(vcot, cic.gather(pmid, disp, M, 0.5, ms[:3]), atol=1e-12)


def _fd4(f, x, idx, eps):
    xs = []
    for k in (-2, -1, 1, 2):
        y = x.copy()
        y[idx] += k * eps
        xs.append(f(y))
    return (xs[0] - 8 * xs[1] + 8 * xs[2] - xs[3]) / (12 * eps)


@pytest.mark.parametrize("mesh_cell", [None, 0.375])  # L = 3 = 8 * 0.375
def test_adjoint_displacement_gradients_match_finite_differences(mesh_cell):
    # tests/test_finite_diff.py:206-255 (4th order central FD, float64)
    rng = np.random.default_rng(3)
    ms = (6, 6, 6)
    pmid = rng.integers(0, 6, (12, 3))
    disp = rng.uniform(-0.9, 0.9, (12, 3))
    v = rng.normal(size=12)
    W = rng.normal(size=ms)
    off = [0.07, -0.11, 0.02]
    ms_mesh = ms if mesh_cell is None else (8, 8, 8)
    W = rng.normal(size=ms_mesh)

    def loss_s(dd):
        return np.sum(cic.scatter(pmid, dd, v, np.zeros(ms_mesh), 0.5, ms, off, mesh_cell) * W)

    def loss_g(dd):
        return np.sum(cic.gather(pmid, dd, W, 0.5, ms, off, mesh_cell) * v)

    ds, _ = cic.scatter_adj(pmid, disp, v, W, 0.5, ms, off, mesh_cell)
    dg, _ = cic.gather_adj(pmid, disp, W, v, 0.5, ms, off, mesh_cell)
    for idx in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        assert _fd4(loss_s, disp, idx, 1e-6) == pytest.approx(ds[idx], rel=1e-7, abs=1e-9)
        assert _fd4(loss_g, disp, idx, 1e-6) == pytest.approx(dg[idx], rel=1e-7, abs=1e-9)


def test_invalid_particles_and_zero_values_give_zero_gradients():
    # tests/test_deep_scatter_fallback.py:42-44, 64-67
    rng = np.random.default_rng(4)
    pmid = rng.integers(0, 4, (6, 3))
    disp = rng.uniform(-0.4, 0.4, (6, 3))
    v = rng.normal(size=6)
    v[2] = 0
    valid = np.array([1, 1, 1, 0, 1, 1], bool)
    W = rng.normal(size=(4, 4, 4))
    dc, vc = cic.scatter_adj(pmid, disp, v, W, 1.0, (4, 4, 4), valid=valid)
    np.testing.assert_array_equal(dc[2], 0)
    np.testing.assert_array_equal(dc[3], 0)
    assert vc[3] == 0
    g = cic.gather(pmid, disp, W, 1.0, (4, 4, 4), valid=valid)
    assert g[3] == 0
