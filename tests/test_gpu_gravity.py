"""HIP PM force step vs the float64 oracle, closed forms and the reference's own gates."""

import numpy as np
import pytest
import torch

from oracle import spectral
from pmpp_amd import Configuration, Particles, gravity

pytestmark = pytest.mark.gpu


def _ptcl(conf, disp_scale, seed, device, valid_frac=None):
    base = Particles.gen_grid(conf, device=device)
    rng = np.random.default_rng(seed)
    d = base.disp.cpu().numpy() + (rng.normal(size=base.disp.shape) * disp_scale * conf.ptcl_spacing)
    unused = None
    if valid_frac is not None:
        unused = torch.as_tensor(rng.random(len(base)) > valid_frac, device=device)
    return base.replace(disp=torch.as_tensor(d.astype(np.float32), device=device),
                        unused_index=unused if unused is not None else base.unused_index)


def _oracle_acc(conf, ptcl, omega):
    pm = ptcl.pmid.cpu().numpy().astype(np.int64)
    dp = ptcl.disp.cpu().numpy().astype(np.float64)
    valid = None if ptcl.unused_index is None else ~ptcl.unused_index.cpu().numpy()
    masks = spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                            conf.cell_size)
    return spectral.gravity(pm, dp, omega, conf.mesh_shape, conf.cell_size, conf.ptcl_num, masks, valid), masks


@pytest.mark.parametrize("ng,ratio,box,scale,valid_frac", [
    (8, 1, 8.0, 0.3, None),
    (8, 2, 16.0, 0.3, None),
    (16, 2, 64.0, 1.5, None),
    (8, 4, 8.0, 0.2, None),
    (16, 2, 32.0, 0.5, 0.8),
    (32, 2, 128.0, 2.0, None),
])
def test_force_matches_float64_oracle(device, ng, ratio, box, scale, valid_frac):
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ptcl = _ptcl(conf, scale, ng * ratio, device, valid_frac)
    acc = gravity(1.0, ptcl, 0.3, conf).cpu().numpy()
    ref, _ = _oracle_acc(conf, ptcl, 0.3)
    # float32 path vs float64: relative error of the field, stated tolerance 1e-5
    err = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err
    np.testing.assert_allclose(acc, ref, atol=3e-5 * np.abs(ref).max())
    if valid_frac is not None:
        assert np.all(acc[ptcl.unused_index.cpu().numpy()] == 0)


@pytest.mark.parametrize("brick", [(4, 2, 8), (1, 1, 1), (16, 16, 16), (32, 1, 2), (16, 16, 32), (3, 3, 3),
                                   (32, 32, 32)])
def test_force_is_independent_of_the_brick_tiling(device, brick):
    # (3,3,3) does not divide the mesh and (32,32,32) overflows the LDS tile: both take the
    # untiled sorted path; the others exercise periodic brick neighbours including nb = 1,
    # and (16,16,32) the compile-time-shape kernels
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2, brick=brick)
    ptcl = _ptcl(conf, 1.0, 3, device, valid_frac=0.9)
    acc = gravity(1.0, ptcl, 0.3, conf).cpu().numpy()
    ref, _ = _oracle_acc(conf, ptcl, 0.3)
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5
    assert np.all(acc[ptcl.unused_index.cpu().numpy()] == 0)


@pytest.mark.parametrize("grid,ratio", [((8, 16, 4), 2), ((16, 4, 8), 4), ((4, 8, 1), 2), ((2, 4, 8), 2)])
def test_banded_spectrum_on_non_cubic_meshes(device, grid, ratio):
    # the particle-Nyquist band differs per axis: packed y rows and kz columns of different
    # widths, a z axis of 2 points (band = DC only) and an x axis of 4 points
    conf = Configuration(1.5, grid, mesh_shape=ratio)
    ptcl = _ptcl(conf, 0.6, sum(grid), device)
    acc = gravity(1.0, ptcl, 0.3, conf).cpu().numpy()
    ref, masks = _oracle_acc(conf, ptcl, 0.3)
    assert len(masks) == 3
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("brick", [(8, 8, 8), (16, 16, 32)])
def test_clustered_particles_exceed_one_gather_chunk(device, brick):
    # 90 % of the particles in one brick: several gather chunks of 2048 particles (the last
    # one partial) and heavy LDS-atomic contention; the rest leave most bricks and some brick
    # columns empty.  (8, 8, 8) runs the runtime-shape tile kernels, (16, 16, 32) the
    # compile-time ones.
    conf = Configuration(1.0, (32, 32, 32), mesh_shape=1, brick=brick)
    rng = np.random.default_rng(9)
    n = 32 ** 3
    pm = rng.integers(0, 8, (n, 3))
    far = rng.random(n) < 0.1
    pm[far] = rng.integers(0, 32, (int(far.sum()), 3))
    pmid = torch.as_tensor(pm.astype(np.int16), device=device)
    disp = torch.as_tensor(rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32), device=device)
    ptcl = Particles(conf, pmid, disp)
    acc = gravity(1.0, ptcl, 0.3, conf).cpu().numpy()
    ref, _ = _oracle_acc(conf, ptcl, 0.3)
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5


def test_uniform_lattice_is_force_free(device):
    # tests/test_deep_gravity_spectral.py:253-264
    conf = Configuration(1.25, (4, 4, 4), mesh_shape=1)
    grid = Particles.gen_grid(conf, device=device)
    acc = gravity(0.8, grid, 0.3, conf)
    assert float(acc.abs().max()) <= 3e-7
    disp = grid.disp.clone()
    disp[0, 0] = 1e-3
    pert = gravity(0.8, grid.replace(disp=disp), 0.3, conf).cpu().numpy()
    assert np.linalg.norm(pert) > 1e-7
    assert np.linalg.norm(pert.sum(axis=0)) < 2e-5


def test_refined_force_mesh_does_not_alias_particle_lattice(device):
    # tests/test_gravity_particle_nyquist_filter.py:11-40 restated
    res, box, amp = 16, 1000.0, 0.5

    def modes(mesh_shape):
        conf = Configuration(box / res, (res,) * 3, mesh_shape=mesh_shape)
        axis = np.arange(res, dtype=np.float32) * np.float32(box / res)
        x, y, z = np.meshgrid(axis, axis, axis, indexing="ij")
        pos = np.stack([x, y, z], -1).reshape(-1, 3)
        pos[:, 0] = (pos[:, 0] + np.float32(amp) * np.sin(np.float32(2 * np.pi / box) * pos[:, 0])) % np.float32(box)
        ptcl = Particles.from_pos(conf, torch.as_tensor(pos, device=device))
        acc = gravity(1.0, ptcl, 0.3, conf).cpu().numpy()
        mean_x = acc[:, 0].reshape(res, res, res).mean(axis=(1, 2))
        return np.abs(np.fft.rfft(mean_x)) / (res / 2)

    base = modes(1)
    for ms in (2, 4):
        m = modes(ms)
        assert np.isclose(m[1], base[1], rtol=0.03)
        assert m[2] < 0.10 * base[1]


@pytest.mark.parametrize("ng,ratio", [(8, 2), (16, 1)])
def test_force_vjp_matches_oracle(device, ng, ratio):
    conf = Configuration(2.0, (ng,) * 3, mesh_shape=ratio)
    ptcl = _ptcl(conf, 0.4, 7, device)
    disp = ptcl.disp.clone().requires_grad_(True)
    omega = torch.tensor(0.3, device=device, requires_grad=True)
    acc = gravity(1.0, ptcl.replace(disp=disp), type("C", (), {"Omega_m": omega})(), conf)
    rng = np.random.default_rng(11)
    ac = rng.normal(size=acc.shape).astype(np.float32)
    (acc * torch.as_tensor(ac, device=device)).sum().backward()
    masks = spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                            conf.cell_size)
    dref, oref = spectral.gravity_vjp(ptcl.pmid.cpu().numpy().astype(np.int64),
                                      ptcl.disp.cpu().numpy().astype(np.float64), 0.3, ac, conf.mesh_shape,
                                      conf.cell_size, conf.ptcl_num, masks)
    got = disp.grad.cpu().numpy()
    # gradients within 1e-3 (relative, field norm)
    assert np.linalg.norm(got - dref) / np.linalg.norm(dref) < 1e-3
    assert float(omega.grad) == pytest.approx(oref, rel=1e-3)


@pytest.mark.parametrize("brick,clustered,valid_frac", [
    ((8, 8, 8), False, 0.8),    # runtime-shape tiled reverse, invalid particles zeroed
    ((16, 16, 32), False, None),  # compile-time tiled reverse
    ((16, 16, 32), True, 0.9),  # one brick holds several 2048-particle chunks
    ((3, 3, 3), False, 0.8),    # brick does not divide the mesh: untiled reverse
])
def test_force_vjp_across_brick_tilings(device, brick, clustered, valid_frac):
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2, brick=brick)
    rng = np.random.default_rng(23)
    n = 16 ** 3
    if clustered:
        pm = rng.integers(0, 12, (n, 3))
        far = rng.random(n) < 0.1
        pm[far] = rng.integers(0, 32, (int(far.sum()), 3))
        base = Particles(conf, torch.as_tensor(pm.astype(np.int16), device=device),
                         torch.as_tensor(rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32), device=device))
    else:
        base = _ptcl(conf, 0.6, 29, device)
    unused = None if valid_frac is None else torch.as_tensor(rng.random(n) > valid_frac, device=device)
    ptcl = base.replace(unused_index=unused)
    disp = ptcl.disp.clone().requires_grad_(True)
    omega = torch.tensor(0.3, device=device, requires_grad=True)
    acc = gravity(1.0, ptcl.replace(disp=disp), type("C", (), {"Omega_m": omega})(), conf)
    ac = rng.normal(size=acc.shape).astype(np.float32)
    (acc * torch.as_tensor(ac, device=device)).sum().backward()
    masks = spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                            conf.cell_size)
    valid = None if unused is None else ~unused.cpu().numpy()
    dref, oref = spectral.gravity_vjp(ptcl.pmid.cpu().numpy().astype(np.int64),
                                      ptcl.disp.cpu().numpy().astype(np.float64), 0.3, ac, conf.mesh_shape,
                                      conf.cell_size, conf.ptcl_num, masks, valid=valid)
    got = disp.grad.cpu().numpy()
    assert np.linalg.norm(got - dref) / np.linalg.norm(dref) < 1e-3
    assert float(omega.grad) == pytest.approx(oref, rel=1e-3)
    if valid is not None:
        assert np.all(got[~valid] == 0)


@pytest.mark.parametrize("brick,clustered,valid_frac", [
    ((8, 8, 8), False, 0.8),
    ((16, 16, 32), False, None),
    ((16, 16, 32), True, 0.9),
    ((3, 3, 3), False, 0.8),
])
def test_density_cotangent_pass_matches_oracle_scatter_adjoint(device, brick, clustered, valid_frac):
    """pm_gravity_density_cot (the particles binned by brick, the tiled density term of the
    VJP's displacement pass, added to disp_cot) against the float64 oracle's CIC scatter adjoint
    with the step's particle value; invalid particles' rows keep what disp_cot held."""
    from oracle import cic
    from pmpp_amd.corrections.training import density_cot_add

    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2, brick=brick)
    rng = np.random.default_rng(37)
    n = 16 ** 3
    if clustered:
        pm = rng.integers(0, 12, (n, 3))
        base = Particles(conf, torch.as_tensor(pm.astype(np.int16), device=device),
                         torch.as_tensor(rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32), device=device))
    else:
        base = _ptcl(conf, 0.6, 41, device)
    unused = None if valid_frac is None else torch.as_tensor(rng.random(n) > valid_frac, device=device)
    ptcl = base.replace(unused_index=unused)
    dens_cot = rng.normal(size=conf.mesh_shape).astype(np.float32)
    start = rng.normal(size=(n, 3)).astype(np.float32)
    got = density_cot_add(ptcl, 0.3, conf, torch.from_numpy(dens_cot).to(device),
                          torch.from_numpy(start).to(device)).cpu().numpy().astype(np.float64)
    valid = None if unused is None else ~unused.cpu().numpy()
    pm64, dp64 = ptcl.pmid.cpu().numpy().astype(np.int64), ptcl.disp.cpu().numpy().astype(np.float64)
    val = spectral.default_value(n, conf.mesh_shape, conf.ptcl_num, valid)
    dref, _ = cic.scatter_adj(pm64, dp64, val, dens_cot.astype(np.float64), conf.cell_size, conf.mesh_shape)
    ref = start.astype(np.float64) + dref
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(dref).max())
    if valid is not None:
        np.testing.assert_array_equal(got[~valid], start[~valid])


def test_force_and_vjp_on_1024_point_z_lines(device):
    # an 8 x 8 x 1024 mesh: the banded z passes of 1024-point lines (r2c, the derivative c2r,
    # and the adjoint's plain c2r) against the oracle
    conf = Configuration(1.0, (4, 4, 512), mesh_shape=2)
    ptcl = _ptcl(conf, 0.5, 31, device)
    ref, masks = _oracle_acc(conf, ptcl, 0.3)
    disp = ptcl.disp.clone().requires_grad_(True)
    omega = torch.tensor(0.3, device=device, requires_grad=True)
    acc = gravity(1.0, ptcl.replace(disp=disp), type("C", (), {"Omega_m": omega})(), conf)
    got = acc.detach().cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    ac = np.random.default_rng(5).normal(size=acc.shape).astype(np.float32)
    (acc * torch.as_tensor(ac, device=device)).sum().backward()
    dref, oref = spectral.gravity_vjp(ptcl.pmid.cpu().numpy().astype(np.int64),
                                      ptcl.disp.cpu().numpy().astype(np.float64), 0.3, ac, conf.mesh_shape,
                                      conf.cell_size, conf.ptcl_num, masks)
    g = disp.grad.cpu().numpy()
    assert np.linalg.norm(g - dref) / np.linalg.norm(dref) < 1e-3
    assert float(omega.grad) == pytest.approx(oref, rel=1e-3)


@pytest.mark.parametrize("brick,valid_frac", [((16, 16, 32), 0.85), ((3, 3, 3), None)])
def test_kept_force_and_resumed_vjp_equal_the_plain_calls(device, brick, valid_frac):
    # pm_gravity_keep / pm_gravity_vjp_resume (the reverse sweep's pair at one set of positions)
    # against pm_gravity / pm_gravity_vjp: bit for bit with brick tiling (fixed-point scatters,
    # per-particle gathers); the untiled path scatters with float atomics, whose order (and so
    # the last bits) differs between any two calls
    from pmpp_amd.nbody.gravity import KeptForce, gravity_keep, gravity_vjp

    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2, brick=brick)
    ptcl = _ptcl(conf, 0.7, 37, device, valid_frac)
    cosmo = type("C", (), {"Omega_m": 0.3})()
    ac = torch.as_tensor(np.random.default_rng(41).normal(size=(len(ptcl), 3)).astype(np.float32), device=device)
    kept = KeptForce()
    acc_keep = gravity_keep(ptcl, cosmo, conf, kept)
    acc = gravity(1.0, ptcl, 0.3, conf)
    d_res, o_res = gravity_vjp(ptcl, cosmo, conf, ac, kept=kept)
    assert kept.key is None  # used once
    d_ref, o_ref = gravity_vjp(ptcl, cosmo, conf, ac)
    if brick == (16, 16, 32):
        assert torch.equal(acc_keep, acc)
        assert torch.equal(d_res, d_ref)
        assert o_res == o_ref
    else:
        assert float((acc_keep - acc).norm() / acc.norm()) < 1e-6
        assert float((d_res - d_ref).norm() / d_ref.norm()) < 1e-6
        assert o_res == pytest.approx(o_ref, rel=1e-6)


def test_particle_subsets_empty_sets_and_all_invalid(device):
    # a particle count that is not a whole lattice (odd, bucketed in plain index chunks), the
    # empty set, and a set whose particles are all invalid
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2)
    full = _ptcl(conf, 0.8, 17, device)
    idx = torch.from_numpy(np.random.default_rng(2).permutation(len(full))[:3001]).to(device)
    sub = full.replace(pmid=full.pmid[idx], disp=full.disp[idx],
                       unused_index=None if full.unused_index is None else full.unused_index[idx])
    acc = gravity(1.0, sub, 0.3, conf).cpu().numpy()
    ref, _ = _oracle_acc(conf, sub, 0.3)
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5
    empty = full.replace(pmid=full.pmid[:0], disp=full.disp[:0],
                         unused_index=None if full.unused_index is None else full.unused_index[:0])
    assert tuple(gravity(1.0, empty, 0.3, conf).shape) == (0, 3)
    dead = full.replace(unused_index=torch.ones(len(full), dtype=torch.bool, device=device))
    assert torch.all(gravity(1.0, dead, 0.3, conf) == 0)
