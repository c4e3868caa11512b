"""Host-side logic: configuration geometry and spectral caches vs the oracle."""

import numpy as np
import pytest

from oracle import spectral
from pmpp_amd import Configuration


def test_configuration_geometry_and_validation():
    conf = Configuration(1000 / 512, (512,) * 3, mesh_shape=2)
    assert conf.mesh_shape == (1024,) * 3
    assert conf.cell_size == pytest.approx(1000 / 1024)
    assert conf.mesh_size == 1024 ** 3 and conf.ptcl_num == 512 ** 3
    assert conf.brick == (16, 16, 32)
    with pytest.raises(ValueError, match="cannot be smaller"):
        Configuration(1.0, (8, 8, 8), mesh_shape=(4, 8, 8))
    with pytest.raises(ValueError, match="aspect ratios"):
        Configuration(1.0, (8, 8, 8), mesh_shape=(16, 8, 8))
    with pytest.raises(ValueError, match="symplectic"):
        Configuration(1.0, (8, 8, 8), symp_splits=((0, 0.5), (1, 0.6)))


@pytest.mark.parametrize("ratio", [1, 2, 4])
def test_spectral_caches_match_oracle(ratio):
    conf = Configuration(2.0, (8, 8, 8), mesh_shape=ratio)
    for a, k in enumerate(conf.kvec):
        np.testing.assert_allclose(k, spectral.kvec(conf.mesh_shape, conf.cell_size)[a].astype(np.float32))
    masks = spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                            conf.cell_size)
    assert len(conf.particle_nyquist_masks) == len(masks)
    for m, mo in zip(conf.particle_nyquist_masks, masks):
        np.testing.assert_array_equal(m, mo)
    for a, nik in enumerate(conf.neg_ik):
        np.testing.assert_allclose(nik.imag, spectral.neg_k(conf.mesh_shape, conf.cell_size)[a], rtol=1e-6)


def test_nbody_schedule():
    conf = Configuration(2.0, (8, 8, 8), a_start=1 / 64, a_stop=1.0, a_nbody_maxstep=(1 - 1 / 64) / 10)
    a = conf.a_nbody
    assert len(a) == 11 and a[0] == pytest.approx(1 / 64) and a[-1] == pytest.approx(1.0)


def test_step_factors_match_the_oracle():
    # host float64 growth table (Dormand-Prince) vs the oracle's (scipy DOP853), same nodes
    import numpy as np
    from oracle import cosmology as oc
    from pmpp_amd import Configuration
    from pmpp_amd.initial_conditions import LCDM
    from pmpp_amd.nbody import integrator

    conf = Configuration(2.0, (8, 8, 8), mesh_shape=2)
    for cos in (LCDM(), LCDM(Omega_m=0.25, Omega_k=0.05, w_0=-0.9, w_a=0.1)):
        kw = dict(Ok=cos.Omega_k, w0=cos.w_0, wa=cos.w_a)
        t = oc.growth_table(conf.growth_a, cos.Omega_m, conf.a_lpt_step, **kw)
        for a0, a1 in zip(conf.a_nbody[:-1], conf.a_nbody[1:]):
            am = 0.5 * (a0 + a1)
            d = oc.drift_factor(am, a0, a1, t, conf.growth_a, cos.Omega_m, **kw)
            k = oc.kick_factor(a0, a0, am, t, conf.growth_a, cos.Omega_m, **kw)
            assert integrator.drift_factor(am, a0, a1, cos, conf) == pytest.approx(d, rel=1e-9)
            assert integrator.kick_factor(a0, a0, am, cos, conf) == pytest.approx(k, rel=1e-9)
    assert np.isclose(integrator._factor_dOm("kick_factor", 0.5, 0.5, 0.55, LCDM(), conf),
                      (integrator.kick_factor(0.5, 0.5, 0.55, LCDM(Omega_m=0.30001), conf)
                       - integrator.kick_factor(0.5, 0.5, 0.55, LCDM(Omega_m=0.29999), conf)) / 2e-5, rtol=1e-5)


def test_correction_parameter_layout_and_initialisation():
    import math

    import torch

    from pmpp_amd.corrections.mesh_cnn import initial_params, param_layout

    for depth, block, nconv in [(1, "residual", 2), (2, "residual", 4), (4, "residual", 8), (2, "plain", 3),
                                (4, "plain", 5)]:
        layers, gain, total = param_layout(depth, block)
        assert [L.name for L in layers][0] == "stem" and layers[-1].name == "out"
        assert len(layers) == nconv
        ends = [L.offset + L.size for L in layers]
        # 64-element alignment (16-byte aligned bf16 weights and float32 biases), no overlap
        assert all(L.offset % 64 == 0 for L in layers) and gain % 64 == 0
        assert all(e <= nxt.offset for e, nxt in zip(ends, layers[1:])) and ends[-1] <= gain
        assert total >= gain + 4
    layers, gain, total = param_layout(2)
    p = initial_params(seed=3, depth=2)
    assert p.shape == (total,)
    out = layers[-1]
    assert torch.all(p[out.offset:out.offset + out.size] == 0)  # zero output layer: zero residual
    stem = p[:layers[0].nweights].view(27, 16, 8)
    assert torch.all(stem[:, :, 4:] == 0)  # padding input channels
    std = math.sqrt(1 / (27 * 4)) / 0.87962566103423978
    assert stem[:, :, :4].abs().max() <= 2 * std + 1e-7
    assert abs(float(stem[:, :, :4].std()) - math.sqrt(1 / (27 * 4))) < 0.1 * math.sqrt(1 / (27 * 4))
    conv1 = p[layers[1].offset:layers[1].offset + layers[1].nweights]
    assert abs(float(conv1.std()) - math.sqrt(1 / (27 * 16))) < 0.05 * math.sqrt(1 / (27 * 16))
    assert torch.all(p[layers[0].nweights:layers[0].size] == 0)  # zero biases
    assert 0 < float(p[gain:gain + 3].abs().max()) < 0.05 and p[gain + 3] == 0


def test_validity_mask_is_converted_once_per_mask_and_follows_in_place_changes():
    """The force step's uint8 validity mask (1 = valid) is the negation of ``unused_index``; it
    is converted once while the same mask tensor is passed unchanged (an N-body run carries one
    mask through every step), again after an in-place change, and separately for another mask."""
    import gc
    from types import SimpleNamespace

    import torch

    from pmpp_amd.nbody.gravity import _VALID_MASKS, _valid_u8

    ui = torch.tensor([False, True, False, False])
    p = SimpleNamespace(unused_index=ui, pmid=torch.zeros(4, 3, dtype=torch.int16))
    first = _valid_u8(p)
    assert first.dtype == torch.uint8 and first.tolist() == [1, 0, 1, 1]
    assert _valid_u8(p) is first
    ui[3] = True
    changed = _valid_u8(p)
    assert changed is not first and changed.tolist() == [1, 0, 1, 0]
    other = SimpleNamespace(unused_index=ui.clone(), pmid=p.pmid)
    assert _valid_u8(other) is not changed and _valid_u8(other).tolist() == [1, 0, 1, 0]
    assert _valid_u8(SimpleNamespace(unused_index=None, pmid=p.pmid)) is None
    with pytest.raises(ValueError, match="unused_index has shape"):
        _valid_u8(SimpleNamespace(unused_index=ui[:2], pmid=p.pmid))
    n = len(_VALID_MASKS)
    del p, other, ui, first, changed
    gc.collect()
    assert len(_VALID_MASKS) < n  # entries go with their masks
