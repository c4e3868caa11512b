"""The learned force correction: corrected force step vs the float64 oracle, the network and
its backward vs a float32 PyTorch restatement, the loss and optimizer kernels, and training on
coarse/fine force pairs.

Tolerances: the corrected force and its adjoint carry the force step's float32 bar (1e-5
relative field error vs float64).  The network runs bf16 activations with float32
accumulation, so it is compared with a float32 PyTorch network fed the same bf16-rounded
inputs, weights and activations: 2e-2 relative (norm) on the residual and on each layer's
gradient, which the backward's bf16 rounding of the pre-activation gradients sets.
"""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import cic, spectral
from pmpp_amd import Configuration, Particles, gravity
from pmpp_amd.corrections import (build_correction_optimizer, density, evaluate_residual, force_loss,
                                  gravity_with_residual, gravity_with_residual_vjp,
                                  init_mesh_cnn_potential_correction, residual_backward, train_step)
from pmpp_amd.initial_conditions import LCDM, particles_2lpt

pytestmark = pytest.mark.gpu


def _ptcl(conf, disp_scale, seed, device):
    base = Particles.gen_grid(conf, device=device)
    rng = np.random.default_rng(seed)
    d = base.disp.cpu().numpy() + rng.normal(size=base.disp.shape) * disp_scale * conf.ptcl_spacing
    return base.replace(disp=torch.as_tensor(d.astype(np.float32), device=device))


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _masks(conf):
    return spectral.particle_nyquist_masks(conf.mesh_shape, conf.ptcl_grid_shape, conf.ptcl_spacing,
                                           conf.cell_size)


@pytest.mark.parametrize("ng,ratio,box", [(16, 1, 32.0), (8, 2, 16.0), (16, 2, 48.0)])
def test_corrected_force_and_adjoint_match_oracle(device, ng, ratio, box):
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ptcl = _ptcl(conf, 0.5, ng + ratio, device)
    rng = np.random.default_rng(7)
    rho_add = (0.2 * rng.normal(size=conf.mesh_shape)).astype(np.float32)
    ra = torch.from_numpy(rho_add).to(device)
    acc = _np(gravity_with_residual(ptcl, 0.3, conf, ra))
    pm, dp = ptcl.pmid.cpu().numpy().astype(np.int64), _np(ptcl.disp)
    args = (pm, dp, 0.3)
    ref = spectral.gravity(*args, conf.mesh_shape, conf.cell_size, conf.ptcl_num, _masks(conf), rho_add=rho_add)
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5
    # no residual is the uncorrected step, bit for bit (same kernels, the added load skipped)
    np.testing.assert_array_equal(_np(gravity_with_residual(ptcl, 0.3, conf, None)),
                                  _np(gravity(1.0, ptcl, 0.3, conf)))
    ac = rng.normal(size=acc.shape).astype(np.float32)
    dc, oc, rc = gravity_with_residual_vjp(ptcl, 0.3, conf, ra, torch.from_numpy(ac).to(device))
    dc0, oc0, rc0 = spectral.gravity_vjp(*args, ac, conf.mesh_shape, conf.cell_size, conf.ptcl_num, _masks(conf),
                                         rho_add=rho_add)
    assert np.linalg.norm(_np(rc) - rc0) / np.linalg.norm(rc0) < 1e-5
    assert np.linalg.norm(_np(dc) - dc0) / np.linalg.norm(dc0) < 1e-4
    assert float(oc) == pytest.approx(oc0, rel=1e-4)


def test_density_is_the_step_cic_mesh(device):
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2)
    ptcl = _ptcl(conf, 0.8, 3, device)
    rho = _np(density(ptcl, 0.3, conf))
    pm, dp = ptcl.pmid.cpu().numpy().astype(np.int64), _np(ptcl.disp)
    val = spectral.default_value(len(pm), conf.mesh_shape, conf.ptcl_num)
    ref = cic.scatter(pm, dp, val, np.zeros(conf.mesh_shape), conf.cell_size, conf.mesh_shape)
    np.testing.assert_allclose(rho, ref, atol=2e-6 * np.abs(ref).max())
    assert rho.sum() == pytest.approx(np.prod(conf.mesh_shape), rel=1e-6)


@pytest.mark.parametrize("ng,ratio,box", [(16, 1, 32.0), (8, 2, 16.0), (16, 2, 48.0), (32, 2, 64.0)])
def test_potential_matches_oracle_and_is_symmetric(device, ng, ratio, box):
    """pm_gravity_potential (the network's potential channel) against the float64 oracle at the
    force step's float32 bar; the operator with shift 0 is symmetric (<P u, v> = <u, P v>, so it
    is its own adjoint); the accumulate form equals the separate sum (up to the fused
    multiply-add of the output scaling)."""
    from pmpp_amd.nbody.gravity import gravity_potential

    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ms = conf.mesh_shape
    rng = np.random.default_rng(ng + ratio)
    rho = (1 + 0.5 * rng.normal(size=ms)).astype(np.float32)
    om = 0.3
    got = _np(gravity_potential(torch.from_numpy(rho).to(device), om, conf))
    ref = spectral.potential(rho, om, ms, conf.cell_size, _masks(conf))
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    u = torch.from_numpy(rng.normal(size=ms).astype(np.float32)).to(device)
    v = torch.from_numpy(rng.normal(size=ms).astype(np.float32)).to(device)
    pu = gravity_potential(u, om, conf, shift=0.0, scale=1.0)
    pv = gravity_potential(v, om, conf, shift=0.0, scale=1.0)
    lhs, rhs = float((pu.double() * v.double()).sum()), float((u.double() * pv.double()).sum())
    assert lhs == pytest.approx(rhs, rel=1e-5)
    w = torch.from_numpy(rng.normal(size=ms).astype(np.float32)).to(device)
    acc = w.clone()
    gravity_potential(u, om, conf, out=acc, shift=0.0, scale=0.45, add=acc)
    sep = _np(w + gravity_potential(u, om, conf, shift=0.0, scale=0.45))
    np.testing.assert_allclose(_np(acc), sep, rtol=1e-6, atol=1e-6 * np.abs(sep).max())


def _randomised(model, seed, out_scale=0.05):
    """Non-zero output and gain layers so that every parameter carries gradient."""
    g = torch.Generator().manual_seed(seed)
    p = model.params.cpu()
    out = model.layers[-1]
    p[out.offset:out.offset + out.size] = torch.randn(out.size, generator=g) * out_scale
    p[model.gain_offset:model.gain_offset + 4] = torch.randn(4, generator=g) * 0.3
    for layer in model.layers[:-1]:  # non-zero biases
        p[layer.offset + layer.nweights:layer.offset + layer.size] = torch.randn(16, generator=g) * 0.1
    model.params.copy_(p.to(model.params.device))
    model.sync_bf16()
    return model


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def _torch_conv(x, w, b, cin):
    # x (X,Y,Z,C) float32; w flat bf16 [tap][cout][cin] with tap = (dx+1)*9 + (dy+1)*3 + (dz+1)
    k = w.to(torch.float32).view(3, 3, 3, 16, cin).permute(3, 4, 0, 1, 2)
    xc = x.permute(3, 0, 1, 2)[None]
    xc = F.pad(xc, (1, 1, 1, 1, 1, 1), mode="circular")
    return F.conv3d(xc, k, b)[0].permute(1, 2, 3, 0)


def _green(conf):
    """mask * (-1/k^2) of the configuration (the oracle's float64 symbol) as a torch tensor."""
    return torch.from_numpy(spectral.green_multiplier(conf.mesh_shape, conf.cell_size, _masks(conf)))


def _torch_residual(model, rho, a, om, params, green=None):
    """float32 restatement of the network with the kernels' bf16 roundings (differentiable in
    the float32 parameter vector ``params``, in ``rho`` and in a tensor ``om``).  A model with
    the potential channel reads phi = irfftn(green * rfftn(S)) (``green`` from ``_green``)."""
    r, s8 = model.max_residual, model.sigma8_value
    om = torch.as_tensor(om, dtype=torch.float32)
    c = 1.5 * om
    S = (rho - 1.0) * c
    sigma = torch.sqrt((S.double() ** 2).mean().float() + 1e-6)
    ones = torch.ones_like(rho)
    if model.potential:
        phi = torch.fft.irfftn(torch.fft.rfftn(S.double()) * green, s=tuple(rho.shape)).float()
        sp = torch.sqrt((phi.double() ** 2).mean().float() + 1e-6)
        x0 = torch.stack([S / sigma, phi / sp, a * ones, om * ones, s8 * ones] + [0 * ones] * 3, -1)
    else:
        x0 = torch.stack([S / sigma, a * ones, om * ones, s8 * ones] + [0 * ones] * 4, -1)
    x0 = _bf(x0)

    def layer(i):
        L = model.layers[i]
        return _bf(params[L.offset:L.offset + L.nweights]), params[L.offset + L.nweights:L.offset + L.size]

    gelu = lambda z: F.gelu(z, approximate="tanh")  # noqa: E731
    w, b = layer(0)
    x = _bf(gelu(_torch_conv(x0, w, b, 8)))
    for i in range(model.depth - 1):
        if model.block == "plain":
            w1, b1 = layer(1 + i)
            x = _bf(gelu(_torch_conv(x, w1, b1, 16)))
            continue
        w1, b1 = layer(1 + 2 * i)
        w2, b2 = layer(2 + 2 * i)
        h = _bf(gelu(_torch_conv(x, w1, b1, 16)))
        x = _bf(gelu(_torch_conv(h, w2, b2, 16) + x))
    w, b = layer(len(model.layers) - 1)
    out = _torch_conv(x, w, b, 16)[..., 0]
    gp = params[model.gain_offset:model.gain_offset + 4]
    g = r * torch.tanh(gp[0] * a + gp[1] * om + gp[2] * s8 + gp[3])
    return g * (rho - 1.0) + r * torch.tanh(out) * sigma / c


@pytest.mark.parametrize("depth,shape,block,potential", [
    (1, (8, 8, 8), "residual", False), (2, (16, 8, 24), "residual", False), (3, (16, 16, 16), "residual", False),
    (2, (16, 8, 24), "plain", False), (3, (8, 16, 16), "plain", False), (2, (16, 8, 32), "plain", True),
    (3, (16, 16, 16), "residual", True)])
def test_network_and_backward_match_float32_torch(device, depth, shape, block, potential):
    model = _randomised(init_mesh_cnn_potential_correction(seed=depth, depth=depth, device=device, block=block,
                                                           potential=potential), depth)
    rng = np.random.default_rng(depth)
    rho = torch.from_numpy((1 + 0.5 * rng.normal(size=shape)).astype(np.float32)).to(device)
    a, om = 0.7, 0.31
    conf = Configuration(2.0, tuple(v // 2 for v in shape), mesh_shape=2) if potential else None
    rho_add, tape = evaluate_residual(model, rho, a, om, keep=True, conf=conf)
    params = model.params.detach().cpu().clone().requires_grad_(True)
    ref = _torch_residual(model, rho.cpu(), a, om, params, _green(conf) if potential else None)
    got = rho_add.cpu()
    assert torch.linalg.norm(got - ref.detach()) / torch.linalg.norm(ref.detach()) < 2e-2
    rho_cot = torch.from_numpy(rng.normal(size=shape).astype(np.float32))
    grads = residual_backward(model, tape, rho_cot.to(device)).cpu()
    (ref * rho_cot).sum().backward()
    gref = params.grad
    for L in model.layers:
        sl = slice(L.offset, L.offset + L.size)
        err = torch.linalg.norm(grads[sl] - gref[sl]) / torch.linalg.norm(gref[sl])
        assert err < 2e-2, (L.name, float(err))
    gs = slice(model.gain_offset, model.gain_offset + 4)
    np.testing.assert_allclose(grads[gs].numpy(), gref[gs].numpy(), rtol=2e-2, atol=1e-6)
    # the layout's padding never receives gradient
    live = torch.zeros(model.params.numel(), dtype=torch.bool)
    for L in model.layers:
        live[L.offset:L.offset + L.size] = True
    live[gs] = True
    assert torch.all(grads[~live] == 0)


@pytest.mark.parametrize("shape,potential", [((16, 8, 24), False), ((32, 32, 48), False), ((64, 64, 64), False),
                                             ((16, 8, 32), True), ((64, 64, 64), True)])
def test_inference_stem_from_density_matches_the_feature_path(device, shape, potential):
    """Inference reads the density straight into the stem, with the constant input channels
    folded into its bias; training writes the 8-channel input x0 and runs the generic conv.
    Both compute the same float32 pre-activations up to summation order, so the stem's bf16
    activations agree within one bf16 ulp (at most 2^-7 relative, plus an absolute floor for
    values near zero), on a few per cent of the values at most, and the residual within the
    network's bf16 bar.  The meshes hold wrapped
    and interior blocks and, at 64^3, more blocks than the persistent grid's workgroups.  With
    the potential channel the stem adds a second MFMA step over the staged potential."""
    model = _randomised(init_mesh_cnn_potential_correction(seed=7, depth=2, device=device, potential=potential), 7)
    rng = np.random.default_rng(8)
    rho = torch.from_numpy((1 + 0.5 * rng.normal(size=shape)).astype(np.float32)).to(device)
    a, om = 0.45, 0.29
    conf = Configuration(1.0, tuple(v // 2 for v in shape), mesh_shape=2) if potential else None
    fused, _ = evaluate_residual(model, rho, a, om, conf=conf)
    ref, tape = evaluate_residual(model, rho, a, om, keep=True, conf=conf)
    # stem activations of both paths (the tape keeps the feature path's)
    from pmpp_amd import _lib
    from pmpp_amd._torch import ptr, stream_handle
    act = torch.empty(shape + (16,), dtype=torch.bfloat16, device=device)
    _lib.call("pm_conv3d_stem", ptr(rho), ptr(tape.phi), _lib.int3(shape), om, a, float(model.sigma8_value),
              ptr(tape.stats), ptr(model.weight(model.layers[0])), ptr(model.bias(model.layers[0])), ptr(act),
              stream_handle(device))
    got, exp = act.float().cpu(), tape.acts[0].float().cpu()
    tol = 2.0 ** -7 * exp.abs() + 1e-3 * exp.abs().max()
    assert torch.all((got - exp).abs() <= tol), float(((got - exp).abs() - tol).max())
    assert float((got != exp).float().mean()) < 0.02
    err = torch.linalg.norm((fused - ref).cpu()) / torch.linalg.norm(ref.cpu())
    assert err < 2e-3, float(err)
    # the fused output layer + residual stage against the output conv and the residual kernel
    # on the same activations: the same MFMA sums and the same elementwise arithmetic
    x = tape.acts[-1]
    L = model.layers[-1]
    rho_add = torch.empty_like(rho)
    _lib.call("pm_conv3d_out_residual", ptr(x), ptr(model.weight(L)), ptr(model.bias(L)), ptr(rho),
              _lib.int3(shape), om, float(model.max_residual), ptr(tape.stats), ptr(rho_add), stream_handle(device))
    np.testing.assert_allclose(rho_add.cpu().numpy(), ref.cpu().numpy(), rtol=1e-6,
                               atol=1e-6 * float(ref.abs().max()))


@pytest.mark.parametrize("ng,ratio", [(16, 2), (32, 2), (16, 1)])
def test_corrected_force_resumes_the_density_step(device, ng, ratio):
    """corrected_gravity buckets and scatters once (pm_gravity_density), runs the network and
    resumes the same step (pm_gravity_corrected_resume); that equals the force step run from
    scratch with the same residual (pm_gravity_corrected), bit for bit."""
    from pmpp_amd.corrections.training import corrected_gravity
    conf = Configuration(2.0, (ng,) * 3, mesh_shape=ratio)
    ptcl = _ptcl(conf, 0.7, ng + 1, device)
    model = _randomised(init_mesh_cnn_potential_correction(seed=3, depth=2, device=device), 3)
    a = 0.6
    got = corrected_gravity(a, ptcl, 0.3, conf, model)
    rho_add, _ = evaluate_residual(model, density(ptcl, 0.3, conf), a, 0.3)
    ref = gravity_with_residual(ptcl, 0.3, conf, rho_add)
    np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())


def test_zero_initialised_correction_is_exactly_the_plain_force(device):
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=1)
    ptcl = _ptcl(conf, 0.7, 5, device)
    model = init_mesh_cnn_potential_correction(seed=0, device=device)
    with torch.no_grad():
        model.params[model.gain_offset:model.gain_offset + 4] = 0  # gain 0, output layer 0
    model.sync_bf16()
    acc = gravity(1.0, ptcl, 0.3, conf, correction=model)
    np.testing.assert_array_equal(_np(acc), _np(gravity(1.0, ptcl, 0.3, conf)))


def test_force_loss_and_its_gradient(device):
    rng = np.random.default_rng(2)
    acc = rng.normal(size=(1000, 3)).astype(np.float32)
    tgt = rng.normal(size=(1000, 3)).astype(np.float32)
    loss, cot = force_loss(torch.from_numpy(acc).to(device), torch.from_numpy(tgt).to(device))
    a64, t64 = acc.astype(np.float64), tgt.astype(np.float64)
    d32 = (acc - tgt).astype(np.float64)  # the kernel differences in float32, sums in float64
    ref = np.sum(d32 ** 2) / np.sum(t64 ** 2)
    assert float(loss[0]) == pytest.approx(ref, rel=1e-12)
    np.testing.assert_allclose(_np(cot), 2 * (a64 - t64) / np.sum(t64 ** 2), rtol=1e-6)
    # the loss of the target itself is zero
    loss0, _ = force_loss(torch.from_numpy(tgt).to(device), torch.from_numpy(tgt).to(device), want_grad=False)
    assert float(loss0[0]) == 0.0


def test_adamax_first_step_clipping_and_non_finite_skip(device):
    model = init_mesh_cnn_potential_correction(seed=1, device=device)
    p0 = model.params.clone()
    n = p0.numel()
    rng = np.random.default_rng(3)
    g = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(device)
    opt = build_correction_optimizer(1e-3, gradient_clip_norm=1e9)
    st = opt.init(model.params)
    opt.step(model, g, st)
    # first Adamax step: m_hat = g, u = |g| + eps, so the move is lr g / (|g| + eps)
    np.testing.assert_allclose(_np(model.params - p0), _np(-1e-3 * g / (g.abs() + 1e-8)), rtol=1e-4, atol=2e-8)
    assert int(st.count.item()) == 1
    np.testing.assert_array_equal(_np(model.params_bf16.float()), _np(model.params.to(torch.bfloat16).float()))
    # clipping: the clipped gradient g c / |g| enters the moments
    model2 = init_mesh_cnn_potential_correction(seed=1, device=device)
    opt2 = build_correction_optimizer(1e-3, gradient_clip_norm=1.0)
    st2 = opt2.init(model2.params)
    opt2.step(model2, g, st2)
    gc = g / torch.linalg.norm(g.double()).float()
    np.testing.assert_allclose(_np(st2.mu), _np(0.1 * gc), rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(_np(st2.nu), _np(gc.abs() + 1e-8), rtol=1e-6)
    # a non-finite gradient leaves parameters, moments and the counter untouched
    before = [t.clone() for t in (model2.params, st2.mu, st2.nu)]
    bad = g.clone()
    bad[5] = float("nan")
    opt2.step(model2, bad, st2)
    for t0, t1 in zip(before, (model2.params, st2.mu, st2.nu)):
        assert torch.equal(t0, t1)
    assert int(st2.count.item()) == 1
    # second step against a float64 restatement of clip + adamax
    g2 = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(device)
    p, m, u = _np(model2.params), _np(st2.mu), _np(st2.nu)
    gg = _np(g2)
    gg = gg / np.linalg.norm(gg) if np.linalg.norm(gg) >= 1.0 else gg
    m = 0.9 * m + 0.1 * gg
    u = np.maximum(0.999 * u, np.abs(gg) + 1e-8)
    p = p - 1e-3 * (m / (1 - 0.9 ** 2)) / u
    opt2.step(model2, g2, st2)
    np.testing.assert_allclose(_np(model2.params), p, rtol=1e-5, atol=1e-9)


def _pair(conf, conf_fine, cosmo, seed, device):
    ptcl = particles_2lpt(conf, cosmo, seed=seed, device=device, a=1.0)
    fine = Particles.from_pos(conf_fine, ptcl.pos())
    target = gravity(1.0, fine, cosmo.Omega_m, conf_fine)
    return ptcl, target


def test_training_on_fine_mesh_forces_lowers_held_out_error(device):
    # 32^3 particles in a 64 Mpc/h box at a = 1 (2LPT), coarse mesh 32^3, target mesh 128^3
    cosmo = LCDM()
    conf = Configuration(2.0, (32, 32, 32), mesh_shape=1)
    conf_fine = Configuration(2.0, (32, 32, 32), mesh_shape=4)
    train = [_pair(conf, conf_fine, cosmo, s, device) for s in range(3)]
    held, held_t = _pair(conf, conf_fine, cosmo, 11, device)
    model = init_mesh_cnn_potential_correction(seed=0, depth=2, device=device)
    opt = build_correction_optimizer(3e-3)
    st = opt.init(model.params)

    def held_loss():
        acc = gravity(1.0, held, cosmo.Omega_m, conf, correction=model)
        return float(force_loss(acc, held_t, want_grad=False)[0][0])

    before = held_loss()
    plain = float(force_loss(gravity(1.0, held, cosmo.Omega_m, conf), held_t, want_grad=False)[0][0])
    assert before == pytest.approx(plain, rel=1e-3)  # zero output layer: only the small gain acts
    losses = []
    for it in range(60):
        ptcl, tgt = train[it % len(train)]
        losses.append(float(train_step(model, opt, st, ptcl, 1.0, cosmo, conf, tgt)))
    after = held_loss()
    print(f"held-out relative force error^2: plain {plain:.4f} -> corrected {after:.4f}; "
          f"train first {losses[0]:.4f} last {losses[-1]:.4f}")
    assert np.all(np.isfinite(losses))
    assert after < 0.8 * plain


def test_correction_trained_on_fine_runs_moves_pk_towards_the_fine_run(device):
    # pairs generated by the method: fine-mesh (4x per axis) N-body runs from 3 seeds, every
    # step's state re-anchored on the coarse mesh with its fine-mesh force as the target; the
    # judge is a held-out seed run plain / corrected on the coarse mesh and on the fine mesh
    from pmpp_amd.corrections.training import compare_runs, fine_run_pairs

    cosmo = LCDM()
    conf = Configuration(2.0, (32, 32, 32), mesh_shape=1, a_start=1 / 16, a_nbody_maxstep=1 / 8)
    conf_fine = conf.replace(mesh_shape=4)
    pairs = fine_run_pairs(cosmo, conf, conf_fine, seeds=(0, 1, 2), device=device)
    model = init_mesh_cnn_potential_correction(seed=0, depth=2, device=device)
    opt = build_correction_optimizer(3e-3)
    st = opt.init(model.params)

    def pair_loss(corr):
        tot = 0.0
        for a, ptcl, tgt in pairs:
            acc = gravity(a, ptcl, cosmo.Omega_m, conf, correction=corr)
            tot += float(force_loss(acc, tgt, want_grad=False)[0][0])
        return tot / len(pairs)

    plain_loss = pair_loss(None)
    for it in range(150):
        a, ptcl, tgt = pairs[it % len(pairs)]
        train_step(model, opt, st, ptcl, a, cosmo, conf, tgt)
    trained_loss = pair_loss(model)
    res = compare_runs(model, cosmo, conf, conf_fine, seed=11, device=device)
    k = res["k"].cpu().numpy()
    band = (k > 0.3 * np.pi / conf.ptcl_spacing) & (k < np.pi / conf.ptcl_spacing)
    fine = res["pk_fine"].cpu().numpy()[band]
    err_plain = np.abs(res["pk_plain"].cpu().numpy()[band] / fine - 1).mean()
    err_corr = np.abs(res["pk_corrected"].cpu().numpy()[band] / fine - 1).mean()
    r_plain = res["r_plain"].cpu().numpy()[band].mean()
    r_corr = res["r_corrected"].cpu().numpy()[band].mean()
    print(f"pairs {len(pairs)}; force loss {plain_loss:.4f} -> {trained_loss:.4f}; "
          f"P(k) error vs fine {err_plain:.4f} -> {err_corr:.4f}; r {r_plain:.5f} -> {r_corr:.5f}")
    assert trained_loss < 0.95 * plain_loss
    assert err_corr < 0.8 * err_plain
    assert r_corr > r_plain - 0.03  # the phases stay as close to the fine run's


@pytest.mark.parametrize("case", ["density", "potential", "potential_only"])
def test_input_path_cotangents_match_torch(device, case):
    """The density and Omega_m cotangents through the network's inputs alone (residual_backward
    with inputs: the network's data gradient, pm_correction_input_bwd and, with the potential
    channel, the potential's adjoint) against float32 torch autograd of the restated network.
    ``potential_only`` zeroes the stem's density-channel weights, so every density cotangent
    passes through the potential (and the statistics)."""
    potential = case != "density"
    conf = Configuration(2.0, (16, 16, 16), mesh_shape=2)
    model = _randomised(init_mesh_cnn_potential_correction(seed=11, depth=2, device=device, block="plain",
                                                           potential=potential), 11, out_scale=0.3)
    if case == "potential_only":
        p = model.params.detach().clone()
        stem = model.layers[0]
        w = p[stem.offset:stem.offset + stem.nweights].view(27, 16, 8)
        w[:, :, 0] = 0
        model.params.copy_(p)
        model.sync_bf16()
    rng = np.random.default_rng(13)
    rho = torch.from_numpy((1 + 0.6 * rng.normal(size=conf.mesh_shape)).astype(np.float32)).to(device)
    rho_cot = torch.from_numpy(rng.normal(size=conf.mesh_shape).astype(np.float32)).to(device)
    a, om = 0.6, 0.3
    _, tape = evaluate_residual(model, rho, a, om, keep=True, conf=conf)
    _, rho_bar, omega_bar = residual_backward(model, tape, rho_cot, inputs=True)
    params = model.params.detach().cpu()
    rho_t = rho.cpu().clone().requires_grad_(True)
    om_t = torch.tensor(om, dtype=torch.float32, requires_grad=True)
    ref = _torch_residual(model, rho_t, a, om_t, params, _green(conf) if potential else None)
    (ref * rho_cot.cpu()).sum().backward()
    got, exp = rho_bar.cpu(), rho_t.grad
    err = float(torch.linalg.norm(got - exp) / torch.linalg.norm(exp))
    print(f"[input path] {case}: rho_bar rel err {err:.3e}, omega {float(omega_bar):.6e} vs {float(om_t.grad):.6e}",
          flush=True)
    assert err < 1e-2
    assert float(omega_bar) == pytest.approx(float(om_t.grad), rel=2e-2, abs=1e-3 * float(exp.abs().sum()))


@pytest.mark.parametrize("potential", [False, True])
def test_corrected_force_reverse_through_the_network_inputs(device, potential):
    # disp and Omega_m cotangents of the corrected force: the HIP chain (force-step adjoint at the
    # corrected source, network backward to its input, input-path kernel, scatter adjoint)
    # against the float64 oracle force adjoint plus float32 torch autograd of the network
    # (straight-through bf16 roundings) and the oracle scatter adjoint
    from pmpp_amd.corrections.training import corrected_gravity_vjp, density

    conf = Configuration(2.0, (16, 16, 16), mesh_shape=1)
    ptcl = _ptcl(conf, 0.6, 21, device)
    model = _randomised(init_mesh_cnn_potential_correction(seed=4, depth=2, device=device, potential=potential), 4,
                        out_scale=0.2)
    om, a = 0.3, 0.8
    rng = np.random.default_rng(12)
    ac = rng.normal(size=(conf.ptcl_num, 3)).astype(np.float32)
    dc, oc = corrected_gravity_vjp(a, ptcl, om, conf, model, torch.from_numpy(ac).to(device))
    # reference
    rho = density(ptcl, om, conf).cpu()
    params = model.params.detach().cpu()
    rho_t = rho.clone().requires_grad_(True)
    om_t = torch.tensor(om, dtype=torch.float32, requires_grad=True)
    rho_add = _torch_residual(model, rho_t, a, om_t, params, _green(conf))
    pm, dp = ptcl.pmid.cpu().numpy().astype(np.int64), _np(ptcl.disp)
    dc0, oc0, rc0 = spectral.gravity_vjp(pm, dp, om, ac, conf.mesh_shape, conf.cell_size, conf.ptcl_num,
                                         _masks(conf), rho_add=_np(rho_add))
    (rho_add * torch.from_numpy(rc0.astype(np.float32))).sum().backward()
    val = spectral.default_value(len(pm), conf.mesh_shape, conf.ptcl_num)
    dsc, _ = cic.scatter_adj(pm, dp, val, rho_t.grad.numpy().astype(np.float64), conf.cell_size, conf.mesh_shape)
    ref = dc0 + dsc
    err = np.linalg.norm(_np(dc) - ref) / np.linalg.norm(ref)
    assert err < 2e-2, err
    assert np.linalg.norm(dsc) > 0.05 * np.linalg.norm(dc0)  # the input path is not negligible here
    assert oc == pytest.approx(oc0 + float(om_t.grad), rel=2e-2)
    # and the autograd wrapper gives the same
    disp = ptcl.disp.clone().requires_grad_(True)
    omt = torch.tensor(om, dtype=torch.float32, device=device, requires_grad=True)

    class _C:
        Omega_m = omt

    acc = gravity(a, ptcl.replace(disp=disp), _C(), conf, correction=model)
    (acc * torch.from_numpy(ac).to(device)).sum().backward()
    np.testing.assert_allclose(_np(disp.grad), _np(dc), rtol=1e-5, atol=1e-6 * np.abs(_np(dc)).max())


def test_nbody_reverse_sweep_with_a_correction_matches_finite_differences(device):
    # the N-body adjoint with the corrected force (per-step recomputation through the network)
    # against central differences of the corrected forward run along one random direction;
    # float32 forward, so eps and the tolerance are loose
    from pmpp_amd.nbody.solver import nbody

    conf = Configuration(4.0, (16, 16, 16), mesh_shape=1, a_start=0.5, a_nbody_maxstep=0.25)
    cos = LCDM()
    base = _ptcl(conf, 0.3, 31, device)
    base = base.replace(vel=torch.zeros_like(base.disp))
    model = _randomised(init_mesh_cnn_potential_correction(seed=5, depth=1, device=device), 5, out_scale=0.1)
    rng = np.random.default_rng(3)
    cot = torch.from_numpy(rng.normal(size=(conf.ptcl_num, 3)).astype(np.float32)).to(device)
    disp = base.disp.clone().requires_grad_(True)
    out = nbody(base.replace(disp=disp), cos, conf, correction=model)
    (out.disp * cot).sum().backward()
    v = torch.from_numpy(rng.normal(size=(conf.ptcl_num, 3)).astype(np.float32)).to(device)
    eps = 2e-2
    with torch.no_grad():
        fp = nbody(base.replace(disp=base.disp + eps * v), cos, conf, correction=model).disp
        fm = nbody(base.replace(disp=base.disp - eps * v), cos, conf, correction=model).disp
    fd = float(((fp - fm) * cot).sum()) / (2 * eps)
    assert float((disp.grad * v).sum()) == pytest.approx(fd, rel=5e-2)


@pytest.mark.parametrize("slab,potential", [(8, False), (16, False), (8, True)])
def test_network_reverse_by_slabs_equals_the_whole_mesh(device, slab, potential):
    """The corrected force's reverse with the network's backward run over x-slabs with a halo
    (network_input_vjp) against the same backward on the whole mesh: the slab's own planes
    see exactly the whole-mesh arithmetic, so the cotangents agree to summation order."""
    from pmpp_amd.corrections.training import corrected_gravity_vjp

    conf = Configuration(2.0, (32, 16, 16), mesh_shape=1)
    ptcl = _ptcl(conf, 0.6, 41, device)
    model = _randomised(init_mesh_cnn_potential_correction(seed=6, depth=2, device=device, block="plain",
                                                           potential=potential), 6, out_scale=0.2)
    rng = np.random.default_rng(14)
    ac = torch.from_numpy(rng.normal(size=(conf.ptcl_num, 3)).astype(np.float32)).to(device)
    d_full, o_full = corrected_gravity_vjp(0.7, ptcl, 0.3, conf, model, ac)
    d_slab, o_slab = corrected_gravity_vjp(0.7, ptcl, 0.3, conf, model, ac, slab=slab)
    np.testing.assert_allclose(_np(d_slab), _np(d_full), rtol=1e-5, atol=1e-6 * np.abs(_np(d_full)).max())
    assert o_slab == pytest.approx(o_full, rel=1e-5)


def test_whole_run_loss_parameter_gradient_matches_central_differences(device):
    """The end-to-end training gradient (log P(k) and cross-correlation loss of a corrected run,
    through the P(k) and r adjoints, the scatter adjoint and the N-body reverse sweep with the
    network's parameter gradient) against central differences along a random direction of the
    float32 biases and gain parameters (the bf16 weight mirror does not see small steps)."""
    from pmpp_amd.corrections.training import RunTarget, run_loss

    cos = LCDM()
    conf = Configuration(4.0, (16, 16, 16), mesh_shape=2, a_start=1 / 16, a_nbody_maxstep=0.25)
    tgt = RunTarget(3, cos, conf, conf.replace(mesh_shape=4), device=device)
    model = _randomised(init_mesh_cnn_potential_correction(seed=8, depth=2, block="plain", max_residual=1.0,
                                                           device=device), 8, out_scale=0.2)
    grads = torch.zeros_like(model.params)
    run_loss(model, tgt, cos, conf, grads=grads)
    g = torch.Generator().manual_seed(9)
    v = torch.zeros(model.params.numel())
    for layer in model.layers:
        v[layer.offset + layer.nweights:layer.offset + layer.size] = torch.randn(16, generator=g)
    v[model.gain_offset:model.gain_offset + 4] = torch.randn(4, generator=g)
    v = v.to(device)
    got = float((grads * v).sum())
    base = model.params.detach().clone()
    eps = 1e-2

    def at(t):
        model.params.copy_(base + t * v)
        model.sync_bf16()
        return run_loss(model, tgt, cos, conf)

    fd = (at(eps) - at(-eps)) / (2 * eps)
    at(0.0)
    print(f"whole-run loss gradient . v: adjoint {got:.6e}, central differences {fd:.6e}")
    assert got == pytest.approx(fd, rel=5e-2)


@pytest.mark.parametrize("shape,potential", [((8, 8, 16), False), ((32, 32, 48), False), ((64, 64, 64), False),
                                             ((128, 128, 128), False), ((256, 64, 64), False), ((200, 128, 128), False),
                                             ((8, 8, 16), True), ((64, 64, 64), True), ((128, 128, 128), True)])
def test_fused_stem_and_hidden_layer_equal_the_two_kernels(device, shape, potential):
    """pm_conv3d_stem_hidden (the stem computed on the hidden layer's staged tile in LDS) against
    pm_conv3d_stem followed by the hidden conv: the same arithmetic in the same order, so the
    activations agree bit for bit; meshes of one block along y and z (every halo wraps) and of
    several blocks along each axis; with and without the potential channel.  On the small meshes
    every block is its own x run; at 128^3 the runs are 8 blocks long, at 256 x 64 x 64 four
    (two staged planes carried from the previous block in LDS for 7 of 8 and 3 of 4 blocks) and
    at 200 x 128 x 128 thirteen, the last run of each x line eleven."""
    from pmpp_amd import _lib
    from pmpp_amd._torch import ptr, stream_handle
    from pmpp_amd.corrections import conv3d

    model = _randomised(init_mesh_cnn_potential_correction(seed=9, depth=2, block="plain", device=device,
                                                           potential=potential), 9)
    rng = np.random.default_rng(10)
    rho = torch.from_numpy((1 + 0.5 * rng.normal(size=shape)).astype(np.float32)).to(device)
    a, om = 0.55, 0.31
    conf = Configuration(1.0, tuple(v // 2 for v in shape), mesh_shape=2) if potential else None
    _, tape = evaluate_residual(model, rho, a, om, keep_out=True, conf=conf)  # the statistics (and phi)
    L0, L1 = model.layers[0], model.layers[1]
    s = stream_handle(device)
    stem = torch.empty(shape + (16,), dtype=torch.bfloat16, device=device)
    _lib.call("pm_conv3d_stem", ptr(rho), ptr(tape.phi), _lib.int3(shape), om, a, float(model.sigma8_value),
              ptr(tape.stats), ptr(model.weight(L0)), ptr(model.bias(L0)), ptr(stem), s)
    _, ref = conv3d(stem, model.weight(L1), model.bias(L1), gelu=True, want_pre=False)
    got = torch.empty(shape + (16,), dtype=torch.bfloat16, device=device)
    _lib.call("pm_conv3d_stem_hidden", ptr(rho), ptr(tape.phi), _lib.int3(shape), om, a, float(model.sigma8_value),
              ptr(tape.stats), ptr(model.weight(L0)), ptr(model.bias(L0)), ptr(model.weight(L1)), ptr(model.bias(L1)),
              ptr(got), s)
    np.testing.assert_array_equal(got.float().cpu().numpy(), ref.float().cpu().numpy())
