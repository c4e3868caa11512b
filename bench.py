"""Benchmark of the PM force step on one MI355X.

A step is one PM force evaluation over the whole particle set: cell sort, CIC scatter, r2c FFT,
fused mask x Green x gradient, the three c2r FFTs and the CIC gather (``pm_gravity``), with the
particles already resident in HBM.  Default workload: 512^3 particles on a 1024^3 float32 mesh
in a 1000 Mpc/h box (LCDM Omega_m=0.3, Omega_b=0.05, h=0.7, n_s=0.96, sigma_8=0.8), positions
= Lagrangian grid + U(-1/2,1/2) l_p jitter (seed 1) + 2LPT displacement at a=1 (white-noise
seed 0).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5|c2] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (particles per side, mesh ratio, box Mpc/h)
    "c5": (512, 2, 1000.0),
    "c2": (256, 2, 256.0),
    "small": (64, 2, 128.0),
}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def num_bricks(mesh, brick):
    nb = 1
    for m, b in zip(mesh, brick):
        nb *= -(-m // b)
    return nb


def kernel_step_bytes(name, L, n_p, brick):
    """Algorithmic (compulsory) HBM bytes per step of every launch of one kernel name.

    R = real float32 mesh; S1 / S2 = the complex64 spectrum band the force passes keep
    (kz < nkz; all y rows / the packed |ky| band), pitch excluded; n_p particles of 18 B
    (int16 pmid + float32 disp).  The tiled scatter hands the upper x/y planes of each brick
    column to the fold; tile reads of the gather count the mesh once.
    """
    R = 4 * L["nreal"]
    Ch = 8 * L["nx"] * L["ny"] * (L["nz"] // 2 + 1)
    S1 = 8 * L["nx"] * L["ny"] * L["nkz"]  # banded spectrum, full y rows
    S2 = 8 * L["nx"] * L["nky"] * L["nkz"]  # banded spectrum, packed y band
    nb = num_bricks((L["nx"], L["ny"], L["nz"]), brick)
    b0, b1, b2 = brick
    nb0, nb1 = L["nx"] // b0, L["ny"] // b1
    planes = 4 * nb0 * nb1 * (b1 + 1 + b0) * L["nz"]  # upper x/y planes of every brick column
    fold = 8 * (nb0 * L["ny"] + (L["nx"] - nb0) * nb1) * L["nz"]  # lower x/y plane cells read + written
    table = {
        "fft_z_r2c": R + S1,
        "fft_cols_fwd": S1 + S2,
        "fft_cols_green": 3 * S2,
        "fft_cols_inv2": S2 + 2 * S1,
        "fft_cols_inv": S2 + S1,
        "fft_z_c2r": 3 * (S1 + R),
        "k_bucket_count": 18 * n_p,
        "k_bucket_place": (18 + 4 + 18 + 1) * n_p,
        "k_zero_i32": 4 * (nb + 2),
        "scan": 8 * (nb + 2),
        "k_scatter_columns": 18 * n_p + R + planes,
        "k_fold_xy": planes + fold,
        "k_gather_columns": (18 + 4 + 12) * n_p + 3 * R,
        "k_fill": R,
        "k_scatter_sorted": 19 * n_p + R,
        "k_gather3_sorted": (19 + 4 + 12) * n_p + 3 * R,
    }
    return table.get(name)


def official_step_bytes(n_p, L):
    """Step bytes under the accounting of the reference analysis: each 3-D FFT as 3 axis passes."""
    R = 4 * L["nreal"]
    C = 8 * L["nx"] * L["ny"] * (L["nz"] // 2 + 1)
    return 66 * n_p + (18 * n_p + R) + (R + 5 * C) + 3 * (R + 5 * C) + (34 * n_p + 3 * R)


def spectral_band(conf):
    """Stored kz columns and y rows of the force passes: with the particle-Nyquist filter only
    |k_i| <= pi/l_p (1 + 8 eps_f32) survive the Green multiply (same rule as the library)."""
    nz, ny = conf.mesh_shape[2], conf.mesh_shape[1]
    nkz, nky = nz // 2 + 1, ny
    if conf.mask_active:
        def half_width(n):
            kmax = math.pi / conf.ptcl_spacing * (1 + 8 * 1.1920928955078125e-07) * (1 + 1e-4)
            return math.floor(kmax / (2 * math.pi / (n * conf.cell_size)))
        kz, ky = half_width(nz), half_width(ny)
        if kz < nz // 2:
            nkz = kz + 1
        if 2 * ky + 1 < ny:
            nky = 2 * ky + 1
    return {"nkz": nkz, "nky": nky}


def strict_step_bytes(n_p, L):
    """Lower bound with every 3-D FFT as a single read + write of its input and output."""
    R = 4 * L["nreal"]
    C = 8 * L["nx"] * L["ny"] * (L["nz"] // 2 + 1)
    return 66 * n_p + (18 * n_p + R) + (R + C) + 3 * (C + R) + (34 * n_p + 3 * R)


def measured_traffic(kernel, config):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, if any.

    ``profiles/traffic_<config>.txt`` holds lines ``<name> <fetch KiB> <fetch factor> <write KiB>
    <write factor>`` per dispatch, where each factor is what the counter under-reads a plain
    1 GiB device copy by (bytes moved = KiB * 1024 * factor).
    """
    path = os.path.join(ROOT, "profiles", f"traffic_{config}.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 5 and parts[0] == kernel:
                    fk, ff, wk, wf = (float(v) for v in parts[1:])
                    return round((fk * ff + wk * wf) * 1024)
    except (OSError, ValueError):
        return None
    return None


def cpu_baseline(dev, omega_m):
    """The reference's operator sequence restated in float32 on the host cores: the 256^3-particle /
    512^3-mesh force step (lattice + jitter + 2LPT at a = 1 in a 256 Mpc/h box, the same input the
    GPU takes for that workload), 1 warm-up then the median of 5; and the whole 64^3-particle /
    128^3-mesh / 128 Mpc/h run (white noise -> 2LPT at a = 1/64 -> 10 KDK steps), 1 warm-up then
    one timed run.  CIC in C/OpenMP, FFTs in scipy with one worker per core."""
    import numpy as np

    from oracle.cpu_baseline import force_step, nbody_run
    from pmpp_amd import Configuration
    from pmpp_amd.initial_conditions import LCDM, particles_2lpt

    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    ng, ratio, box = CONFIGS["c2"]
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ptcl = particles_2lpt(conf, LCDM(), seed=0, jitter_seed=1, device=dev, a=1.0)
    pm = ptcl.pmid.cpu().numpy()
    dp = ptcl.disp.cpu().numpy()
    del ptcl
    args = (pm, dp, omega_m, conf.mesh_shape, conf.cell_size, conf.ptcl_grid_shape, conf.ptcl_spacing, cores)
    force_step(*args)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        force_step(*args)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    nbody_run(64, 2, 128.0, 10, 0, LCDM(), workers=cores)  # warm-up
    t0 = time.perf_counter()
    nbody_run(64, 2, 128.0, 10, 0, LCDM(), workers=cores)
    t_c1 = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": conf.ptcl_num / med,
        "unit": "particles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"median of 5 float32 force steps at {ng}^3 particles on a {ng * ratio}^3 mesh ({box:g} Mpc/h, "
                  f"lattice + jitter + 2LPT at a = 1: {med:.2f} s/step, {1.0 / med:.3f} steps/s; CIC in C/OpenMP, "
                  f"scipy FFTs, {cores} threads); CPU: {model or platform.processor()}",
        "c2_force_step": {"s_per_step": round(med, 4), "steps_per_s": round(1.0 / med, 4),
                          "particles_per_s": conf.ptcl_num / med, "particles": conf.ptcl_num,
                          "mesh": list(conf.mesh_shape)},
        "c1_end_to_end": {"s": round(t_c1, 3), "kdk_steps": 10, "steps_per_s": round(10 / t_c1, 3),
                          "workload": "64^3 particles, 128^3 mesh, 128 Mpc/h: white noise (seed 0) -> 2LPT at "
                                      "a = 1/64 -> 10 KDK steps to a = 1 (float64 ICs, float32 force)"},
    }


def corrected_run(args):
    """Only the corrected forward run of ``--nbody-run`` (512^3 particles, 1024^3 mesh, 1000
    Mpc/h, 2LPT at a = 1/64, ``--run-steps`` KDK steps with the shipped correction), timed after
    one warm run: a short command to profile the corrected step's kernels.  Prints one JSON line
    (not the headline metric)."""
    import torch

    from pmpp_amd import Configuration
    from pmpp_amd.corrections import load_correction
    from pmpp_amd.initial_conditions import LCDM, linear_modes, lpt, white_noise
    from pmpp_amd.nbody.solver import nbody

    dev = torch.device("cuda", 0)
    ng, ratio, box = CONFIGS["c5"]
    a0, nsteps = 1 / 64, args.run_steps
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio, a_start=a0, a_nbody_maxstep=(1 - a0) / nsteps)
    cosmo = LCDM()
    ptcl = lpt(linear_modes(white_noise(0, conf, device=dev), cosmo, conf), cosmo, conf)
    model = load_correction(args.weights, device=dev)
    nbody(ptcl, cosmo, conf, correction=model, observer=lambda a, p: None)  # warm (allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nbody(ptcl, cosmo, conf, correction=model, observer=lambda a, p: None)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    print(json.dumps({"metric": "corrected N-body run steps/s (512^3 particles, 1024^3 mesh, 1000 Mpc/h)",
                      "steps_per_s": nsteps / t, "s": t, "steps": nsteps, "forces": nsteps + 1,
                      "potential_channel": bool(model.potential),
                      "config": {"workload": "corrected N-body run: 512^3 particles, 1024^3 mesh, 1000 Mpc/h"}}),
          flush=True)


def nbody_run(args):
    """The full single-GPU run of the 512^3 / 1024^3 / 1000 Mpc/h workload: 2LPT at a = 1/64,
    20 KDK steps to a = 1 (plain), the reverse sweep of that run, the same run with the shipped
    trained correction (the three-layer BF16 network) and its reverse sweep (the network's
    backward over x-slabs), the memory high-water marks, and the same plain run on a 2048^3 mesh
    (4x the particle grid) for comparison; every forward run is timed after one warm run.
    Prints one JSON line (not the headline metric)."""
    import torch

    from pmpp_amd import Configuration
    from pmpp_amd.corrections import load_correction
    from pmpp_amd.initial_conditions import LCDM, linear_modes, lpt, white_noise
    from pmpp_amd.nbody.solver import nbody, nbody_adj

    dev = torch.device("cuda", 0)
    ng, ratio, box = CONFIGS["c5"]
    a0, nsteps = 1 / 64, args.run_steps
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio, a_start=a0, a_nbody_maxstep=(1 - a0) / nsteps)
    cosmo = LCDM()
    log(f"[run] 2LPT at a = {a0:g}: {ng}^3 particles")
    ptcl = lpt(linear_modes(white_noise(0, conf, device=dev), cosmo, conf), cosmo, conf)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    nbody(ptcl, cosmo, conf, observer=lambda a, p: None)  # warm (allocator, workspace), as for every run
    out, t_plain = timed(lambda: nbody(ptcl, cosmo, conf, observer=lambda a, p: None))
    log(f"[run] plain: {nsteps} steps in {t_plain:.3f} s")
    g = torch.Generator(device="cpu").manual_seed(0)
    cot = out.replace(disp=torch.randn(out.disp.shape, generator=g).to(dev), vel=torch.zeros_like(out.vel),
                      acc=torch.zeros_like(out.acc))
    nbody_adj(out, cot, cosmo, conf)  # warm
    _, t_adj = timed(lambda: nbody_adj(out, cot, cosmo, conf))
    log(f"[run] reverse sweep in {t_adj:.3f} s")
    del cot, out
    corrected = {}
    model = load_correction(args.weights, device=dev)
    nbody(ptcl, cosmo, conf, correction=model, observer=lambda a, p: None)  # warm (allocator)
    out, t_corr = timed(lambda: nbody(ptcl, cosmo, conf, correction=model, observer=lambda a, p: None))
    corrected = {"conv_layers": len(model.layers), "steps_per_s": nsteps / t_corr, "s": t_corr,
                 "weights": "shipped trained correction"}
    log(f"[run] corrected ({len(model.layers)} conv layers, trained): {nsteps} steps in {t_corr:.3f} s")
    peak = torch.cuda.max_memory_allocated() / 2 ** 30
    cot = out.replace(disp=torch.randn(out.disp.shape, generator=g).to(dev), vel=torch.zeros_like(out.vel),
                      acc=torch.zeros_like(out.acc))
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    _, t_cadj = timed(lambda: nbody_adj(out, cot, cosmo, conf, correction=model))
    peak_cadj = torch.cuda.max_memory_allocated() / 2 ** 30
    log(f"[run] reverse sweep with the correction in {t_cadj:.3f} s, memory high-water mark {peak_cadj:.1f} GiB")
    corrected["reverse_steps_per_s"] = nsteps / t_cadj
    corrected["reverse_s"] = t_cadj
    corrected["reverse_memory_high_water_GiB"] = round(peak_cadj, 2)
    del cot, out, model
    torch.cuda.empty_cache()
    conf4 = conf.replace(mesh_shape=4)
    nbody(ptcl, cosmo, conf4, observer=lambda a, p: None)  # warm (allocator), as for the corrected runs
    out4, t4 = timed(lambda: nbody(ptcl, cosmo, conf4, observer=lambda a, p: None))
    log(f"[run] plain on the 2048^3 mesh: {nsteps} steps in {t4:.3f} s")
    del out4
    res = {
        "metric": "N-body run steps/s (512^3 particles, 1024^3 mesh, 1000 Mpc/h, 20 KDK steps)",
        "steps_per_s_plain": nsteps / t_plain, "steps_per_s_reverse": nsteps / t_adj,
        "steps_per_s_corrected": corrected["steps_per_s"], "corrected": corrected,
        "memory_high_water_GiB": round(peak, 2),
        "steps_per_s_plain_2048_mesh": nsteps / t4,
        "corrected_vs_plain_2048_wall": corrected["s"] / t4,
        "n_gpus": 1, "dtype": "f32 (network bf16)",
        "data": "synthetic (2LPT LCDM field from seeded white noise)",
        "config": {"workload": "N-body run: 512^3 particles, 1024^3 mesh, 1000 Mpc/h, a = 1/64 -> 1",
                   "steps": nsteps, "comparison_mesh": 2048},
    }
    print(json.dumps(res), flush=True)


def _parse_floats(text):
    return [float(v) for v in str(text).split(",") if v]


def train_correction_run(args):
    """The learned-correction workload: 128^3 particles in a 250 Mpc/h box (the 1.95 Mpc/h
    particle spacing of the full-scale workload), coarse mesh 256^3 (2x the particle grid), fine
    mesh 512^3 (4x), 2LPT at a = 1/64 then ``--run-steps`` KDK steps to a = 1.

    Stage 1 (force pairs): every other step of the fine runs of seeds 100 .. 100 +
    ``--train-seeds`` - 1, re-anchored on the coarse mesh with the fine force as the target;
    each (kernel symmetry, max_residual, learning rate) of the grid trains the three-layer
    network for ``--train-iters`` Adamax steps on the relative force error.  Stage 2 (whole
    runs): the best stage-1 network is fine-tuned for ``--e2e-iters`` steps on whole corrected
    runs of seeds 100 .. 100 + ``--e2e-seeds`` - 1 (loss: mean squared log P(k) error plus the
    mean 1 - r against the fine run over 0 < k <= 1 h/Mpc, through the N-body reverse sweep),
    keeping the checkpoint with the best validation score.  Every choice uses the validation
    seed 999 alone (score: largest |P/P_fine - 1| up to k = 1 h/Mpc plus 30 x any shortfall of
    the cross-correlation below the plain run's); only the final network is judged, on the
    held-out seeds 1000 and 1001.  Densities are measured on the coarse mesh.  ``--init-from PATH``
    skips stage 1 and fine-tunes saved weights instead (with ``--potential``, a network without
    the potential channel gets it with zero weights first).  ``--save PATH``
    writes the weights there (the package ships them as
    ``pmpp_amd/corrections/weights/mesh_cnn_2x_from_4x.npz``).  Prints one JSON line."""
    import torch

    from pmpp_amd import Configuration
    from pmpp_amd.corrections import (compare_runs, correction_meta, fine_run_pairs, load_correction, mean_force_loss,
                                      pk_errors, save_correction, train_correction, with_potential_channel)
    from pmpp_amd.corrections.training import RunTarget, train_correction_runs
    from pmpp_amd.initial_conditions import LCDM

    dev = torch.device("cuda", 0)
    cosmo = LCDM()
    ng, box, a0 = 128, 250.0, 1 / 64
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=2, a_start=a0, a_nbody_maxstep=(1 - a0) / args.run_steps)
    conf_fine = conf.replace(mesh_shape=4)

    def judge(model, seed):
        return pk_errors(compare_runs(model, cosmo, conf, conf_fine, seed=seed, device=dev, measure_conf=conf))

    def score(v):
        c, p = v["corrected"], v["plain"]
        return c["pk_err_max"] + 30.0 * max(0.0, p["r_mean"] - c["r_mean"])

    t0 = time.perf_counter()
    stage1 = args.init_from is None
    init_meta = {} if stage1 else correction_meta(args.init_from)
    pairs = (fine_run_pairs(cosmo, conf, conf_fine, seeds=range(100, 100 + args.train_seeds), every=2, device=dev)
             if stage1 else [])
    val_pairs = fine_run_pairs(cosmo, conf, conf_fine, seeds=[999], every=2, device=dev)
    torch.cuda.synchronize()
    log(f"[train] {len(pairs)} training pairs from {args.train_seeds if stage1 else 0} fine runs in "
        f"{time.perf_counter() - t0:.1f} s")
    val_plain = mean_force_loss(None, val_pairs, cosmo, conf)
    cands = []
    grid = [(sym, mr, lr) for sym in args.symmetry.split(",") for mr in _parse_floats(args.max_residual)
            for lr in _parse_floats(args.lr)] if stage1 else []
    if not stage1:
        init = load_correction(args.init_from, device=dev)
        if args.potential:
            init = with_potential_channel(init)
        val = judge(init, 999)
        cands.append({"symmetry": init.symmetry, "max_residual": init.max_residual,
                      "lr": float(init_meta.get("lr", 3e-4)), "train_s": 0.0,
                      "init_from": str(args.init_from), "val_force_loss": mean_force_loss(init, val_pairs, cosmo, conf),
                      "val_force_loss_plain": val_plain, "val": val, "score": score(val), "model": init})
        log(f"[train] start from {args.init_from}: val P(k) err to k=1 max {val['corrected']['pk_err_max']:.4f}, "
            f"r {val['corrected']['r_mean']:.6f} (plain {val['plain']['r_mean']:.6f})")
    for sym, mr, lr in grid:
        tag = f"symmetry {sym} max_residual {mr:g} lr {lr:g}"
        t0 = time.perf_counter()
        model, losses = train_correction(
            pairs, cosmo, conf, iters=args.train_iters, learning_rate=lr, max_residual=mr, symmetry=sym, device=dev,
            potential=args.potential, progress=lambda it, m, tag=tag: log(f"[train] {tag} step {it}: mean training loss {m:.6f}"))
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        val_loss = mean_force_loss(model, val_pairs, cosmo, conf)
        val = judge(model, 999)
        q = max(1, len(losses) // 10)
        cands.append({"symmetry": sym, "max_residual": mr, "lr": lr, "train_s": round(t_train, 2),
                      "train_loss_first_tenth": float(losses[:q].mean()),
                      "train_loss_last_tenth": float(losses[-q:].mean()),
                      "val_force_loss": val_loss, "val_force_loss_plain": val_plain, "val": val, "score": score(val),
                      "model": model})
        log(f"[train] {tag}: val force loss {val_loss:.6f} (plain {val_plain:.6f}), "
            f"val P(k) err to k=1: max {val['corrected']['pk_err_max']:.4f} "
            f"(plain {val['plain']['pk_err_max']:.4f}), r {val['corrected']['r_mean']:.6f} "
            f"(plain {val['plain']['r_mean']:.6f})")
    best = min(cands, key=lambda c: c["score"])
    model = best["model"]
    n_pairs = len(pairs)
    del pairs
    stage2 = None
    if args.e2e_iters > 0:
        t0 = time.perf_counter()
        targets = [RunTarget(sd, cosmo, conf, conf_fine, device=dev) for sd in range(100, 100 + args.e2e_seeds)]
        vals = []

        def validate(m):
            v = judge(m, 999)
            vals.append(v)
            log(f"[e2e] validation: P(k) err to k=1 max {v['corrected']['pk_err_max']:.4f} "
                f"mean {v['corrected']['pk_err_mean']:.4f}, r {v['corrected']['r_mean']:.6f} "
                f"(plain {v['plain']['r_mean']:.6f})")
            return score(v)

        e2e_losses, scores = train_correction_runs(
            model, targets, cosmo, conf, iters=args.e2e_iters, learning_rate=args.e2e_lr, r_weight=args.r_weight,
            validate=validate, validate_every=args.e2e_validate_every,
            progress=lambda it, l: log(f"[e2e] step {it}: run loss {l:.6f}") if it % 10 == 0 else None)
        torch.cuda.synchronize()
        stage2 = {"seeds": f"100-{99 + args.e2e_seeds}", "iters": args.e2e_iters, "lr": args.e2e_lr,
                  "r_weight": args.r_weight, "s": round(time.perf_counter() - t0, 2),
                  "loss_first_tenth": float(e2e_losses[:max(1, len(e2e_losses) // 10)].mean()),
                  "loss_last_tenth": float(e2e_losses[-max(1, len(e2e_losses) // 10):].mean()),
                  "validation_scores": [(int(i), round(float(v), 5)) for i, v in scores],
                  "kept_step": int(min(scores, key=lambda t: t[1])[0])}
        del targets
    judged = {}
    for sd in (1000, 1001):
        judge_pairs = fine_run_pairs(cosmo, conf, conf_fine, seeds=[sd], every=2, device=dev)
        judged[sd] = {**judge(model, sd), "force_loss_corrected": mean_force_loss(model, judge_pairs, cosmo, conf),
                      "force_loss_plain": mean_force_loss(None, judge_pairs, cosmo, conf)}
        log(f"[train] judge seed {sd}: {json.dumps(judged[sd])}")
    recipe = {"particles": ng, "box": box, "coarse_mesh": 2 * ng, "fine_mesh": 4 * ng, "kdk_steps": args.run_steps,
              "a_start": a0, "train_seeds": f"100-{99 + args.train_seeds}",
              "pairs": n_pairs if stage1 else int(init_meta.get("pairs", 0)),
              "iters": args.train_iters if stage1 else int(init_meta.get("iters", 0)), "validation_seed": 999,
              "judge_seed": 1000,
              "max_residual": best["max_residual"], "lr": best["lr"], "symmetry": best["symmetry"],
              "e2e_iters": args.e2e_iters, "e2e_seeds": f"100-{99 + args.e2e_seeds}", "e2e_lr": args.e2e_lr,
              "r_weight": args.r_weight, "potential": int(args.potential),
              "init_from": "" if stage1 else str(args.init_from)}
    if args.save:
        save_correction(model, args.save, **recipe)
        log(f"[train] wrote {args.save}")
    out = {
        "metric": "learned correction: P(k) error of the corrected 2x-mesh run vs the 4x-mesh run (held-out seeds)",
        "recipe": recipe,
        "judge": judged,
        "stage1_candidates": [{k: v for k, v in c.items() if k != "model"} for c in cands],
        "stage2": stage2,
        "config": {"workload": f"learned correction: {ng}^3 particles, {box:g} Mpc/h, coarse {2 * ng}^3 vs fine "
                               f"{4 * ng}^3, {args.run_steps} KDK steps"},
    }
    print(json.dumps(out), flush=True)


def pk_error_run(dev, nsteps=20, seed=0, weights=None):
    """The second half of the headline metric at full scale: from one seed's white noise, the
    512^3-particle / 1000 Mpc/h run (2LPT at a = 1/64, ``nsteps`` KDK steps to a = 1) on the
    1024^3 mesh with the shipped trained correction, and the plain run on the 2048^3 mesh (4x
    the particle grid); both final densities are measured on the 1024^3 mesh (CIC, deconvolved),
    and each run is timed after one warm run.  Returns a dict (P(k) error and cross-correlation
    over 0 < k <= 1 h/Mpc for the corrected and the plain 2x run, and the wall times).
    ``weights`` (a path) replaces the shipped correction."""
    import torch

    from pmpp_amd import Configuration
    from pmpp_amd.analysis import density_to_cross_correlation, density_to_pk
    from pmpp_amd.cic import scatter
    from pmpp_amd.corrections import correction_meta, load_correction
    from pmpp_amd.initial_conditions import LCDM, linear_modes, lpt, white_noise
    from pmpp_amd.nbody.particles import Particles
    from pmpp_amd.nbody.solver import nbody

    ng, ratio, box = CONFIGS["c5"]
    a0 = 1 / 64
    cosmo = LCDM()
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio, a_start=a0, a_nbody_maxstep=(1 - a0) / nsteps)
    conf4 = conf.replace(mesh_shape=4)
    model = load_correction(weights, device=dev)
    modes = linear_modes(white_noise(seed, conf, device=dev), cosmo, conf)

    def timed(c, corr):
        p0 = lpt(modes, cosmo, c)
        nbody(p0, cosmo, c, correction=corr, observer=lambda a, p: None)  # warm (allocator, workspaces)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = nbody(p0, cosmo, c, correction=corr, observer=lambda a, p: None)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        del p0
        rho = scatter(Particles.from_pos(conf, out.pos(dtype=torch.float64)), conf)
        del out
        torch.cuda.empty_cache()
        return rho, t

    log("[pk] corrected run on the 1024^3 mesh")
    d_corr, t_corr = timed(conf, model)
    log(f"[pk] corrected: {nsteps} steps in {t_corr:.3f} s; plain run on the 1024^3 mesh")
    d_plain, t_plain = timed(conf, None)
    log(f"[pk] plain 2x: {t_plain:.3f} s; plain run on the 2048^3 mesh")
    d_fine, t_fine = timed(conf4, None)
    log(f"[pk] plain 4x: {t_fine:.3f} s")
    k, pk_fine, _ = density_to_pk(d_fine, conf)
    _, r_c, _, pk_c, _, _ = density_to_cross_correlation(d_corr, d_fine, conf)
    _, r_p, _, pk_p, _, _ = density_to_cross_correlation(d_plain, d_fine, conf)
    k = k.cpu().numpy()
    sel = (k > 0) & (k <= 1.0)
    res = {"kmax_h_per_Mpc": 1.0, "shells": int(sel.sum()), "seed": seed, "kdk_steps": nsteps,
           "measured_on": "1024^3 mesh, CIC deconvolved", "weights": correction_meta(weights),
           "potential_channel": bool(model.potential)}
    for name, pk, r in (("corrected_2x", pk_c, r_c), ("plain_2x", pk_p, r_p)):
        e = abs(pk.cpu().numpy()[sel] / pk_fine.cpu().numpy()[sel] - 1)
        res[name] = {"pk_err_mean": round(float(e.mean()), 5), "pk_err_max": round(float(e.max()), 5),
                     "r_mean": round(float(r.cpu().numpy()[sel].mean()), 6)}
    res["wall_s"] = {"corrected_2x": round(t_corr, 3), "plain_2x": round(t_plain, 3), "plain_4x": round(t_fine, 3)}
    res["corrected_vs_plain_4x_wall"] = round(t_corr / t_fine, 4)
    return res


def network_run(args):
    """The learned correction's inference network alone on the full-scale workload's coarse
    mesh: the CIC density of the 512^3-particle 2LPT state (a = 1) on the 1024^3 mesh, then
    ``--steps`` evaluations of the three-layer BF16 network (statistics, stem from the density,
    hidden GELU conv, output layer fused with the residual).  Prints one JSON line with the
    per-kernel times from the library's HIP events (not the headline metric)."""
    import torch

    from pmpp_amd import Configuration, _lib
    from pmpp_amd._torch import stream_handle
    from pmpp_amd.corrections import init_mesh_cnn_potential_correction
    from pmpp_amd.corrections.mesh_cnn import evaluate_residual
    from pmpp_amd.corrections.training import density
    from pmpp_amd.initial_conditions import LCDM, particles_2lpt

    dev = torch.device("cuda", 0)
    ng, ratio, box = CONFIGS[args.config]
    cosmo = LCDM()
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ptcl = particles_2lpt(conf, cosmo, seed=0, jitter_seed=1, device=dev, a=1.0)
    rho = density(ptcl, cosmo.Omega_m, conf)
    del ptcl
    model = init_mesh_cnn_potential_correction(seed=0, depth=2, block="plain", device=dev, potential=args.potential)
    for _ in range(args.warmup):
        evaluate_residual(model, rho, 1.0, cosmo.Omega_m, conf=conf)
    torch.cuda.synchronize()
    stream = stream_handle(dev)
    cap = 16 * args.steps + 16
    _lib.call("pm_timing_enable", cap)
    _lib.call("pm_timing_mark", b"begin", stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        evaluate_residual(model, rho, 1.0, cosmo.Omega_m, conf=conf)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (ctypes.c_float * cap)()
    names = (ctypes.c_char_p * cap)()
    cnt = ctypes.c_int32(0)
    _lib.call("pm_timing_read", ms, names, cap, ctypes.byref(cnt))
    _lib.call("pm_timing_enable", 0)
    per = {}
    for i in range(cnt.value):
        k = names[i].decode()
        per[k] = per.get(k, 0.0) + ms[i] / args.steps
    nvox = conf.mesh_size
    print(json.dumps({
        "metric": "inference network time per force step (three-layer BF16 mesh CNN)",
        "ms_per_eval": (t1 - t0) * 1e3 / args.steps, "voxels": nvox, "steps": args.steps,
        "kernels_ms": {k: round(v, 4) for k, v in sorted(per.items(), key=lambda kv: -kv[1])},
        "config": {"workload": f"network on the {ng * ratio}^3 density of {ng}^3 particles, {box:g} Mpc/h"},
    }), flush=True)


def vjp_run(args):
    """The reverse of the force step (``pm_gravity_vjp``: forward meshes recomputed, the gather
    transpose, the spectral adjoint, the scatter adjoint) on the headline workload with a
    random acceleration cotangent, ``--steps`` times after ``--warmup``.  Prints one JSON line
    with the per-kernel times from the library's HIP events (not the headline metric)."""
    import torch

    from pmpp_amd import Configuration, _lib
    from pmpp_amd._torch import ensure_init, stream_handle
    from pmpp_amd.initial_conditions import LCDM, particles_2lpt
    from pmpp_amd.nbody.gravity import gravity_vjp

    dev = torch.device("cuda", 0)
    ng, ratio, box = CONFIGS[args.config]
    cosmo = LCDM()
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    ptcl = particles_2lpt(conf, cosmo, seed=0, jitter_seed=1, device=dev, a=1.0)
    ensure_init(dev)
    g = torch.Generator(device=dev).manual_seed(5)
    cot = torch.randn((len(ptcl), 3), device=dev, generator=g)
    for _ in range(args.warmup):
        gravity_vjp(ptcl, cosmo, conf, cot)
    torch.cuda.synchronize()
    stream = stream_handle(dev)
    cap = 64 * args.steps + 16
    _lib.call("pm_timing_enable", cap)
    _lib.call("pm_timing_mark", b"begin", stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gravity_vjp(ptcl, cosmo, conf, cot)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (ctypes.c_float * cap)()
    names = (ctypes.c_char_p * cap)()
    cnt = ctypes.c_int32(0)
    _lib.call("pm_timing_read", ms, names, cap, ctypes.byref(cnt))
    _lib.call("pm_timing_enable", 0)
    per = {}
    for i in range(cnt.value):
        k = names[i].decode()
        per[k] = per.get(k, 0.0) + ms[i] / args.steps
    print(json.dumps({
        "metric": "force-step VJP time (acc cotangent -> disp and Omega_m cotangents)",
        "ms_per_vjp": (t1 - t0) * 1e3 / args.steps, "particles": len(ptcl), "steps": args.steps,
        "kernels_ms": {k: round(v, 4) for k, v in sorted(per.items(), key=lambda kv: -kv[1])},
        "config": {"workload": f"PM force step VJP: {ng}^3 particles, {ng * ratio}^3 mesh, {box:g} Mpc/h"},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c5", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nbody-run", action="store_true",
                    help="time the full 20-step run (plain, reverse sweep, with the learned correction) and a "
                         "plain 2048^3-mesh force step instead of the force-step benchmark")
    ap.add_argument("--run-steps", type=int, default=20)
    ap.add_argument("--corrected-run", action="store_true",
                    help="only the corrected forward run of --nbody-run (for profiling its step)")
    ap.add_argument("--network", action="store_true",
                    help="time the learned correction's inference network on the workload's coarse mesh instead")
    ap.add_argument("--vjp", action="store_true",
                    help="time the force step's VJP on the workload instead")
    ap.add_argument("--train-correction", action="store_true",
                    help="train the learned correction on the 128^3 / 250 Mpc/h coarse (2x) / fine (4x) workload, "
                         "choose it on a validation seed and judge it on a held-out one, instead")
    ap.add_argument("--train-seeds", type=int, default=32)
    ap.add_argument("--train-iters", type=int, default=2000)
    ap.add_argument("--max-residual", default="0.5")
    ap.add_argument("--lr", default="1e-3")
    ap.add_argument("--symmetry", default="none", help="kernel symmetry of the trained network: none, cubic, or both")
    ap.add_argument("--e2e-iters", type=int, default=300)
    ap.add_argument("--e2e-seeds", type=int, default=16)
    ap.add_argument("--e2e-lr", type=float, default=1e-4)
    ap.add_argument("--e2e-validate-every", type=int, default=50)
    ap.add_argument("--r-weight", type=float, default=1.0)
    ap.add_argument("--weights", default=None,
                    help="a saved correction for the P(k) comparison and --nbody-run instead of the shipped one")
    ap.add_argument("--init-from", default=None,
                    help="with --train-correction: skip stage 1 and fine-tune these saved weights")
    ap.add_argument("--potential", action="store_true",
                    help="with --train-correction or --network: the network also reads the potential channel")
    ap.add_argument("--save", default=None,
                    help="with --train-correction: write the chosen weights to this path (the package ships "
                         "pmpp_amd/corrections/weights/mesh_cnn_2x_from_4x.npz)")
    ap.add_argument("--no-pk-error", action="store_true",
                    help="skip the corrected 1024^3 vs plain 2048^3 P(k) comparison of the default run")
    ap.add_argument("--pk-seed", type=int, default=0,
                    help="white-noise seed of that P(k) comparison (default 0)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (the known-size kernel that "
                         "calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU x-slab path is a "
              "design note, not built in this tree", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.nbody_run:
        nbody_run(args)
        return
    if args.corrected_run:
        corrected_run(args)
        return

    if args.network:
        network_run(args)
        return

    if args.vjp:
        vjp_run(args)
        return

    if args.train_correction:
        args.run_steps = args.run_steps or 20
        train_correction_run(args)
        return

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copied 3 x 1 GiB (read 1 GiB + write 1 GiB per copy)")
        return

    from pmpp_amd import Configuration, _lib
    from pmpp_amd._torch import ensure_init, ptr, stream_handle, workspace
    from pmpp_amd.initial_conditions import LCDM, particles_2lpt
    from pmpp_amd.nbody.gravity import gravity_desc

    dev = torch.device("cuda", 0)
    ng, ratio, box = CONFIGS[args.config]
    cosmo = LCDM()
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio)
    log(f"[bench] building 2LPT input: {ng}^3 particles, {ng * ratio}^3 mesh, {box:g} Mpc/h")
    ptcl = particles_2lpt(conf, cosmo, seed=0, jitter_seed=1, device=dev, a=1.0)
    torch.cuda.synchronize()
    ensure_init(dev)
    d = gravity_desc(conf, cosmo.Omega_m)
    n = len(ptcl)
    pm = ptcl.pmid.contiguous()
    dp = ptcl.disp.contiguous()
    nbytes = _lib.size_query("pm_gravity_workspace", d, n)
    ws = workspace(nbytes, dev)
    acc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    stream = stream_handle(dev)
    lib = _lib.load_library()
    log(f"[bench] workspace {nbytes / 2**30:.2f} GiB; warmup {args.warmup}")

    def step():
        _lib.check(lib.pm_gravity(ptr(pm), ptr(dp), None, n, d, ptr(acc), ptr(ws), nbytes, stream), "pm_gravity")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    max_events = args.steps * 48 + 16
    _lib.call("pm_timing_enable", max_events)
    _lib.call("pm_timing_mark", b"begin", stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (ctypes.c_float * max_events)()
    names = (ctypes.c_char_p * max_events)()
    cnt = ctypes.c_int32(0)
    _lib.call("pm_timing_read", ms, names, max_events, ctypes.byref(cnt))
    _lib.call("pm_timing_enable", 0)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    log(f"[bench] {ms_per_step:.3f} ms/step over {args.steps} steps")

    L = {"nx": conf.mesh_shape[0], "ny": conf.mesh_shape[1], "nz": conf.mesh_shape[2],
         "nreal": conf.mesh_size}
    L.update(spectral_band(conf))
    per = {}
    for i in range(cnt.value):
        k = names[i].decode()
        t, c = per.get(k, (0.0, 0))
        per[k] = (t + ms[i], c + 1)
    kernels = {}
    brick = tuple(conf.brick)
    for k, (t, c) in sorted(per.items(), key=lambda kv: -kv[1][0]):
        sb = kernel_step_bytes(k, L, n, brick)
        step_ms = t / args.steps
        kernels[k] = {"ms_per_step": step_ms, "launches_per_step": c / args.steps, "avg_ms": t / c,
                      "bytes_per_step": sb, "GBps": (sb / (step_ms * 1e-3) / 1e9) if sb else None}
        log(f"[bench]   {k:22s} {step_ms:8.3f} ms/step  x{c / args.steps:.0f}  avg {t / c:.3f} ms"
            + (f"  {sb / (step_ms * 1e-3) / 1e9:8.1f} GB/s" if sb else ""))
    dominant = next(iter(kernels))
    dk = kernels[dominant]
    per_launch = dk["bytes_per_step"] / dk["launches_per_step"]
    achieved = per_launch / (dk["avg_ms"] * 1e-3) / 1e9
    off_bytes = official_step_bytes(n, L)
    strict_bytes = strict_step_bytes(n, L)
    roofline = {
        "bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dominant, args.config),
        "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": round(dk["avg_ms"], 4),
        "step_official_bytes": off_bytes,
        "step_frac": round(off_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        "strict": {"step_bytes": strict_bytes,
                   "frac": round(strict_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)},
    }
    del ws, acc, pm, dp, ptcl
    torch.cuda.empty_cache()
    pk_err = None
    from pmpp_amd.corrections.mesh_cnn import TRAINED_DEFAULT, TRAINED_DIR
    if not args.no_pk_error and args.config == "c5":
        if (TRAINED_DIR / TRAINED_DEFAULT).exists():
            log("[bench] P(k) of the corrected 1024^3 run against the plain 2048^3 run")
            pk_err = pk_error_run(dev, seed=args.pk_seed, weights=args.weights)
        else:
            log("[bench] no trained correction weights shipped: P(k) comparison skipped")
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (float32 NumPy/SciPy port)")
        cpu = cpu_baseline(dev, cosmo.Omega_m)
    value = n / (ms_per_step * 1e-3)
    out = {
        "metric": "PM force steps/sec & particles/s (512^3 on 1024^3 mesh, fp32)",
        "value": value,
        "unit": "particles/s",
        "steps_per_s": 1e3 / ms_per_step,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (2LPT LCDM field from seeded white noise + lattice jitter)",
        "config": {"workload": f"PM force step: {ng}^3 particles, {ng * ratio}^3 mesh, {box:g} Mpc/h",
                   "particles": n, "mesh": list(conf.mesh_shape), "box": box, "brick": list(conf.brick)},
        "roofline": roofline,
        "pk_error_vs_4x_mesh": pk_err,
        "cpu_baseline": cpu,
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
