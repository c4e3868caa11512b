"""The learned-correction study: train on coarse/fine pairs the method generates and judge the
correction by P(k) and cross-correlation against a held-out fine run.

    python -m pmpp_amd.corrections.study [--particles 128] [--box 256] [--coarse 2] [--fine 4]
                                         [--seeds 32] [--iters 800] [--steps 10] [--lr 1e-4]

Defaults: 128^3 particles in a 256 Mpc/h box, a coarse mesh of 256^3 (``--coarse 2``) and fine
runs on a mesh 4x finer per axis (1024^3), 2LPT at a = 1/64 then ``--steps`` KDK steps to a = 1,
training pairs from every other step of ``--seeds`` fine runs, and seed 1000 held out; the
network is the three-layer one (``--depth 2 --block plain``: stem, one hidden GELU conv,
output).  Prints one JSON line: the relative P(k) error against the fine run over the band
0.3 k_Nyq < k < k_Nyq (k_Nyq the particle Nyquist) and up to k = 1 h/Mpc (mean and maximum),
the mean cross-correlation in the band, for the plain and the corrected coarse runs, the
training loss before and after, and the inference network's time per evaluation with its
throughput from HIP events (useful and MFMA-issued TFLOP/s, against the 2.5 PFLOP/s dense BF16
peak).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time

import numpy as np
import torch

from .. import _lib
from .._torch import stream_handle
from ..core.configuration import Configuration
from ..cosmology import LCDM
from ..nbody.gravity import gravity
from .mesh_cnn import CHANNELS, TAPS, evaluate_residual, init_mesh_cnn_potential_correction
from .training import build_correction_optimizer, compare_runs, density, fine_run_pairs, force_loss, train_step

BF16_PEAK_TFLOPS = 2500.0


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def conv_throughput(model, rho, a, omega_m, reps=5):
    """(TFLOP/s, ms) of the forward convolutions on ``rho``'s mesh, from the library's own
    per-kernel HIP events."""
    dev = rho.device
    stream = stream_handle(dev)
    evaluate_residual(model, rho, a, omega_m)  # warm
    torch.cuda.synchronize()
    cap = 64 * reps + 16
    _lib.call("pm_timing_enable", cap)
    _lib.call("pm_timing_mark", b"begin", stream)
    for _ in range(reps):
        evaluate_residual(model, rho, a, omega_m)
    torch.cuda.synchronize()
    ms = (ctypes.c_float * cap)()
    names = (ctypes.c_char_p * cap)()
    cnt = ctypes.c_int32(0)
    _lib.call("pm_timing_read", ms, names, cap, ctypes.byref(cnt))
    _lib.call("pm_timing_enable", 0)
    conv = (b"k_conv_fwd", b"k_conv_stem", b"k_conv_out_residual")
    t = sum(ms[i] for i in range(cnt.value) if names[i] in conv) / reps
    n = rho.numel()
    # useful: the stem's one live spatial channel (its constant channels fold into the bias),
    # 16 -> 16 hidden layers, the output layer's one channel; issued: what the MFMAs compute
    # (the stem's 27 taps in one K = 32 step, 15 steps of 16 x 16 x 32 per hidden / output tile)
    useful = 2.0 * TAPS * CHANNELS * n  # stem
    issued = 2.0 * 32 * CHANNELS * n
    for layer in model.layers[1:]:
        cout = 1 if layer.name == "out" else CHANNELS
        useful += 2.0 * TAPS * CHANNELS * cout * n
        issued += 2.0 * 15 * 32 * CHANNELS * n
    return useful / (t * 1e-3) / 1e12, issued / (t * 1e-3) / 1e12, t


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--particles", type=int, default=128)
    ap.add_argument("--box", type=float, default=256.0)
    ap.add_argument("--coarse", type=int, default=2)
    ap.add_argument("--fine", type=int, default=4)
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--iters", type=int, default=800)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--lr", default="1e-4", help="learning rate, or a comma-separated list to compare")
    ap.add_argument("--max-residual", type=float, default=0.1)
    ap.add_argument("--depth", type=int, default=2)
    ap.add_argument("--block", default="plain", choices=["plain", "residual"])
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    cosmo = LCDM()
    a0 = 1 / 64
    conf = Configuration(args.box / args.particles, (args.particles,) * 3, mesh_shape=args.coarse, a_start=a0,
                         a_nbody_maxstep=(1 - a0) / args.steps)
    conf_fine = conf.replace(mesh_shape=args.coarse * args.fine)
    t0 = time.perf_counter()
    pairs = fine_run_pairs(cosmo, conf, conf_fine, seeds=range(args.seeds), every=2, device=dev)
    torch.cuda.synchronize()
    _log(f"[study] {len(pairs)} pairs from {args.seeds} fine runs in {time.perf_counter() - t0:.1f} s")
    def pair_loss(corr, sample):
        tot = 0.0
        for a, p, tgt in sample:
            acc = gravity(a, p, cosmo.Omega_m, conf, correction=corr)
            tot += float(force_loss(acc, tgt, want_grad=False)[0][0])
        return tot / len(sample)

    sample = pairs[:: max(1, len(pairs) // 16)]
    loss_plain = pair_loss(None, sample)
    results = []
    model = None
    for lr in [float(v) for v in str(args.lr).split(",")]:
        model = init_mesh_cnn_potential_correction(seed=0, depth=args.depth, max_residual=args.max_residual,
                                                   device=dev, block=args.block)
        opt = build_correction_optimizer(lr)
        st = opt.init(model.params)
        rng = np.random.default_rng(0)
        t0 = time.perf_counter()
        grads = torch.zeros_like(model.params)
        curve = []
        for it in range(args.iters):
            a, p, tgt = pairs[int(rng.integers(len(pairs)))]
            train_step(model, opt, st, p, a, cosmo, conf, tgt, grads=grads)
            if (it + 1) % 100 == 0:
                curve.append(round(pair_loss(model, sample), 6))
                _log(f"[study] lr {lr:g} iteration {it + 1}: held-in loss {curve[-1]:.6f} (plain {loss_plain:.6f})")
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        res = compare_runs(model, cosmo, conf, conf_fine, seed=1000, device=dev)
        k = res["k"].cpu().numpy()
        kn = np.pi / conf.ptcl_spacing
        band = (k > 0.3 * kn) & (k < kn)
        fine = res["pk_fine"].cpu().numpy()[band]
        k1 = (k > 0) & (k <= 1.0)
        e1p = np.abs(res["pk_plain"].cpu().numpy()[k1] / res["pk_fine"].cpu().numpy()[k1] - 1)
        e1c = np.abs(res["pk_corrected"].cpu().numpy()[k1] / res["pk_fine"].cpu().numpy()[k1] - 1)
        results.append({
            "lr": lr, "train_s": round(t_train, 2), "loss_curve": curve,
            "force_loss_plain": loss_plain, "force_loss_corrected": pair_loss(model, sample),
            "pk_err_plain": float(np.abs(res["pk_plain"].cpu().numpy()[band] / fine - 1).mean()),
            "pk_err_corrected": float(np.abs(res["pk_corrected"].cpu().numpy()[band] / fine - 1).mean()),
            "pk_err_k1_plain": {"mean": float(e1p.mean()), "max": float(e1p.max())},
            "pk_err_k1_corrected": {"mean": float(e1c.mean()), "max": float(e1c.max())},
            "r_plain": float(res["r_plain"].cpu().numpy()[band].mean()),
            "r_corrected": float(res["r_corrected"].cpu().numpy()[band].mean()),
        })
        _log(f"[study] lr {lr:g}: {json.dumps(results[-1])}")
    out = {"particles": args.particles, "coarse_mesh": conf.mesh_shape[0], "fine_mesh": conf_fine.mesh_shape[0],
           "box": args.box, "seeds": args.seeds, "pairs": len(pairs), "iters": args.iters,
           "max_residual": args.max_residual, "depth": args.depth, "block": args.block,
           "conv_layers": len(model.layers) if model is not None else None, "runs": results}
    rho = density(pairs[-1][1], cosmo.Omega_m, conf)
    useful, issued, ms = conv_throughput(model, rho, 1.0, cosmo.Omega_m)  # the last model's layers
    out["conv_fwd_ms"] = round(ms, 4)
    out["conv_fwd_tflops_useful"] = round(useful, 1)
    out["conv_fwd_tflops_issued"] = round(issued, 1)
    out["conv_mfma_frac"] = round(issued / BF16_PEAK_TFLOPS, 4)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
