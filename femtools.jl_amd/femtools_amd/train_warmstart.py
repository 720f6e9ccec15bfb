"""Trains the learned warm start on the GPU and writes its weights.

    python -m femtools_amd.train_warmstart --samples 2000 --epochs 30 --out weights.npz

1. Data: DR solutions of the viscous inclusion generator (warmstart.
   random_inclusions: 1..10 circles, r in U[0.02, 0.1], log10 eta in
   U[-2, 3], eta = 1 matrix, pure shear, incompressible) on 128 x 128-cell T7
   meshes, solved by this library (FP32-bulk DR, FP64 residuals) to the given
   tolerance in batched calls. Seeds 0..samples-1 train, seeds from 10^7
   validate, seeds from 2 10^7 evaluate the gain.
2. Training: the device network's own forward/backward kernels, Adam, cosine
   learning-rate decay, mean-squared error on the scaled targets.
3. Output: BF16 weights and FP32 biases (.npz) and a JSON summary on stdout,
   with the iterations and wall time to 1e-8 of the first two steps of the
   visco-elasto-plastic series generator from the zero, pure-shear,
   previous-step and learned guesses (warmstart.evaluate_gain).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

T0 = time.perf_counter()


def log(msg):
    print(f"# [{time.perf_counter() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--max-iterations", type=int, default=100_000)
    ap.add_argument("--batch-solves", type=int, default=64, help="problems per batched solve call")
    ap.add_argument("--gain", default="256:2,512:1", help="n:count pairs of the evaluation of the learned guess")
    ap.add_argument("--data-budget-s", type=float, default=1e9, help="stop generating data after this long")
    ap.add_argument("--out", default=None)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)

    import femtools_amd as fa
    from femtools_amd import warmstart as WS
    out = args.out or WS.WEIGHTS
    seeds = list(range(args.samples))
    vseeds = list(range(10_000_000, 10_000_000 + args.val))
    t0 = time.perf_counter()
    Xv, Tv, sv = WS.generate_dataset(vseeds, tolerance=args.tolerance, max_iterations=args.max_iterations,
                                     batch=args.batch_solves)
    parts, stats = [], []
    chunk = 4 * args.batch_solves
    for i0 in range(0, len(seeds), chunk):
        if time.perf_counter() - t0 > args.data_budget_s:
            log(f"data budget spent after {i0} samples")
            break
        X, T, st = WS.generate_dataset(seeds[i0:i0 + chunk], tolerance=args.tolerance,
                                       max_iterations=args.max_iterations, batch=args.batch_solves)
        parts.append((X, T))
        stats += st
        log(f"data {i0 + len(st)}/{len(seeds)}: {sum(s['converged'] for s in stats)} converged")
    X = np.concatenate([p[0] for p in parts])
    T = np.concatenate([p[1] for p in parts])
    t_data = time.perf_counter() - t0
    net = WS.Network(fa.default_context(), max_batch=args.batch)
    net.set_params(WS.init_params(args.seed))
    v0 = WS.evaluate(net, Xv, Tv)
    zero = float((Tv ** 2).mean())
    log(f"{len(X)} samples in {t_data:.0f}s; validation loss at init {v0:.4e}, of the zero prediction {zero:.4e}")
    t1 = time.perf_counter()
    hist = WS.train(net, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log,
                    val=(Xv, Tv))
    t_train = time.perf_counter() - t1
    params = net.get_params()
    meta = {"samples": len(X), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "tolerance": args.tolerance, "val_loss": hist[-1].get("val_loss"), "zero_loss": zero}
    WS.save(params, out, meta)
    gain = {}
    net1 = WS.Network(fa.default_context(), max_batch=1)
    net1.set_params(params)
    for item in filter(None, args.gain.split(",")):
        n, cnt = (int(v) for v in item.split(":"))
        gain[n] = WS.evaluate_gain(net1, n, list(range(20_000_000, 20_000_000 + cnt)), log=log)
    summary = {"samples": len(X), "converged": int(sum(s["converged"] for s in stats)),
               "iterations_mean": float(np.mean([s["iter"] for s in stats])), "data_s": round(t_data, 1),
               "train_s": round(t_train, 1), "epochs": args.epochs, "final_loss": hist[-1]["loss"],
               "val_loss": hist[-1].get("val_loss"), "val_loss_zero_prediction": zero, "weights": out,
               "history": hist, "gain": gain}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
