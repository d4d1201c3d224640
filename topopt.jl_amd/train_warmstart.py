#!/usr/bin/env python3
"""Train the learned warm-start network on (x~, u) pairs from SIMP runs.

Dataset: SIMP runs of the cantilever on an nx x ny x nz grid (default
64 x 32 x 32) with the volume fraction drawn from U[0.25, 0.6] and the unit
-y point load at a node of the free face drawn uniformly, one seed per run;
the filtered densities and converged displacements are kept every
`--keep-every` iterations (and at the first).  All solves, the dataset and
the training (BF16 MFMA convolutions, Adam) run on the GPU.

Writes the weights (FP32) with the training record to --out (.npz).
Usage: python topopt.jl_amd/train_warmstart.py --pairs 2000 --out w.npz
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _pkg  # noqa: E402


def log(*a):
    print(*a, flush=True)


def bench_training(NN, nels, batch, steps, lr, seed):
    """Training throughput on synthetic (x~, f, u) samples: densities
    U[0.05, 1], a free-face point load and a smooth displacement field."""
    import json
    g = np.random.default_rng(seed)
    net = NN.WarmStartNet(nels, max_batch=batch, max_samples=2 * batch, seed=seed)
    nx, ny, nz = nels
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                             indexing="ij")
    xi = ii.ravel() / nx
    for i in range(2 * batch):
        u = np.zeros((net.nnodes, 3))
        u[:, 1] = -(xi ** 2) * (1 + 0.1 * i)
        net.add_sample(g.uniform(0.05, 1.0, net.ncells),
                       NN.cantilever_load(nels, ny // 2, nz // 2), u.ravel())
    net.train(2, batch, lr, seed=seed)  # warm-up
    t0 = time.perf_counter()
    net.train(steps, batch, lr, seed=seed + 1)
    dt = (time.perf_counter() - t0) / steps
    fwd = NN.conv_flops(nels, batch)
    flops = 3 * fwd  # forward + input gradient + weight gradient
    out = {"nels": list(nels), "batch": batch, "steps": steps, "ms_per_step": dt * 1e3,
           "params": net.nparams, "train_gflop_per_step": flops / 1e9,
           "tflops": flops / dt / 1e12, "bf16_peak_tflops": 2500.0,
           "mfma_frac": flops / dt / 2.5e15}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nels", default="64,32,32")
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--iters", type=int, default=30, help="SIMP iterations per run")
    ap.add_argument("--keep-every", type=int, default=5)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--holdout", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "warmstart_64x32x32.npz"))
    ap.add_argument("--init", default=None, help="start from these weights (fine-tuning)")
    ap.add_argument("--eval-runs", type=int, default=4,
                    help="held-out SIMP runs (seeds after the training runs) compared with "
                         "zero / previous-iterate / learned warm starts")
    ap.add_argument("--eval-iters", type=int, default=20)
    ap.add_argument("--rmin", type=float, default=1.5, help="density filter radius")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "mixed_mg"],
                    help="solver of the data-generating SIMP runs and of the evaluation")
    ap.add_argument("--volfrac-range", default="0.25,0.6",
                    help="volume fractions of the runs, drawn uniformly from this range")
    ap.add_argument("--bench-steps", type=int, default=0,
                    help="only time this many training steps on synthetic samples (no solver) "
                         "and report the convolutions' FLOP rate")
    args = ap.parse_args()
    _pkg.load_package()
    import topopt_jl_amd.nn as NN
    nels = tuple(int(v) for v in args.nels.split(","))
    if args.bench_steps:
        bench_training(NN, nels, args.batch, args.bench_steps, args.lr, args.seed)
        return
    ctx, fixed = NN.make_context(nels)
    prec = _pkg.load_package()._lib.PRECISIONS[args.precision]
    vlo, vhi = (float(v) for v in args.volfrac_range.split(","))
    net = NN.WarmStartNet(nels, max_batch=args.batch, max_samples=args.pairs, seed=args.seed,
                          fixed=fixed)
    if args.init:
        net.load(args.init)
        log(f"initial weights: {args.init}")
    log(f"network: {net.nparams} parameters on {nels}")
    t0 = time.perf_counter()
    run = 0
    pcg = []
    while net.nsamples < args.pairs:
        rng = np.random.default_rng(args.seed * 100003 + run)
        spec = NN.SimpRunSpec(volfrac=float(rng.uniform(vlo, vhi)),
                              load_j=int(rng.integers(0, nels[1] + 1)),
                              load_k=int(rng.integers(0, nels[2] + 1)), seed=run)
        for xp, f, u, its in NN.simp_pairs(ctx, nels, spec, args.iters, args.keep_every,
                                           rmin=args.rmin, precision=prec):
            if net.nsamples >= args.pairs:
                break
            net.add_sample(xp, f, u)
            pcg.append(its)
        run += 1
        if run % 20 == 0:
            log(f"  runs {run}: {net.nsamples} pairs, {time.perf_counter() - t0:.1f}s")
    t_data = time.perf_counter() - t0
    log(f"dataset: {net.nsamples} pairs from {run} runs in {t_data:.1f}s")
    n_hold = max(1, int(args.holdout * net.nsamples))
    n_train = net.nsamples - n_hold  # the last runs' pairs are held out
    t1 = time.perf_counter()
    losses = []
    chunk = 100
    for s0 in range(0, args.steps, chunk):
        lr = args.lr if s0 < args.steps // 2 else args.lr * 0.3
        if s0 >= 3 * args.steps // 4:
            lr = args.lr * 0.1
        h = net.train(min(chunk, args.steps - s0), args.batch, lr, seed=args.seed + s0,
                      first=0, count=n_train)
        losses.extend(h.tolist())
        hold = net.eval(n_train, n_hold)
        log(f"  step {s0 + len(h)}: train {np.mean(h[-20:]):.4f} holdout {hold:.4f} "
            f"({time.perf_counter() - t1:.1f}s)")
    t_train = time.perf_counter() - t1
    params = net.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    hold = net.eval(n_train, n_hold)
    # held-out runs: PCG iterations with each warm start
    evals = []
    for e in range(args.eval_runs):
        r = run + 1000 + e
        rng = np.random.default_rng(args.seed * 100003 + r)
        spec = NN.SimpRunSpec(volfrac=float(rng.uniform(vlo, vhi)),
                              load_j=int(rng.integers(0, nels[1] + 1)),
                              load_k=int(rng.integers(0, nels[2] + 1)), seed=r)
        res = NN.compare_warm_starts(ctx, net, nels, spec, args.eval_iters, rmin=args.rmin,
                                     precision=prec)
        row = [res[k][0] for k in ("zero", "previous", "learned")]
        evals.append(np.stack(row))
        log(f"  held-out run {r} (volfrac {spec.volfrac:.3f}, load j={spec.load_j} "
            f"k={spec.load_k}): PCG total zero {row[0].sum()} previous {row[1].sum()} "
            f"learned {row[2].sum()}; first solve zero {row[0][0]} learned {row[2][0]}; "
            f"max rel obj diff {np.max(np.abs(res['learned'][1] / res['previous'][1] - 1)):.2e}")
    evals = np.array(evals, np.int64).reshape(len(evals), 3, -1)
    np.savez_compressed(args.out, params=params, nels=np.array(nels), loss=np.array(losses),
                        holdout_loss=hold, pairs=net.nsamples, runs=run,
                        data_seconds=t_data, train_seconds=t_train,
                        mean_pcg_iters=float(np.mean(pcg)) if pcg else 0.0,
                        eval_pcg_iters=evals)
    log(f"saved {args.out}: holdout loss {hold:.4f}, data {t_data:.1f}s, train {t_train:.1f}s")


if __name__ == "__main__":
    main()
