"""Training and use of the learned pressure-correction guess (mc_mpnn_* in
include/mcflow.h).  All network arithmetic runs in the HIP library (forward,
backward and Adam kernels); this module only drives it.

    python -m minicombust_app_amd.learned train --m 64 --pairs 2000 --steps 3000

Data: (pressure system, solution) pairs recorded by the library during the
pressure solves of a box-mesh run whose gas velocity is scaled by a seeded
uniform factor in [1 - 0.2, 1 + 0.2] per case (mc_options.inlet_scale,
mc_reset_fields between cases).  Training: Adam on single pairs in a seeded
random order.  The weights file is raw little-endian float32 in the
library's parameter order (mc_mpnn_nparams() values).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "mpnn_pressure_guess.f32")
# the training run that produced WEIGHTS: command, data, loss curve, held-out loss
RECORD = os.path.join(HERE, "weights", "mpnn_pressure_guess.train.json")

H, FIN, LAYERS = 64, 4, 3


def init_weights(seed=1234):
    """He-uniform weights for the ReLU layers, zero biases, small decoder and
    residual branches (parameter order of mc_mpnn.hip)."""
    rng = np.random.default_rng(seed)

    def he(k, m):
        lim = np.sqrt(6.0 / k)
        return rng.uniform(-lim, lim, size=(k, m))

    parts = [he(FIN, H), np.zeros(H)]
    for _ in range(LAYERS):
        parts += [he(H, H), he(H, H), rng.uniform(-1.0, 1.0, size=H), np.zeros(H), he(H, H), np.zeros(H),
                  he(2 * H, H), np.zeros(H), 0.1 * he(H, H), np.zeros(H)]
    parts += [rng.uniform(-0.1, 0.1, size=H), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def save_weights(w, path=WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.asarray(w, dtype="<f4").tofile(path)


def load_weights(path=WEIGHTS):
    return np.fromfile(path, dtype="<f4")


def generate(ctx, pairs, seed=20240910, perturb=0.2, iters_per_case=12, log=None):
    """Run perturbed cases on `ctx` until `pairs` pressure systems were stored."""
    rng = np.random.default_rng(seed)
    ctx.mpnn_collect(pairs)
    cases = 0
    while ctx.mpnn_dataset_size() < pairs:
        scale = 1.0 + rng.uniform(-perturb, perturb)
        ctx.set_options(inlet_scale=scale)
        ctx.reset_fields()
        for it in range(iters_per_case):
            ctx.simple_iteration(it % 3)
            if ctx.mpnn_dataset_size() >= pairs:
                break
        cases += 1
        if log:
            log("case %d: inlet scale %.4f, %d pairs" % (cases, scale, ctx.mpnn_dataset_size()))
    ctx.mpnn_collect(0)
    ctx.set_options(inlet_scale=1.0)
    ctx.reset_fields()
    return cases


def train(ctx, steps, lr=1e-3, seed=7, holdout=0.1, log=None, log_every=200):
    """Adam on single stored pairs; the last `holdout` fraction is held out.
    Returns (train losses sampled every log_every steps, held-out loss)."""
    n = ctx.mpnn_dataset_size()
    n_train = max(1, int(round(n * (1.0 - holdout))))
    rng = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        pair = int(rng.integers(0, n_train))
        want = (s % log_every == 0) or s == steps - 1
        lr_s = lr * (0.5 * (1 + np.cos(np.pi * s / steps)))   # cosine decay
        loss = ctx.mpnn_train_step(pair, max(lr_s, 1e-6), want_loss=want)
        if want:
            losses.append((s, loss))
            if log:
                log("step %d loss %.5f" % (s, loss))
    held = [ctx.mpnn_grad(p)[1] for p in range(n_train, n)]
    return losses, (float(np.mean(held)) if held else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--m", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20240910)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--record", default=None, help="training record (JSON); default: next to --out")
    args = ap.parse_args()
    from minicombust_app_amd import mcflow as mc

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = mc.Context(args.m, mode=mc.MODE_SIMPLE, psolve=mc.PSOLVE_FP64, p_rtol=1e-8)
    ctx.mpnn_load(init_weights())
    t0 = time.time()
    cases = generate(ctx, args.pairs, seed=args.seed, log=log)
    t1 = time.time()
    n = ctx.mpnn_dataset_size()
    n_train = max(1, int(round(n * 0.9)))
    held0 = float(np.mean([ctx.mpnn_grad(p)[1] for p in range(n_train, n)])) if n > n_train else None
    losses, held = train(ctx, args.steps, lr=args.lr, log=log)
    t2 = time.time()
    w = ctx.mpnn_save()
    save_weights(w, args.out)
    rec = {"command": "python -m minicombust_app_amd.learned " + " ".join(sys.argv[1:]),
           "m": args.m, "cells": 2 * args.m ** 3, "pairs": n, "train_pairs": n_train, "heldout_pairs": n - n_train,
           "cases": cases, "inlet_scale": "1 + U(-0.2, 0.2), seed %d" % args.seed, "steps": args.steps,
           "lr": args.lr, "schedule": "Adam, cosine decay, one pair per step",
           "generate_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
           "heldout_loss_initial": held0, "heldout_loss": held,
           "loss_curve": [[s, round(l, 6)] for s, l in losses],
           "weights": os.path.relpath(args.out, os.path.dirname(HERE)), "nparams": int(w.size)}
    record = args.record or os.path.splitext(args.out)[0] + ".train.json"
    with open(record, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(json.dumps({k: v for k, v in rec.items() if k != "loss_curve"}))
    ctx.close()


if __name__ == "__main__":
    main()
