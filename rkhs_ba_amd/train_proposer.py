"""Trains the learned pose proposer on the GPU with the library's own forward
and backward kernels (cvo_model_train_step) and writes its weights.

Data: seeded pairs of the planes/spheres generator (two independent samples
of a random scene, the second moved by a uniform random rotation and 2 m), 2048
points per cloud; the label is the generating source-to-target transform.

    python -m rkhs_ba_amd.train_proposer --pairs 50000 --out rkhs_ba_amd/weights/proposer_k8.npy
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _gen(args):
    seeds, npts = args
    from rkhs_ba_amd import proposer
    return proposer.make_batch(seeds, npts)


def dataset(first_seed, n, npts, workers):
    chunks = [(list(range(s, min(s + 256, first_seed + n))), npts) for s in range(first_seed, first_seed + n, 256)]
    with mp.get_context("spawn").Pool(workers) as pool:
        parts = pool.map(_gen, chunks)
    return (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
            np.concatenate([p[2] for p in parts]))


def quat_angle_deg(q, qs):
    d = np.abs(np.sum(q * qs, axis=-1))
    return np.degrees(2 * np.arccos(np.clip(d, -1, 1)))


def evaluate(prop, a, b, t, batch=256):
    best, top = [], []
    for i in range(0, a.shape[0], batch):
        qt, conf, _ = prop.propose(a[i:i + batch], b[i:i + batch])
        ang = quat_angle_deg(qt[:, :, :4], t[i:i + batch, None, :4])
        best.append(ang.min(axis=1))
        top.append(ang[np.arange(ang.shape[0]), conf.argmax(axis=1)])
    best, top = np.concatenate(best), np.concatenate(top)
    return {"median_best_deg": float(np.median(best)), "median_top_conf_deg": float(np.median(top)),
            "best_within_30deg": float(np.mean(best < 30)), "best_within_45deg": float(np.mean(best < 45)),
            "best_within_60deg": float(np.mean(best < 60))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=50000)
    ap.add_argument("--val", type=int, default=512)
    ap.add_argument("--npts", type=int, default=2048)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--time-limit", type=float, default=600.0, help="seconds of training")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(ROOT, "rkhs_ba_amd", "weights", "proposer_k8.npy"))
    args = ap.parse_args()
    from rkhs_ba_amd import CvoGPU
    from rkhs_ba_amd.proposer import PoseProposer
    t0 = time.time()
    print(f"[train] generating {args.pairs} + {args.val} pairs on {args.workers} workers", flush=True)
    a, b, t = dataset(1_000_000 * (args.seed + 1), args.pairs, args.npts, args.workers)
    va, vb, vt = dataset(7_000_000 + args.seed, args.val, args.npts, args.workers)
    print(f"[train] data ready in {time.time() - t0:.1f} s", flush=True)
    cvo = CvoGPU()
    prop = PoseProposer(cvo, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    steps_per_epoch = args.pairs // args.batch
    total_steps = steps_per_epoch * args.epochs
    history = []
    step = 0
    t_train = time.time()
    stop = False
    print(f"[train] initial {evaluate(prop, va, vb, vt)}", flush=True)
    for ep in range(args.epochs):
        perm = rng.permutation(args.pairs)
        losses = []
        for s in range(steps_per_epoch):
            idx = np.sort(perm[s * args.batch:(s + 1) * args.batch])
            lr = args.lr * 0.5 * (1 + np.cos(np.pi * step / max(1, total_steps)))
            losses.append(prop.train_step(a[idx], b[idx], t[idx], float(lr)))
            step += 1
            if step % 100 == 0:
                print(f"[train] step {step} epoch {ep} loss {np.mean(losses[-100:]):.4f} "
                      f"{(time.time() - t_train) / step * 1e3:.1f} ms/step", flush=True)
            if time.time() - t_train > args.time_limit:
                stop = True
                break
        ev = evaluate(prop, va, vb, vt)
        history.append({"epoch": ep, "steps": step, "loss": float(np.mean(losses)), **ev})
        print(f"[train] epoch {ep} {history[-1]}", flush=True)
        if stop:
            break
    prop.save(args.out)
    summary = {"pairs": args.pairs, "npts": args.npts, "batch": args.batch, "lr": args.lr, "steps": step,
               "train_seconds": time.time() - t_train, "ms_per_step": (time.time() - t_train) / max(1, step) * 1e3,
               "history": history}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
