"""Trains the learned pose proposer on the GPU with the library's own forward
and backward kernels (cvo_model_train_step) and writes its weights.

Data: seeded pairs of the planes/spheres generator (two independent samples
of a random scene, the second moved by a uniform random rotation and 2 m), 2048
points per cloud.  Labels (--label cvo, the default): the converged poses of
a batched multi-start CVO registration of each training pair (super-Fibonacci
start rotations with centroid-matching translations, the plain grid's
construction, at the global-registration parameters); the best-cos converged
pose is the label.  A pair whose converged pose lies farther than 0.05 rad /
0.1 m from the generating transform (CVO settled in another basin) is dropped.
--label generator uses the generating transform itself.

Labelling takes ~17 ms per pair, so a large set can be labelled in parts:
--label-only FILE labels the --pairs pairs of --seed and writes (first seed,
count, kept indices, labels) to FILE without training; --labels FILE ...
regenerates those files' pairs from their seeds and adds their kept pairs to
the training set.

    python -m rkhs_ba_amd.train_proposer --pairs 50000 --out rkhs_ba_amd/weights/proposer_k8.npy
    python -m rkhs_ba_amd.train_proposer --seed 1 --pairs 42000 --label-only labels_1.npz
    python -m rkhs_ba_amd.train_proposer --seed 2 --pairs 42000 --labels labels_1.npz
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


def cvo_labels(a, b, t, starts, max_iter, score_ell=0.3, accept_rad=0.05, accept_m=0.1):
    """Labels from converged multi-start CVO: (kept index, labels (n, 7),
    seconds).  The label is the source-to-target transform (align-init
    convention) of the best-cos start after max_iter iterations."""
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    from rkhs_ba_amd.proposer import grid_starts, rot_to_quat
    params = read_params_yaml(os.path.join(ROOT, "rkhs_ba_amd", "configs", "multistart_20k.yaml"))
    params.MAX_ITER = max_iter
    cvo = CvoGPU(params=params)
    keep, labels = [], []
    t0 = time.time()
    try:
        for i in range(a.shape[0]):
            src = synth.Cloud(np.ascontiguousarray(a[i, :, :3]), np.ascontiguousarray(a[i, :, 3:8]))
            tgt = synth.Cloud(np.ascontiguousarray(b[i, :, :3]), np.ascontiguousarray(b[i, :, 3:8]))
            r = cvo.align_multistart(src, tgt, grid_starts(src, tgt, starts), score_ell, mode=0)
            T = np.linalg.inv(np.asarray(r["T_best"], np.float64))  # align output -> init convention
            Tg = np.eye(4)
            Tg[:3, :3] = synth.quat_to_rot(t[i, :4].astype(np.float64))
            Tg[:3, 3] = t[i, 4:7]
            E = np.linalg.inv(Tg) @ T
            ang = float(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))
            if ang <= accept_rad and np.linalg.norm(E[:3, 3]) <= accept_m:
                keep.append(i)
                labels.append(np.concatenate([rot_to_quat(T[:3, :3]), T[:3, 3]]).astype(np.float32))
            if (i + 1) % 2000 == 0:
                print(f"[train] labelled {i + 1}/{a.shape[0]}, kept {len(keep)}, "
                      f"{(time.time() - t0) / (i + 1) * 1e3:.1f} ms per pair", flush=True)
    finally:
        cvo.close()
    return np.array(keep, np.int64), np.array(labels, np.float32).reshape(-1, 7), time.time() - t0


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
    ap.add_argument("--label", choices=("cvo", "generator"), default="cvo")
    ap.add_argument("--label-starts", type=int, default=24, help="multi-start CVO starts per training pair")
    ap.add_argument("--label-iters", type=int, default=300, help="CVO iterations per start")
    ap.add_argument("--label-only", default=None, help="label the pairs, write them to this .npz and stop")
    ap.add_argument("--labels", nargs="*", default=[], help=".npz files of --label-only runs to train on as well")
    args = ap.parse_args()
    if (args.label_only or args.labels) and args.label != "cvo":
        ap.error("--label-only and --labels need --label cvo")
    from rkhs_ba_amd import CvoGPU
    from rkhs_ba_amd.proposer import PoseProposer
    t0 = time.time()
    print(f"[train] generating {args.pairs} + {args.val} pairs on {args.workers} workers", flush=True)
    first_seed = 1_000_000 * (args.seed + 1)
    a, b, t = dataset(first_seed, args.pairs, args.npts, args.workers)
    va, vb, vt = dataset(7_000_000 + args.seed, args.val, args.npts, args.workers)
    print(f"[train] data ready in {time.time() - t0:.1f} s", flush=True)
    label_info = {"label": args.label}
    if args.label == "cvo":
        keep, lab, secs = cvo_labels(a, b, t, args.label_starts, args.label_iters)
        label_info.update(starts=args.label_starts, iterations=args.label_iters, kept=int(keep.size),
                          of=int(a.shape[0]), seconds=secs)
        print(f"[train] CVO labels: kept {keep.size} of {a.shape[0]} pairs in {secs:.1f} s", flush=True)
        if args.label_only:
            np.savez(args.label_only, first_seed=np.int64(first_seed), n=np.int64(a.shape[0]), keep=keep,
                     labels=lab, starts=np.int64(args.label_starts), iterations=np.int64(args.label_iters),
                     seconds=np.float64(secs))
            print(f"[train] wrote {args.label_only}", flush=True)
            return
        a, b, t = a[keep], b[keep], lab
        parts = [dict(first_seed=first_seed, kept=int(keep.size), of=int(args.pairs))]
        for path in args.labels:
            z = np.load(path, allow_pickle=False)
            if int(z["starts"]) != args.label_starts or int(z["iterations"]) != args.label_iters:
                raise SystemExit(f"{path}: labelled with other CVO settings")
            za, zb, _ = dataset(int(z["first_seed"]), int(z["n"]), args.npts, args.workers)
            k = z["keep"]
            a, b, t = np.concatenate([a, za[k]]), np.concatenate([b, zb[k]]), np.concatenate([t, z["labels"]])
            parts.append(dict(first_seed=int(z["first_seed"]), kept=int(k.size), of=int(z["n"])))
            label_info["kept"] += int(k.size)
            label_info["of"] += int(z["n"])
            label_info["seconds"] += float(z["seconds"])
            del za, zb
            print(f"[train] {path}: {k.size} of {int(z['n'])} pairs", flush=True)
        label_info["parts"] = parts
        args.pairs = int(a.shape[0])
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
               "labels": label_info,
               "train_seconds": time.time() - t_train, "ms_per_step": (time.time() - t_train) / max(1, step) * 1e3,
               "history": history}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
