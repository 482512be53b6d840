"""Band predictor of the learned-band form: training on the GPU and model files.

The predictor is an MLP 24 -> 64 -> 64 -> 5 over per-task features; its
classes are band-following frames of 24, 40, 80 and 144 columns and "plain"
(the task's own band in the plain kernel).  Labels are the smallest frame that
holds the task's whole run (computed on the GPU by bsx_band_labels_device), so
a task the predictor sends to too small a frame costs a failed frame attempt
plus the plain run, and one sent to too large a frame costs the difference in
frame width.  Training,
inference and labelling all run through libbsx.so (hand-written HIP kernels);
this module only moves arrays and logs.

    python -m bsbolt_amd.band train --workload cfg2 --out bsbolt_amd/band_model.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import bsx

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "band_model.npy")
SHAPES = (("W1", (64, 24)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (5, 64)), ("b3", (5,)))


def init_params(seed=0):
    """Uniform He initialisation of the weights, zero biases (FP32, 6085 values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, s in SHAPES:
        if k.startswith("b"):
            parts.append(np.zeros(s))
        else:
            lim = np.sqrt(6.0 / s[1])
            parts.append(rng.uniform(-lim, lim, s))
    p = np.concatenate([x.reshape(-1) for x in parts]).astype(np.float32)
    assert p.size == bsx.BAND_PARAMS
    return p


def save_model(path, params, info=None):
    np.save(path, np.asarray(params, np.float32), allow_pickle=False)
    if info is not None:
        with open(os.path.splitext(path)[0] + ".json", "w") as f:
            json.dump(info, f, indent=1, sort_keys=True)


def load_model(path=DEFAULT_MODEL):
    p = np.load(path, allow_pickle=False)
    if p.shape != (bsx.BAND_PARAMS,) or p.dtype != np.float32:
        raise ValueError(f"{path}: not a band model ({p.shape}, {p.dtype})")
    return p


def features_and_labels(eng, batch, scoring=None):
    """(features uint16 [n, 32], labels uint8 [n]) of a host batch, on the GPU."""
    keep, dev = eng.tasks_to_device(batch)
    dev["n"] = len(batch)
    f = eng.empty((len(batch), bsx.BAND_FEATURE_STRIDE), np.uint16)
    lab = eng.empty((len(batch),), np.uint8)
    eng.band_features_device(dev, f.ptr, scoring)
    eng.band_labels_device(dev, lab.ptr, scoring)
    eng.sync()
    return f.numpy(), lab.numpy()


def train(eng, X, y, steps=1500, batch=4096, lr=3e-3, seed=0, init=None, log=None, val=None):
    """Adam on the GPU over shuffled (X, y); returns (params, info)."""
    keep = y != 0xFF
    X, y = X[keep], y[keep]
    perm = np.random.default_rng(seed).permutation(len(y))
    X, y = np.ascontiguousarray(X[perm]), np.ascontiguousarray(y[perm])
    eng.set_band_model(init_params(seed) if init is None else init)
    Xd, yd = eng.to_device(X), eng.to_device(y)
    losses = eng.empty((steps,), np.float32)
    t0 = time.time()
    chunk = 250
    for s0 in range(0, steps, chunk):
        k = min(chunk, steps - s0)
        # consecutive minibatches continue where the previous chunk stopped
        start = (s0 * batch) % len(y)
        eng.band_train_device(Xd.ptr + start * X.shape[1] * 2, yd.ptr + start, len(y) - start, batch, k, lr,
                              losses.ptr + 4 * s0)
        eng.sync()
        if log:
            ls = losses.numpy()[s0:s0 + k]
            log(f"steps {s0 + k}/{steps}: loss {ls.mean():.4f}")
    elapsed = time.time() - t0
    params = eng.get_band_model()
    ls = losses.numpy()
    info = dict(steps=steps, batch=batch, lr=lr, seed=seed, n_train=int(len(y)), train_seconds=elapsed,
                loss_first=float(ls[:20].mean()), loss_last=float(ls[-50:].mean()),
                label_hist=[int((y == c).sum()) for c in range(5)])
    if val is not None:
        Xv, yv = val
        kv = yv != 0xFF
        lg, cl = eng.band_forward(np.ascontiguousarray(Xv[kv]))
        yv = yv[kv].astype(np.int64)
        z = lg - lg.max(axis=1, keepdims=True)
        lp = z - np.log(np.exp(z).sum(axis=1, keepdims=True))
        info.update(val_loss=float(-lp[np.arange(len(yv)), yv].mean()), val_acc=float((cl == yv).mean()),
                    val_confusion=[[int(((yv == a) & (cl == b)).sum()) for b in range(5)] for a in range(5)])
    return params, info


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--workload", default="cfg2")
    t.add_argument("--scale", type=float, default=1.0)
    t.add_argument("--steps", type=int, default=1500)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args(argv)
    from .workload import CONFIGS, Workload

    def log(m):
        print(f"[band] {m}", file=sys.stderr, flush=True)

    eng = bsx.Engine(0)
    cfg = CONFIGS[args.workload]
    if args.scale != 1.0:
        cfg = cfg.scaled(args.scale)
    t0 = time.time()
    wl = Workload(cfg)
    batch, _, _, _ = wl.build_ext_batch(lambda tk: eng.extend(tk, cells=False))
    wl.close()
    log(f"{args.workload}: {len(batch)} tasks in {time.time() - t0:.1f}s")
    X, y = features_and_labels(eng, batch)
    log("label histogram (f24, f40, f80, f144, plain): " + str([int((y == c).sum()) for c in range(5)]))
    nval = min(200_000, len(y) // 10)
    params, info = train(eng, X[nval:], y[nval:], steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                         log=log, val=(X[:nval], y[:nval]))
    info.update(workload=args.workload, scale=args.scale, n_tasks=int(len(y)), n_val=int(nval))
    log(json.dumps({k: info[k] for k in ("loss_first", "loss_last", "val_loss", "val_acc")}))
    save_model(args.out, params, info)
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
