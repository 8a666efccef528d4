"""Autotuner and cost-model training of the batching form.

collect  Times every applicable plan choice (kfft_tune_candidates: radix cap,
         forced four-step, transforms per tile) of sampled length classes
         (length, input kind, batch count) as ragged batches on the GPU, and
         writes the timing database (one row per class and choice).
train    Fits the cost model (kfft_mlp: 19 -> 64 -> 64 -> 1, ReLU, FP32) to
         ln(microseconds) with the library's GPU training kernels, on the
         classes of the training split; reports the regret of its choices on
         the held-out classes against the exhaustive optimum.

    python -m kfft_amd.tune collect --classes 1300 --seed 7 --out a.npz
    python -m kfft_amd.tune collect --classes 1300 --seed 8 --out b.npz
    python -m kfft_amd.tune merge --out db.npz a.npz b.npz
    python -m kfft_amd.tune train --db db.npz --out model.npz
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import cost_features, tune_candidates, tune_measure
from .workloads import ragged_batch

DB_FIELDS = ("n", "kind", "count", "cls", "cand", "us")


def sample_classes(k: int, seed: int):
    """Length classes drawn like the ragged workload: lengths log-uniform over
    [16, 65536] (15% primes, 30% powers of two), half real input, batch counts
    log-uniform over 1..256 with at most 2^22 points per class."""
    n, kind = ragged_batch(k, seed=seed)
    rng = np.random.default_rng(seed + 1)
    count = np.exp(rng.uniform(0, math.log(256), k)).round().astype(np.int64)
    count = np.maximum(1, np.minimum(count, (1 << 22) // n.astype(np.int64)))
    return n.astype(np.int64), kind.astype(np.int64), count


def collect(classes: int, seed: int, reps: int = 3, log=print):
    cands = tune_candidates()
    n, kind, count = sample_classes(classes, seed)
    rows = {f: [] for f in DB_FIELDS}
    t0 = time.time()
    for c in range(classes):
        for j, ch in enumerate(cands):
            us = tune_measure(int(n[c]), int(kind[c]), int(count[c]), ch, reps)
            if us <= 0:
                continue
            for f, v in zip(DB_FIELDS, (n[c], kind[c], count[c], c, j, us)):
                rows[f].append(v)
        if (c + 1) % 10 == 0 or c + 1 == classes:
            log(f"# tune: {c + 1}/{classes} classes, {len(rows['us'])} timings, "
                f"{time.time() - t0:.0f} s", flush=True)
    return {f: np.asarray(v) for f, v in rows.items()}


def features(db) -> np.ndarray:
    cands = tune_candidates()
    return np.stack([cost_features(int(n), int(k), int(c), cands[int(j)])
                     for n, k, c, j in zip(db["n"], db["kind"], db["count"], db["cand"])])


def split(db, test_frac: float = 0.2, seed: int = 0):
    """Class-level split: boolean mask of the held-out rows."""
    cls = np.unique(db["cls"])
    rng = np.random.default_rng(seed)
    test = set(rng.choice(cls, int(len(cls) * test_frac), replace=False).tolist())
    return np.array([c in test for c in db["cls"]])


COST_MARGIN = 0.2231  # ln(1.25): the planner keeps the default unless 25% faster


def choose(pred: np.ndarray, db, rows: np.ndarray, margin: float = COST_MARGIN) -> dict:
    """Per class of `rows`: the row the prediction ranks first, as the tuned
    planner does (the default choice, candidate 0, unless another is predicted
    faster by more than `margin` in ln time)."""
    idx = np.nonzero(rows)[0]
    best = {int(db["cls"][i]): i for i in idx if db["cand"][i] == 0}
    dflt = dict(best)
    for i in idx:
        c = int(db["cls"][i])
        if c in dflt and pred[i] < pred[dflt[c]] - margin and \
                (best[c] == dflt[c] or pred[i] < pred[best[c]]):
            best[c] = i
    return best


def regret(db, picks: dict) -> dict:
    """Measured time of each class's pick against the fastest measured choice
    and against the default choice (candidate 0)."""
    out = {"regret": [], "gain_vs_default": []}
    for c, i in picks.items():
        rows = db["cls"] == c
        t_best = db["us"][rows].min()
        d = rows & (db["cand"] == 0)
        t_def = db["us"][d][0] if d.any() else np.nan
        out["regret"].append(db["us"][i] / t_best - 1.0)
        out["gain_vs_default"].append(t_def / db["us"][i])
    r, g = np.asarray(out["regret"]), np.asarray(out["gain_vs_default"])
    return {"classes": len(r), "mean_regret": float(r.mean()), "median_regret": float(np.median(r)),
            "p90_regret": float(np.quantile(r, 0.9)),
            "geomean_gain_vs_default": float(np.exp(np.nanmean(np.log(g))))}


def train(db, hidden: int = 64, steps: int = 20000, batch: int = 512, lr: float = 2e-3,
          seed: int = 0, log=print):
    import torch

    from . import Mlp

    X = features(db).astype(np.float32)
    y = np.log(db["us"]).astype(np.float32)
    test = split(db, seed=seed)
    Xtr, ytr = X[~test], y[~test]
    xd = torch.from_numpy(Xtr).cuda()
    yd = torch.from_numpy(ytr).cuda()
    m = Mlp(X.shape[1], hidden, seed=seed)
    rng = np.random.default_rng(seed)
    n = len(ytr)
    for s in range(steps):
        idx = torch.from_numpy(rng.integers(0, n, batch)).cuda()
        # cosine decay of the step size
        cur = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
        loss = m.train_step(xd[idx].contiguous(), yd[idx].contiguous(), cur)
        if (s + 1) % 2000 == 0:
            log(f"# train: step {s + 1}/{steps} batch loss {loss:.4f}", flush=True)
    pred = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    mse = lambda r: float(np.mean((pred[r] - y[r]) ** 2))  # noqa: E731
    stats = {"train_mse": mse(~test), "test_mse": mse(test),
             "test": regret(db, choose(pred, db, test)),
             "train": regret(db, choose(pred, db, ~test))}
    return m, stats


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--classes", type=int, default=2500)
    c.add_argument("--seed", type=int, default=7)
    c.add_argument("--reps", type=int, default=3)
    c.add_argument("--out", required=True)
    mg = sub.add_parser("merge")
    mg.add_argument("--out", required=True)
    mg.add_argument("parts", nargs="+")
    t = sub.add_parser("train")
    t.add_argument("--db", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--steps", type=int, default=20000)
    t.add_argument("--hidden", type=int, default=64)
    a = ap.parse_args(argv)
    if a.cmd == "merge":
        parts = [dict(np.load(p)) for p in a.parts]
        off = 0
        for d in parts:  # class ids unique across the parts
            d["cls"] = d["cls"] + off
            off = int(d["cls"].max()) + 1
        db = {f: np.concatenate([d[f] for d in parts]) for f in DB_FIELDS}
        np.savez_compressed(a.out, **db)
        print(f"# merged {len(db['us'])} timings of {len(np.unique(db['cls']))} classes")
    elif a.cmd == "collect":
        db = collect(a.classes, a.seed, a.reps)
        np.savez_compressed(a.out, **db)
        print(f"# wrote {len(db['us'])} timings to {a.out}")
    else:
        db = dict(np.load(a.db))
        m, stats = train(db, hidden=a.hidden, steps=a.steps)
        np.savez(a.out, params=m.get_params(), nin=m.nin, hidden=m.hidden)
        import json

        print(json.dumps(stats))
    return 0


if __name__ == "__main__":
    sys.exit(main())
