"""Trains the learned edge rating on the GPU and writes kahip_amd/data/rating_mlp.npy (+ .json).

Training set: seeded RGG and R-MAT graphs with n = 2^16 .. 2^20 (log2 n drawn uniformly from
the integers 16..20 per graph), alternating kinds, generator seeds 1000 + i (disjoint from the
seeds 0..99 used for evaluation). Each graph is partitioned by this library's own plain
pipeline (k = 64, eps = 3%); an edge is labelled 1 if its endpoints share a block (uncut) and 0
otherwise. Features are those of kahip_edge_features on the input graph with the k = 64 cluster
bound. Per graph an equal number of cut and uncut edges is sampled (numpy Generator, seed 0; at
most per_graph / 2 of each), all examples are shuffled once, and the MLP is trained by Adam with the library's GPU forward/backward kernels
(kahip_mlp_train). Everything is seeded, so a rerun writes the same weights.

    python -m kahip_amd.rating_train [--graphs 200] [--per-graph 40000] [--epochs 6]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import kahip_amd as K

FEATURES, HIDDEN = K.MLP_FEATURES, K.MLP_HIDDEN
OFF_W1 = 0
OFF_B1 = OFF_W1 + HIDDEN * FEATURES
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN


def init_params(seed=0):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(K.MLP_PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / FEATURES), HIDDEN * FEATURES)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    return p


def training_graph(i):
    """(kind, log2 n, device Graph) of training graph i."""
    seed = 1000 + i
    rng = np.random.default_rng(seed)
    lg = int(rng.integers(16, 21))
    if i % 2 == 0:
        return "rgg", lg, K.Graph.rgg(1 << lg, seed)
    return "rmat", lg, K.Graph.rmat(lg, 16 << lg, seed)


def cluster_bound(n, k, imbalance=0.03, factor=18):
    U = int((1 + imbalance) * -(-n // k))
    return -(-U // factor)


def collect(num_graphs, per_graph, k=64, log=print):
    rng = np.random.default_rng(0)
    feats, labels, meta = [], [], []
    for i in range(num_graphs):
        t0 = time.time()
        kind, lg, g = training_graph(i)
        inf = g.info()
        part, st = g.partition(k=k, imbalance=0.03, seed=i, ip_repetitions=1)
        x, a, _, _ = g.to_host()
        f = K.graph_edge_features(g, cluster_bound(inf["n"], k), 7919 * i + 1)
        src = np.repeat(np.arange(inf["n"]), np.diff(x))
        y = (part[src] == part[a]).astype(np.float32)
        # class-balanced sample: the score only ranks the edges around one node, so the model
        # should separate cut from uncut edges within a graph, not one graph from another
        pos, neg = np.nonzero(y > 0.5)[0], np.nonzero(y < 0.5)[0]
        half = min(per_graph // 2, len(pos), len(neg))
        take = np.concatenate([rng.choice(pos, size=half, replace=False),
                               rng.choice(neg, size=half, replace=False)])
        feats.append(f[take])
        labels.append(y[take])
        meta.append({"kind": kind, "log2n": lg, "n": inf["n"], "m2": inf["m2"],
                     "cut": st["cut"], "uncut_frac": float(y.mean())})
        g.free()
        log(f"graph {i:3d} {kind:4s} n=2^{lg} m2={inf['m2']} cut={st['cut']} "
            f"uncut={y.mean():.3f} ({time.time() - t0:.1f}s)")
    return np.concatenate(feats), np.concatenate(labels), meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--per-graph", type=int, default=40000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=os.path.join(K.HERE, "data", "rating_mlp.npy"))
    args = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)

    t0 = time.time()
    X, y, meta = collect(args.graphs, args.per_graph, log=log)
    order = np.random.default_rng(1).permutation(len(y))
    X, y = X[order], y[order]
    nval = len(y) // 20
    Xv, yv, Xt, yt = X[:nval], y[:nval], X[nval:], y[nval:]
    log(f"examples: {len(yt)} train / {nval} validation, uncut fraction {y.mean():.3f}")
    t1 = time.time()
    w, losses = K.mlp_train(Xt, yt, init_params(0), batch=args.batch, epochs=args.epochs,
                            lr=args.lr)
    t2 = time.time()
    zv = K.mlp_forward(Xv, w)
    acc = float(((zv > 0) == (yv > 0.5)).mean())
    vloss = float(np.mean(np.maximum(zv, 0) - zv * yv + np.log1p(np.exp(-np.abs(zv)))))
    base = float(max(yv.mean(), 1 - yv.mean()))
    log(f"train loss per epoch {np.round(losses, 4).tolist()}; validation loss {vloss:.4f}, "
        f"accuracy {acc:.4f} (majority {base:.4f}); training {t2 - t1:.1f}s")
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, w.astype(np.float32))
    info = {"graphs": args.graphs, "per_graph": args.per_graph, "epochs": args.epochs,
            "batch": args.batch, "lr": args.lr, "examples_train": int(len(yt)),
            "examples_validation": int(nval), "train_loss": [float(v) for v in losses],
            "validation_loss": vloss, "validation_accuracy": acc, "majority_accuracy": base,
            "seconds_collect": t1 - t0, "seconds_train": t2 - t1, "graph_meta": meta}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(info, fh, indent=1)
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
