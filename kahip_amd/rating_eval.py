"""Learned vs plain edge rating on held-out graphs (RGG, unit weights, eps = 3%).

For every seed the same GPU-generated graph is partitioned with the plain clustering (edge
weights, the FASTSOCIAL preset), with the expansion*^2 rating and with the learned rating
(committed weights, or --weights); the exact edge cut and the GPU time of coarsening + refinement
(including the rating) are reported, with geometric means of the cut ratios.

    python -m kahip_amd.rating_eval [--n 32768] [--k 2] [--seeds 10] [--levels 64]
"""
from __future__ import annotations

import argparse
import json
import math
import sys

import numpy as np

import kahip_amd as K


def gpu_ms(st):
    return st["ms_coarsen"] + st["ms_refine"] + st["ms_cut"]


def evaluate(n, k, seeds, weights, levels=64, first_seed=0, log=print,
             learned_rounds=(8, 6)):
    """Per seed: plain, expansion*^2, learned (10 clustering rounds) and learned with fewer
    clustering rounds (learned_rounds) -- the learned form at the plain form's GPU time."""
    variants = [("plain", {}),
                ("expansion_star2", {"rating": K.RATING_EXPANSION_STAR2, "rating_levels": levels}),
                ("learned", {"rating": K.RATING_LEARNED, "rating_levels": levels,
                             "mlp_weights": weights})]
    for r in learned_rounds:
        variants.append((f"learned_{r}_rounds", {"rating": K.RATING_LEARNED,
                                                 "rating_levels": levels, "mlp_weights": weights,
                                                 "cluster_rounds": r}))
    rows = []
    for s in range(first_seed, first_seed + seeds):
        g = K.Graph.rgg(n, s)
        row = {"seed": s}
        for name, kw in variants:
            _, st = g.partition(k=k, imbalance=0.03, seed=s, want_part=False, **kw)
            row[name] = {"cut": st["cut"], "balance": round(st["balance"], 5),
                         "gpu_ms": round(gpu_ms(st), 2), "rating_ms": round(st["ms_rating"], 2)}
        g.free()
        rows.append(row)
        log(json.dumps(row))

    def gm(key):
        return math.exp(np.mean([math.log(r[key]["cut"] / r["plain"]["cut"]) for r in rows]))

    names = [v[0] for v in variants]
    summary = {"n": n, "k": k, "seeds": seeds, "rating_levels": levels,
               "geomean_cut_ratio_vs_plain": {m: round(gm(m), 5) for m in names[1:]},
               "mean_gpu_ms": {m: round(float(np.mean([r[m]["gpu_ms"] for r in rows])), 2)
                               for m in names},
               "rows": rows}
    summary["geomean_cut_ratio_learned_vs_plain"] = summary["geomean_cut_ratio_vs_plain"]["learned"]
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--k", type=int, default=2)
    ap.add_argument("--seeds", type=int, default=10)
    ap.add_argument("--first-seed", type=int, default=0)
    ap.add_argument("--levels", type=int, default=64)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--learned-rounds", default="8,6,5,4")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    w = K.load_rating_weights(args.weights)
    res = evaluate(args.n, args.k, args.seeds, w, args.levels, args.first_seed,
                   log=lambda s: print(s, flush=True),
                   learned_rounds=[int(r) for r in args.learned_rounds.split(",") if r])
    res.pop("rows") if not args.out else None
    print(json.dumps({k: v for k, v in res.items() if k != "rows"}), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(res, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
