"""CPU mirror of the library's whole FASTSOCIAL pipeline, composed from the oracle's R2 steps
(synchronous LP clustering and refinement, first-appearance relabel, contraction, projection)
and the library's own CPU initial partitioner, with the library's per-level seeds:
    cluster seed of level l = hash3(seed, 0xC1, l), refinement seed = hash3(seed, 0x4E, l),
    rating seed = hash3(seed, 0x5A, l).
The GPU pipeline must reproduce its partition bit for bit. Test infrastructure only."""
from __future__ import annotations

import math

import numpy as np

import oracle as O


def ip_repetitions(k):
    reps = 2
    if k > 1:
        reps = max(int(math.ceil(16 / math.log2(k))), 2)
    return min(4, reps)


def ip_seed(seed):
    return ((seed & 0xFFFFFFFF) * 0x9E3779B97F4A7C15 + 0x1234567) & 0xFFFFFFFFFFFFFFFF


def mirror_partition(xadj, adjncy, vwgt, adjwgt, k, imbalance=0.03, seed=0, cluster_rounds=10,
                     refine_rounds=25, subrounds=4, cluster_factor=18, stop_factor=5000,
                     ip=None, scores=None, rating_levels=64):
    """Returns (part, info). `ip(xadj, adjncy, vwgt, adjwgt, k, U, reps, seed)` partitions the
    coarsest graph (the library's kahip_amd.initial_partition by default). `scores(xadj, adjncy,
    vwgt, adjwgt, B, rating_seed)` (optional) gives the int32 edge scores that replace the edge
    weights in the clustering of levels < rating_levels."""
    if ip is None:
        import kahip_amd
        ip = kahip_amd.initial_partition
    n = len(xadj) - 1
    W = n if vwgt is None else int(np.asarray(vwgt, np.int64).sum())
    U = O.upper_bound(W, k, imbalance)
    B = O.cluster_bound(U, cluster_factor)
    levels = [(np.asarray(xadj, np.int32), np.asarray(adjncy, np.int32),
               None if vwgt is None else np.asarray(vwgt, np.int32),
               None if adjwgt is None else np.asarray(adjwgt, np.int32))]
    maps = []
    level = 0
    again = n > 0
    while again:
        x, a, vw, ew = levels[level]
        cs = O.hash3(seed, 0xC1, level)
        rw = ew
        if scores is not None and level < rating_levels and int(x[-1]) > 0:
            rw = np.asarray(scores(x, a, vw, ew, B, O.hash3(seed, 0x5A, level)), np.int32)
        cl, nc, _ = O.lp_cluster_sync(x, a, vw, rw, B, cluster_rounds, subrounds, cs)
        cx, ca, cw, cv = O.contract(x, a, vw, ew, cl, nc)
        maps.append(cl)
        levels.append((cx, ca, cv, cw))
        fine_n = len(x) - 1
        level += 1
        again = O.stop_rule_continue(fine_n, nc, k, stop_factor)
    x, a, vw, ew = levels[level]
    part = np.asarray(ip(x, a, vw, ew, k, U, ip_repetitions(k), ip_seed(seed)), np.int32)
    initial = part.copy()
    for l in range(level, -1, -1):
        x, a, vw, ew = levels[l]
        if l != level:
            part = O.project(maps[l], part)
        rs = O.hash3(seed, 0x4E, l)
        part, _ = O.lp_refine_sync(x, a, vw, ew, k, U, refine_rounds, subrounds, rs, part)
    info = {"U": U, "B": B, "levels": level,
            "level_n": [len(lv[0]) - 1 for lv in levels], "initial": initial,
            "coarsest": levels[level]}
    return part, info
