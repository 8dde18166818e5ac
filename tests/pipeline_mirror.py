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


def reorder_wanted(xadj, mode=-1):
    n = len(xadj) - 1
    if mode == 0 or n == 0:
        return False
    if mode > 0:
        return True
    deg = np.diff(np.asarray(xadj, np.int64))
    return int(deg.max()) * n > 32 * int(xadj[-1])


def reorder_by_degree(xadj, adjncy, vwgt, adjwgt):
    """Node i of the result = the i-th node by (degree desc, id asc); rows keep their edge order
    with mapped ids. Returns (graph, inv) with inv[old] = new."""
    x = np.asarray(xadj, np.int64)
    deg = np.diff(x)
    order = np.argsort(-deg, kind="stable")
    inv = np.empty(len(deg), np.int64)
    inv[order] = np.arange(len(deg))
    nx = np.concatenate([[0], np.cumsum(deg[order])])
    # old edge index of every new entry: rows of order[] concatenated
    starts = np.repeat(x[order] - nx[:-1], deg[order])
    old_e = np.arange(int(nx[-1])) + starts
    na = inv[np.asarray(adjncy, np.int64)[old_e]].astype(np.int32)
    nw = None if adjwgt is None else np.asarray(adjwgt, np.int32)[old_e]
    nv = None if vwgt is None else np.asarray(vwgt, np.int32)[order]
    return (nx.astype(np.int32), na, nv, nw), inv


def ensemble_factor(seed, level, run):
    return 10 + O.hash3(seed, 0xE5 + level, run) % 21


def ensemble_seed(seed, level, run):
    return O.hash3(seed, 0xE6 + level, run)


def overlay(a, b):
    """Smallest node id among the nodes sharing (a[v], b[v])."""
    key = np.asarray(a, np.int64) * (len(a) + 1) + np.asarray(b, np.int64)
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    return first[inv].astype(np.int32)


def mirror_partition(xadj, adjncy, vwgt, adjwgt, k, imbalance=0.03, seed=0, cluster_rounds=10,
                     refine_rounds=25, subrounds=4, cluster_factor=18, stop_factor=5000,
                     ip=None, scores=None, rating_levels=64, reorder=-1, ensemble=1, fm_rounds=0):
    """As mirror_partition_plain, after the library's degree renumbering of skewed inputs
    (reorder: -1 automatic, 0 off, 1 on); the partition is returned in the input numbering."""
    if reorder_wanted(xadj, reorder):
        g, inv = reorder_by_degree(xadj, adjncy, vwgt, adjwgt)
        part, info = mirror_partition_plain(*g, k, imbalance, seed, cluster_rounds,
                                            refine_rounds, subrounds, cluster_factor,
                                            stop_factor, ip, scores, rating_levels, ensemble,
                                            fm_rounds)
        info["reordered"] = True
        return part[inv], info
    part, info = mirror_partition_plain(xadj, adjncy, vwgt, adjwgt, k, imbalance, seed,
                                        cluster_rounds, refine_rounds, subrounds, cluster_factor,
                                        stop_factor, ip, scores, rating_levels, ensemble,
                                        fm_rounds)
    info["reordered"] = False
    return part, info


def mirror_partition_plain(xadj, adjncy, vwgt, adjwgt, k, imbalance=0.03, seed=0,
                           cluster_rounds=10, refine_rounds=25, subrounds=4, cluster_factor=18,
                           stop_factor=5000, ip=None, scores=None, rating_levels=64, ensemble=1,
                           fm_rounds=0):
    """Returns (part, info). `ip(xadj, adjncy, vwgt, adjwgt, k, U, reps, seed)` partitions the
    coarsest graph (the library's kahip_amd.initial_partition by default). `scores(xadj, adjncy,
    vwgt, adjwgt, B, rating_seed)` (optional) gives the int32 edge scores that replace the edge
    weights in the clustering of levels < rating_levels. fm_rounds > 0 applies the library's CPU
    pairwise FM (kahip_amd.fm_refine) to every level after its LP refinement, as the pipeline
    does (seed fm_seed(seed) + level)."""
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
        for run in range(1, ensemble):
            eb = O.cluster_bound(U, ensemble_factor(seed, level, run))
            other, _, _ = O.lp_cluster_sync(x, a, vw, rw, eb, cluster_rounds, subrounds,
                                            ensemble_seed(seed, level, run))
            cl = overlay(cl, other)
        if ensemble > 1:
            cl, nc = O.remap_first_appearance(cl)
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
        if fm_rounds > 0 and len(x) > 1:
            import kahip_amd
            part, _ = kahip_amd.fm_refine(x, a, vw, ew, k, U, part, rounds=fm_rounds,
                                          seed=(kahip_amd.fm_seed(seed) + l) & (2**64 - 1))
    info = {"U": U, "B": B, "levels": level,
            "level_n": [len(lv[0]) - 1 for lv in levels], "initial": initial,
            "coarsest": levels[level]}
    return part, info


def mirror_batch(graphs, k, imbalance=0.03, seed=0, cluster_rounds=10, refine_rounds=25,
                 subrounds=4, cluster_factor=18, stop_factor=5000, ip=None, ip_tries=-1):
    """CPU mirror of the batched mode: the disjoint union of `graphs` ((xadj, adjncy, vwgt,
    adjwgt) with local ids) under the synchronous schedule with per-label bounds (the cluster
    bound of a node's graph; the block bound U_g of labels g*k + b), per-graph stop rules and
    per-graph coarsest levels; ip_tries: BFS-grown starts per bisection of the initial
    partitioner (-1: its default). Returns (local parts per graph, info)."""
    if ip is None:
        import kahip_amd

        def ip(*args):
            return kahip_amd.initial_partition(*args, tries=ip_tries)
    G = len(graphs)
    # the union
    noff = [0]
    xs, adjs, vws, ews = [], [], [], []
    eoff = 0
    for x, a, vw, ew in graphs:
        n = len(x) - 1
        xs.append(np.asarray(x[:-1], np.int64) + eoff)
        adjs.append(np.asarray(a, np.int64) + noff[-1])
        vws.append(np.ones(n, np.int64) if vw is None else np.asarray(vw, np.int64))
        ews.append(np.ones(int(x[-1]), np.int64) if ew is None else np.asarray(ew, np.int64))
        eoff += int(x[-1])
        noff.append(noff[-1] + n)
    ux = np.concatenate(xs + [np.array([eoff])]).astype(np.int32)
    ua = np.concatenate(adjs).astype(np.int32) if eoff else np.zeros(0, np.int32)
    uv = np.concatenate(vws).astype(np.int32)
    uw = np.concatenate(ews).astype(np.int32) if eoff else np.zeros(0, np.int32)
    W = [int(v.sum()) for v in vws]
    U = [O.upper_bound(w, k, imbalance) for w in W]
    B = [O.cluster_bound(u, cluster_factor) for u in U]
    levels = [(ux, ua, uv, uw)]
    offs = [np.array(noff, np.int64)]
    maps = []
    on = np.ones(G, bool)
    coarsest = np.zeros(G, np.int64)
    level = 0
    while on.any():
        x, a, vw, ew = levels[level]
        n = len(x) - 1
        o = offs[level]
        ng = np.repeat(np.arange(G), np.diff(o))
        nbound = np.array(B, np.int64)[ng].astype(np.int32)
        include = on[ng].astype(np.uint8)
        cs = O.hash3(seed, 0xC1, level)
        cl, nc, _ = O.lp_cluster_sync_ex(x, a, vw, ew, nbound, include, cluster_rounds,
                                         subrounds, cs)
        cx, ca, cw, cv = O.contract(x, a, vw, ew, cl, nc)
        maps.append(cl)
        levels.append((cx, ca, cv, cw))
        co = np.array([cl[o[g]] for g in range(G)] + [nc], np.int64)
        offs.append(co)
        level += 1
        for g in range(G):
            if on[g] and not O.stop_rule_continue(int(o[g + 1] - o[g]), int(co[g + 1] - co[g]),
                                                  k, stop_factor):
                on[g] = False
                coarsest[g] = level
    # initial partitions per graph at its coarsest level
    reps = ip_repetitions(k)
    base = ip_seed(seed)
    ip_part = {}
    for g in range(G):
        l = int(coarsest[g])
        x, a, vw, ew = levels[l]
        o = offs[l]
        lo, hi = int(o[g]), int(o[g + 1])
        e0, e1 = int(x[lo]), int(x[hi])
        sx = (np.asarray(x[lo:hi + 1], np.int64) - e0).astype(np.int32)
        sa = (np.asarray(a[e0:e1], np.int64) - lo).astype(np.int32)
        sv = np.asarray(vw[lo:hi], np.int32)
        sw = np.asarray(ew[e0:e1], np.int32)
        gseed = base ^ ((O.hash3(seed, 0xB7, g) << 1) & 0xFFFFFFFFFFFFFFFF)
        ip_part[g] = np.asarray(ip(sx, sa, sv, sw, k, U[g], reps, gseed), np.int64)
    # uncoarsening
    lbound = np.repeat(np.array(U, np.int64), k).astype(np.int32)
    part = None
    for l in range(level, -1, -1):
        x, a, vw, ew = levels[l]
        o = offs[l]
        ng = np.repeat(np.arange(G), np.diff(o))
        if l == level:
            part = (ng * k).astype(np.int32)
        else:
            part = O.project(maps[l], part)
        for g in range(G):
            if coarsest[g] == l and l > 0:
                part[int(o[g]):int(o[g + 1])] = g * k + ip_part[g]
        include = (coarsest[ng] >= l).astype(np.uint8)
        rs = O.hash3(seed, 0x4E, l)
        part, _ = O.lp_refine_sync_ex(x, a, vw, ew, G * k, lbound, include, refine_rounds,
                                      subrounds, rs, part)
    out = []
    for g in range(G):
        out.append(part[noff[g]:noff[g + 1]] - g * k)
    info = {"U": U, "B": B, "levels": level, "coarsest": coarsest.tolist(),
            "union": levels[0], "node_off": noff}
    return out, info
