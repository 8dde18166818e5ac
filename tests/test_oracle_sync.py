"""The synchronous schedule (R2) that the GPU kernels implement has no reference counterpart; it is
pinned here by invariants and special cases whose answers are fixed by the definition."""
import numpy as np
import pytest

import oracle as O
from conftest import example, manual_graph, random_graph


def clique(n, offset=0):
    edges = [(i + offset, j + offset) for i in range(n) for j in range(n) if i != j]
    return edges


def csr(n, directed_edges, w=None):
    u = np.array([a for a, b in directed_edges], np.int64)
    v = np.array([b for a, b in directed_edges], np.int64)
    o = np.lexsort((v, u))
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, u + 1, 1)
    return np.cumsum(xadj).astype(np.int32), v[o].astype(np.int32)


def test_hash3_values_are_fixed():
    # the constants of the schedule; a change here changes every GPU result
    assert O.hash3(0, 0, 0) == O.hash3(0, 0, 0)
    assert O.hash3(1, 2, 3) != O.hash3(1, 3, 2)
    vals = [O.hash3(7, r, v) % 4 for r in range(1) for v in range(4000)]
    counts = np.bincount(vals, minlength=4)
    assert counts.min() > 900  # sub-rounds are balanced


@pytest.mark.parametrize("S", [1, 4, 8])
def test_cluster_weights_below_bound_and_labels_valid(S):
    for g in (example("rgg_n_2_15_s0"), random_graph(300, 0.03, 7, weighted=True)):
        x, a, vw, ew = g
        n = len(x) - 1
        vwv = np.ones(n, np.int64) if vw is None else vw.astype(np.int64)
        for B in (5, 40, 938):
            cl, nc, moves = O.lp_cluster_sync(x, a, vw, ew, B, 10, S, 3)
            assert cl.min() == 0 and cl.max() == nc - 1
            cw = np.bincount(cl, weights=vwv)
            assert np.all((cw < B) | (np.bincount(cl) == 1))


def test_sync_cluster_special_cases():
    # edgeless: identity
    x, a = np.zeros(8, np.int32), np.zeros(0, np.int32)
    cl, nc, mv = O.lp_cluster_sync(x, a, None, None, 100, 10, 4, 0)
    assert nc == 7 and np.array_equal(cl, np.arange(7)) and mv.sum() == 0
    # bound 1: no merge possible (1 + 1 < 1 is false)
    x, a, _, _ = random_graph(40, 0.2, 1)
    cl, nc, mv = O.lp_cluster_sync(x, a, None, None, 1, 10, 4, 0)
    assert nc == 40 and mv.sum() == 0
    # bound 3 with unit weights: clusters of at most 2 nodes, and every merge is an edge
    cl, nc, mv = O.lp_cluster_sync(x, a, None, None, 3, 10, 4, 0)
    sizes = np.bincount(cl)
    assert sizes.max() <= 2
    src = np.repeat(np.arange(40), np.diff(x))
    for c in np.flatnonzero(sizes == 2):
        u, v = np.flatnonzero(cl == c)
        assert np.any((src == u) & (a == v))


def test_sync_cluster_dumbbell_splits_into_cliques():
    edges = clique(8) + clique(8, 8) + [(0, 8), (8, 0)]
    x, a = csr(16, edges)
    for seed in range(5):
        cl, nc, _ = O.lp_cluster_sync(x, a, None, None, 9, 10, 4, seed)
        assert nc == 2
        assert len(set(cl[:8])) == 1 and len(set(cl[8:])) == 1


def test_sync_cluster_complete_graph_collapses():
    x, a = csr(12, clique(12))
    for seed in range(5):
        cl, nc, _ = O.lp_cluster_sync(x, a, None, None, 100, 10, 4, seed)
        assert nc == 1


def test_sync_cluster_deterministic_and_seed_sensitive():
    x, a, vw, ew = example("delaunay_n15")
    c1 = O.lp_cluster_sync(x, a, vw, ew, 938, 10, 4, 11)[0]
    c2 = O.lp_cluster_sync(x, a, vw, ew, 938, 10, 4, 11)[0]
    c3 = O.lp_cluster_sync(x, a, vw, ew, 938, 10, 4, 12)[0]
    assert np.array_equal(c1, c2) and not np.array_equal(c1, c3)


@pytest.mark.parametrize("k", [2, 8, 64])
def test_sync_refine_respects_bound_and_keeps_labels(k):
    x, a, vw, ew = example("rgg_n_2_15_s0")
    n = len(x) - 1
    U = O.upper_bound(n, k, 0.03)
    start = (np.arange(n) * k // n).astype(np.int32)
    p, moves = O.lp_refine_sync(x, a, vw, ew, k, U, 25, 4, 5, start)
    assert p.min() >= 0 and p.max() < k
    assert O.block_weights(vw, k, p).max() <= U
    assert O.edge_cut(x, a, ew, p) < O.edge_cut(x, a, ew, start)
    assert moves[0] > 0


def test_sync_refine_drains_overloaded_block():
    x, a, vw, ew = example("delaunay_n15")
    n = len(x) - 1
    k = 4
    U = O.upper_bound(n, k, 0.03)
    rng = np.random.default_rng(0)
    start = np.where(rng.random(n) < 0.6, 0, rng.integers(1, k, n)).astype(np.int32)
    assert O.block_weights(vw, k, start)[0] > U
    p, _ = O.lp_refine_sync(x, a, vw, ew, k, U, 25, 4, 1, start)
    assert O.block_weights(vw, k, p)[0] < O.block_weights(vw, k, start)[0]


def test_sync_refine_frontier_stops():
    x, a, vw, ew = manual_graph()
    p, moves = O.lp_refine_sync(x, a, vw, ew, 2, 4, 25, 4, 0, np.array([0, 0, 1, 1, 1], np.int32))
    assert moves[-1] == 0  # converges well before 25 rounds
