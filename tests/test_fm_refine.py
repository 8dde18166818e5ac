"""CPU checks of the pairwise 2-way FM (kahip_fm_refine): the reported gain is the cut change,
balanced partitions stay balanced and never get a higher cut, an overloaded block is relieved,
and after convergence no single move between adjacent blocks that respects the bound lowers the
cut (a brute-force check of every node and block). On the reference's golden fixtures it lowers
the LP pipeline's cut."""
import math

import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from conftest import example, random_graph
from pipeline_mirror import mirror_partition


def best_single_move(x, a, vw, ew, part, k, U):
    """Largest cut reduction of moving one node to a neighbouring block within U (brute force)."""
    n = len(x) - 1
    vw = np.ones(n, np.int64) if vw is None else np.asarray(vw, np.int64)
    ew = np.ones(len(a), np.int64) if ew is None else np.asarray(ew, np.int64)
    W = np.bincount(part, weights=vw, minlength=k)
    best = 0
    for v in range(n):
        conn = {}
        for e in range(x[v], x[v + 1]):
            b = int(part[a[e]])
            conn[b] = conn.get(b, 0) + int(ew[e])
        own = conn.get(int(part[v]), 0)
        for b, c in conn.items():
            if b != part[v] and W[b] + vw[v] <= U:
                best = max(best, c - own)
    return best


@pytest.mark.parametrize("seed", range(4))
def test_gain_balance_and_local_optimum(seed):
    x, a, vw, ew = random_graph(300, 0.03, seed, weighted=True)
    k = 4
    U = O.upper_bound(int(vw.sum()), k, 0.03)
    rng = np.random.default_rng(seed)
    # a balanced start: greedy fill of the blocks in a random node order
    part = np.zeros(300, np.int32)
    W = np.zeros(k, np.int64)
    for v in rng.permutation(300):
        b = int(np.argmin(W))
        part[v] = b
        W[b] += vw[v]
    assert W.max() <= U
    c0 = O.edge_cut(x, a, ew, part)
    p2, gain = K.fm_refine(x, a, vw, ew, k, U, part, rounds=50, seed=seed)
    c1 = O.edge_cut(x, a, ew, p2)
    assert gain == c0 - c1 and gain > 0
    assert O.block_weights(vw, k, p2).max() <= U
    assert best_single_move(x, a, vw, ew, p2, k, U) <= 0
    p3, g3 = K.fm_refine(x, a, vw, ew, k, U, part, rounds=50, seed=seed)
    assert np.array_equal(p2, p3) and g3 == gain  # deterministic


def test_overloaded_block_is_relieved():
    x, a, vw, ew = random_graph(200, 0.05, 7)
    k = 2
    U = O.upper_bound(200, k, 0.03)
    part = np.zeros(200, np.int32)
    part[:20] = 1  # block 0 far above U
    p2, _ = K.fm_refine(x, a, vw, ew, k, U, part, rounds=20, seed=1)
    w = np.bincount(p2, minlength=k)
    assert w[0] < 180 and w.max() <= max(U, w[0])


def test_single_block_and_empty_graph():
    x, a, vw, ew = random_graph(50, 0.1, 3)
    part = np.zeros(50, np.int32)
    p2, g = K.fm_refine(x, a, vw, ew, 2, 26, part, rounds=3, seed=0)
    assert g >= 0 and np.bincount(p2, minlength=2).max() <= 50
    p3, g3 = K.fm_refine(np.zeros(1, np.int32), np.zeros(0, np.int32), None, None, 2, 1,
                         np.zeros(0, np.int32))
    assert len(p3) == 0 and g3 == 0


@pytest.mark.slow
@pytest.mark.parametrize("graph,k", [("delaunay_n15", 2), ("delaunay_n15", 8),
                                     ("rgg_n_2_15_s0", 64)])
def test_lowers_the_lp_pipeline_cut(graph, k):
    x, a, vw, ew = example(graph)
    ratios = []
    for seed in range(5):
        lp, info = mirror_partition(x, a, vw, ew, k, 0.03, seed)
        fm, _ = mirror_partition(x, a, vw, ew, k, 0.03, seed, fm_rounds=K.PRESET_FM_ROUNDS)
        assert O.block_weights(vw, k, fm).max() <= info["U"]
        ratios.append(O.edge_cut(x, a, ew, fm) / O.edge_cut(x, a, ew, lp))
    gm = math.exp(np.mean(np.log(ratios)))
    print(f"{graph} k={k}: LP + pairwise FM / LP = {gm:.4f}")
    assert max(ratios) <= 1.0 and gm < 0.99
