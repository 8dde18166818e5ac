"""Closed forms, brute force on small graphs and invariants that pin the oracle independently of
the reference, plus planted mistakes that these checks must catch."""
import itertools

import numpy as np
import pytest

import oracle as O
from conftest import manual_graph, random_graph
from kahip_amd import graphs


def src_of(xadj):
    return np.repeat(np.arange(len(xadj) - 1), np.diff(xadj))


def cut_numpy(xadj, adjncy, adjwgt, part):
    w = np.ones(len(adjncy), np.int64) if adjwgt is None else adjwgt.astype(np.int64)
    s = src_of(xadj)
    return int(w[part[s] != part[adjncy]].sum()) // 2


def contract_by_sort(xadj, adjncy, vwgt, adjwgt, cl, nc):
    """Quotient graph by an independent route: group cut edges by (p, q), sum weights, keep the
    smallest fine edge index, order each row by it (the first-appearance order of
    complete_boundary.h:511-562 expressed as a sort)."""
    n = len(xadj) - 1
    s = src_of(xadj)
    p, q = cl[s], cl[adjncy]
    w = np.ones(len(adjncy), np.int64) if adjwgt is None else adjwgt.astype(np.int64)
    e = np.arange(len(adjncy))
    keep = p != q
    p, q, w, e = p[keep], q[keep], w[keep], e[keep]
    groups = {}
    for pi, qi, wi, ei in zip(p.tolist(), q.tolist(), w.tolist(), e.tolist()):
        g = groups.setdefault((pi, qi), [0, ei])
        g[0] += wi
        g[1] = min(g[1], ei)
    rows = [[] for _ in range(nc)]
    for (pi, qi), (wi, ei) in groups.items():
        rows[pi].append((ei, qi, wi))
    cx = [0]
    ca, cw = [], []
    for r in rows:
        r.sort()
        ca += [qi for _, qi, _ in r]
        cw += [wi for _, _, wi in r]
        cx.append(len(ca))
    vw = np.ones(n, np.int64) if vwgt is None else vwgt.astype(np.int64)
    cv = np.bincount(cl, weights=vw, minlength=nc).astype(np.int64)
    return (np.array(cx, np.int32), np.array(ca, np.int32), np.array(cw, np.int32),
            cv.astype(np.int32))


GRAPHS = [manual_graph(), random_graph(30, 0.2, 1), random_graph(60, 0.1, 2, weighted=True),
          random_graph(40, 0.3, 3, weighted=True)]


@pytest.mark.parametrize("gi", range(len(GRAPHS)))
def test_edge_cut_block_weights_balance_brute_force(gi):
    x, a, vw, ew = GRAPHS[gi]
    n = len(x) - 1
    rng = np.random.default_rng(gi)
    for k in (1, 2, 3, 7):
        part = rng.integers(0, k, n).astype(np.int32)
        assert O.edge_cut(x, a, ew, part) == cut_numpy(x, a, ew, part)
        vwv = np.ones(n, np.int64) if vw is None else vw.astype(np.int64)
        bw = np.bincount(part, weights=vwv, minlength=k).astype(np.int64)
        assert np.array_equal(O.block_weights(vw, k, part), bw)
        assert O.balance(vw, k, part) == bw.max() / np.ceil(vwv.sum() / k)


def test_remap_closed_form():
    rng = np.random.default_rng(5)
    for n in (1, 7, 50):
        lab = rng.integers(0, n, n)
        got, nc = O.remap_first_appearance(lab)
        # new id of a cluster = rank of its smallest member among all clusters' smallest members
        mins = {c: np.flatnonzero(lab == c).min() for c in np.unique(lab)}
        order = sorted(mins, key=lambda c: mins[c])
        want = np.array([order.index(c) for c in lab], np.int32)
        assert nc == len(mins) and np.array_equal(got, want)


@pytest.mark.parametrize("gi", range(len(GRAPHS)))
def test_contract_matches_sort_formulation(gi):
    x, a, vw, ew = GRAPHS[gi]
    n = len(x) - 1
    rng = np.random.default_rng(10 + gi)
    for nclus in (1, 3, n // 3 + 1, n):
        cl, nc = O.remap_first_appearance(rng.integers(0, nclus, n))
        got = O.contract(x, a, vw, ew, cl, nc)
        want = contract_by_sort(x, a, vw, ew, cl, nc)
        for g_, w_ in zip(got, want):
            assert np.array_equal(g_, w_)
        cx, ca, cw, cv = got
        graphs.check_graph(cx, ca, cv, cw)  # symmetric, no self-loops, positive weights
        assert cv.sum() == (n if vw is None else vw.sum())
        wsum = len(a) if ew is None else int(ew.sum())
        intra = wsum - 2 * cut_numpy(x, a, ew, cl)
        assert int(cw.sum()) == wsum - intra


def test_contract_brute_force_all_partitions_of_a_tiny_graph():
    x, a, vw, ew = manual_graph()
    n = len(x) - 1
    for labels in itertools.product(range(3), repeat=n):
        cl, nc = O.remap_first_appearance(np.array(labels))
        got = O.contract(x, a, vw, ew, cl, nc)
        want = contract_by_sort(x, a, vw, ew, cl, nc)
        for g_, w_ in zip(got, want):
            assert np.array_equal(g_, w_)


def test_project_and_stop_rule_closed_forms():
    rng = np.random.default_rng(1)
    m = rng.integers(0, 10, 100)
    cp = rng.integers(0, 4, 10)
    assert np.array_equal(O.project(m, cp), cp[m])
    # stop_rules.h:108-111: continue iff n_f / n_c >= 1.1 and n_c >= 5000 k
    assert O.stop_rule_continue(4194304, 400000, 64) is True
    assert O.stop_rule_continue(4194304, 300000, 64) is False      # below 5000 k = 320000
    assert O.stop_rule_continue(1100000, 1000001, 2) is False      # rate < 1.1
    assert O.stop_rule_continue(1100000, 1000000, 2) is True       # rate == 1.1


def test_lp_cluster_async_special_cases():
    # edgeless graph: every node stays a singleton
    x, a = np.zeros(6, np.int32), np.zeros(0, np.int32)
    cl, nc = O.lp_cluster_async(x, a, None, None, 100, 10, 0)
    assert nc == 5 and np.array_equal(cl, np.arange(5))
    # bound 1 forbids every merge (size + 1 < 1 never holds)
    x, a, _, _ = random_graph(30, 0.3, 4)
    cl, nc = O.lp_cluster_async(x, a, None, None, 1, 10, 0)
    assert nc == 30
    # cluster weights stay below the bound
    for B in (2, 5, 17):
        cl, nc = O.lp_cluster_async(x, a, None, None, B, 10, 1)
        assert np.bincount(cl).max() < B


# ------------------------------------------------------------------------- planted mistakes
def test_planted_contraction_order_mistake_is_caught():
    """Emitting each coarse row in ascending neighbour id (a plausible slip) differs from the
    first-appearance order on the reference example, so the order checks above are sharp."""
    from conftest import example
    x, a, vw, ew = example("rgg_n_2_15_s0")
    cl, nc = O.lp_cluster_async(x, a, vw, ew, 938, 10, 0)
    cx, ca, cw, cv = O.contract(x, a, vw, ew, cl, nc)
    wrong = ca.copy()
    for p in range(nc):
        wrong[cx[p]:cx[p + 1]] = np.sort(ca[cx[p]:cx[p + 1]])
    assert not np.array_equal(wrong, ca)


def test_planted_cut_mistake_is_caught():
    """Forgetting the halving of the directed sum, or counting intra-block edges, changes the
    value the brute-force check compares against."""
    x, a, vw, ew = GRAPHS[2]
    part = np.random.default_rng(0).integers(0, 3, len(x) - 1).astype(np.int32)
    assert O.edge_cut(x, a, ew, part) * 2 != O.edge_cut(x, a, ew, part)
    assert cut_numpy(x, a, ew, np.zeros_like(part)) == 0


def test_upper_bound_perfectly_balanced_mode():
    """app/balance_configuration.h:41-43: imbalance 0 partitions under 1.01 ceil(W/k)."""
    assert O.upper_bound(32768, 2, 0.0) == int(1.01 * 16384) == 16547
    assert O.upper_bound(32768, 7, 0.0) == int(1.01 * 4682)
    assert O.upper_bound(1000, 3, 0.0) == int(1.01 * 334)
    assert O.upper_bound(32768, 2, 0.03) == 16875
