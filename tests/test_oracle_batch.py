"""CPU checks of the batched schedule (oracle lp_*_sync_ex and the batch mirror): per-label
bounds, inclusion masks and the per-graph structure of the results."""
import numpy as np

import oracle as O
from conftest import example, random_graph
from kahip_amd import graphs as GR
from pipeline_mirror import mirror_batch, mirror_partition


def union(graphs):
    xs, adjs = [], []
    noff, eoff = [0], 0
    for x, a, _, _ in graphs:
        xs.append(np.asarray(x[:-1], np.int64) + eoff)
        adjs.append(np.asarray(a, np.int64) + noff[-1])
        eoff += int(x[-1])
        noff.append(noff[-1] + len(x) - 1)
    ux = np.concatenate(xs + [np.array([eoff])]).astype(np.int32)
    return ux, np.concatenate(adjs).astype(np.int32), np.array(noff)


def test_cluster_ex_with_uniform_bounds_equals_plain():
    x, a, vw, ew = random_graph(400, 0.02, 3, weighted=True)
    want, nc, mv = O.lp_cluster_sync(x, a, vw, ew, 25, 5, 4, 99)
    got, nc2, mv2 = O.lp_cluster_sync_ex(x, a, vw, ew, np.full(400, 25), None, 5, 4, 99)
    assert nc == nc2 and np.array_equal(want, got) and np.array_equal(mv, mv2)
    part = np.arange(400) % 4
    p1, _ = O.lp_refine_sync(x, a, vw, ew, 4, 300, 6, 4, 7, part)
    p2, _ = O.lp_refine_sync_ex(x, a, vw, ew, 4, np.full(4, 300), None, 6, 4, 7, part)
    assert np.array_equal(p1, p2)


def test_cluster_ex_respects_per_graph_bounds_and_masks():
    gs = [GR.gen_rgg(2000, 1), random_graph(300, 0.05, 2), GR.gen_rgg(800, 3)]
    x, a, off = union(gs)
    ng = np.repeat(np.arange(3), np.diff(off))
    B = np.array([5, 30, 12])
    include = (ng != 2).astype(np.uint8)  # graph 2 stays out
    cl, nc, _ = O.lp_cluster_sync_ex(x, a, None, None, B[ng], include, 10, 4, 5)
    # clusters never cross graphs, respect their graph's bound, and graph 2 is untouched
    w = np.bincount(cl, minlength=nc)
    for c in range(nc):
        members = np.nonzero(cl == c)[0]
        gset = set(ng[members].tolist())
        assert len(gset) == 1
        assert w[c] < B[gset.pop()] or len(members) == 1
    lo, hi = off[2], off[3]
    assert len(set(cl[lo:hi].tolist())) == hi - lo


def test_refine_ex_keeps_blocks_inside_graphs():
    gs = [GR.gen_rgg(1500, 4), random_graph(200, 0.05, 5)]
    x, a, off = union(gs)
    k = 4
    ng = np.repeat(np.arange(2), np.diff(off))
    rng = np.random.default_rng(0)
    part = (ng * k + rng.integers(0, k, len(ng))).astype(np.int32)
    U = np.array([O.upper_bound(1500, k, 0.03), O.upper_bound(200, k, 0.03)])
    lb = np.repeat(U, k)
    p, mv = O.lp_refine_sync_ex(x, a, None, None, 2 * k, lb, None, 25, 4, 3, part)
    assert np.all(p // k == ng)
    for g in range(2):
        w = np.bincount(p[ng == g] - g * k, minlength=k)
        assert w.max() <= max(U[g], np.bincount(part[ng == g] - g * k, minlength=k).max())
    assert mv.sum() > 0


def test_batch_mirror_per_graph_results():
    gs = [GR.gen_rgg(3000, 1), GR.gen_rmat(10, 16 << 10, 2), example("example_weighted")]
    parts, info = mirror_batch(gs, 4, 0.03, 2, stop_factor=5)
    for (x, a, vw, ew), p, U in zip(gs, parts, info["U"]):
        assert p.min() >= 0 and p.max() < 4
        assert O.block_weights(vw, 4, p).max() <= U
    assert info["levels"] >= 2
    # the single-graph pipeline on a comparable graph gives a comparable cut
    x, a, vw, ew = gs[0]
    single, _ = mirror_partition(x, a, vw, ew, 4, 0.03, 2, stop_factor=5)
    c_single = O.edge_cut(x, a, ew, single)
    c_batch = O.edge_cut(x, a, ew, parts[0])
    assert c_batch <= 1.5 * c_single + 20
