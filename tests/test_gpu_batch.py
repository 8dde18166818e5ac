"""GPU tests of the batched mode: the generated union against the numpy generators, the batched
partition against the CPU mirror of the union schedule (bit for bit), hipGraph replay against
eager launches, per-graph cuts and balance, and the bin autotuner."""
import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from kahip_amd import graphs as GR
from conftest import random_graph
from pipeline_mirror import mirror_batch

pytestmark = pytest.mark.gpu


def small_batch():
    return [GR.gen_rgg(3000, 1), GR.gen_rmat(10, 16 << 10, 2), GR.gen_rgg(1500, 3),
            GR.gen_rgg(200, 4), random_graph(60, 0.1, 5, weighted=True), (np.array([0, 0],
            np.int32), np.zeros(0, np.int32), None, None)]


def test_generated_batch_matches_numpy():
    b = K.Batch.generate(6, 8.0, 11.0, seed=3)
    inf = b.info()
    x, a, vw, ew = b.to_host()
    off = inf["node_off"]
    for i in range(6):
        gx, ga, _, gw = GR.batch_graph(3, i, 8.0, 11.0)
        lo, hi = off[i], off[i + 1]
        assert hi - lo == len(gx) - 1
        e0 = x[lo]
        assert np.array_equal(x[lo:hi + 1] - e0, gx)
        assert np.array_equal(a[e0:x[hi]] - lo, ga)
        want_w = np.ones(len(ga), np.int32) if gw is None else gw
        assert np.array_equal(ew[e0:x[hi]], want_w)
    b.free()


@pytest.mark.parametrize("k,seed,stop", [(8, 5, 5000), (4, 1, 5), (2, 7, 20)])
def test_batch_matches_mirror(k, seed, stop):
    gs = small_batch()
    b = K.Batch.from_graphs(gs)
    part, cuts, mbw, st = b.partition(k=k, seed=seed, stop_factor=stop)
    want, info = mirror_batch(gs, k, 0.03, seed, stop_factor=stop)
    off = info["node_off"]
    for g, (x, a, vw, ew) in enumerate(gs):
        p = part[off[g]:off[g + 1]]
        assert np.array_equal(p, want[g]), g
        assert cuts[g] == O.edge_cut(x, a, ew, p)
        w = O.block_weights(vw, k, p)
        assert mbw[g] == w.max()
    assert st["levels"] == info["levels"]
    b.free()


@pytest.mark.skip(reason="hipGraph-captured LP rounds are disabled: that path hit a host "
                         "segmentation fault under rocprofv3 whose cause was not found")
def test_batch_hipgraph_equals_eager():
    gs = small_batch()
    b = K.Batch.from_graphs(gs)
    p1, c1, _, _ = b.partition(k=4, seed=2, stop_factor=5)
    p2, c2, _, _ = b.partition(k=4, seed=2, stop_factor=5, use_graph=True)
    p3, c3, _, _ = b.partition(k=4, seed=2, stop_factor=5, use_graph=True)
    assert np.array_equal(p1, p2) and np.array_equal(p1, p3)
    assert np.array_equal(c1, c2) and np.array_equal(c1, c3)
    b.free()


def test_generated_batch_balance_and_cuts():
    b = K.Batch.generate(64, 10.0, 13.0, seed=11)
    inf = b.info()
    x, a, vw, ew = b.to_host()
    part, cuts, mbw, st = b.partition(k=8, seed=0)
    off = inf["node_off"]
    for g in range(inf["graphs"]):
        lo, hi = off[g], off[g + 1]
        gx = x[lo:hi + 1] - x[lo]
        ga = a[x[lo]:x[hi]] - lo
        gw = ew[x[lo]:x[hi]]
        p = part[lo:hi]
        assert p.min() >= 0 and p.max() < 8
        assert cuts[g] == O.edge_cut(gx, ga, gw, p)
        U = O.upper_bound(hi - lo, 8, 0.03)
        assert mbw[g] <= U, (g, mbw[g], U)
    b.free()


@pytest.mark.parametrize("name", ["rmat13", "rgg16"])
def test_tile_shapes_keep_results(name):
    """The LP kernels' tile shapes (nodes per lane group, small wave kernel limit) and the wave
    bin change launches only: the partition is the same bits."""
    g = K.Graph.rmat(13, 16 << 13, 4) if name == "rmat13" else K.Graph.rgg(1 << 16, 5)
    want, s0 = g.partition(k=8, seed=1)
    for items, small, wave_hi in ((1, 64, 96), (4, 256, 512), (1, 256, 256), (4, 64, 128)):
        part, st = g.partition(k=8, seed=1, group_items=items, wave_small=small,
                               bins=[8, 16, 32, 64, wave_hi, 0x7fffffff])
        assert np.array_equal(part, want), (items, small, wave_hi)
        assert st["cut"] == s0["cut"]
    g.free()


def test_autotune_tile_shapes_keeps_results():
    g = K.Graph.rmat(13, 16 << 13, 4)
    tuned, ms = K.autotune(g, k=8)
    assert len(ms) == K.AUTOTUNE_EVALS and all(t > 0 for t in ms)
    assert tuned["group_items"] in (1, 2, 4) and tuned["wave_small"] in (64, 128, 256)
    assert tuned["bins"][:4] == [8, 16, 32, 64] and 96 <= tuned["bins"][4] <= 512
    p1, _ = g.partition(k=8, seed=1)
    p2, _ = g.partition(k=8, seed=1, **tuned)
    assert np.array_equal(p1, p2)
    g.free()


def test_autotune_bins_keeps_results():
    g = K.Graph.rmat(13, 16 << 13, 4)
    bins, ms = K.autotune_bins(g, k=8)
    assert bins[:4] == [8, 16, 32, 64] and 96 <= bins[4] <= 512
    assert all(t > 0 for t in ms)
    p1, s1 = g.partition(k=8, seed=1)
    p2, s2 = g.partition(k=8, seed=1, bins=bins)
    assert np.array_equal(p1, p2)
    b = K.Batch.from_graphs(small_batch())
    bins_b, _ = K.autotune_bins(b, k=4)
    q1 = b.partition(k=4, seed=3)[0]
    q2 = b.partition(k=4, seed=3, bins=bins_b)[0]
    assert np.array_equal(q1, q2)
    g.free()
    b.free()
