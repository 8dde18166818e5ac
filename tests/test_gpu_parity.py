"""GPU parity: every HIP step of the path against the oracle, bit for bit, through the C ABI.

The oracle's R2 functions define the synchronous schedule; R1 functions (pinned against the
reference build and the golden fixtures) define contraction, relabel, cut and projection.
"""
import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from conftest import manual_graph, random_graph
from pipeline_mirror import mirror_partition

pytestmark = pytest.mark.gpu


def star(n_leaves, weighted=False):
    """A hub of degree n_leaves plus a path through the leaves (exercises the hub kernels)."""
    n = n_leaves + 1
    edges = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, n - 1)]
    u = np.array([a for a, b in edges] + [b for a, b in edges])
    v = np.array([b for a, b in edges] + [a for a, b in edges])
    o = np.lexsort((v, u))
    u, v = u[o], v[o]
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, u + 1, 1)
    w = (np.arange(len(u)) % 5 + 1).astype(np.int32) if weighted else None
    return np.cumsum(xadj).astype(np.int32), v.astype(np.int32), None, w


def graphs(rgg15, delaunay15, weighted132):
    out = {
        "manual": manual_graph(),
        "er200w": random_graph(200, 0.05, 1, weighted=True),
        "er500": random_graph(500, 0.02, 2),
        "rgg15": rgg15,
        "delaunay15": delaunay15,
        "weighted132": weighted132,
        "star1500": star(1500),
        "star700w": star(700, weighted=True),
        "rmat12": K.gen_rmat(12, 16 << 12, 7),
        "star10k": star(10000),
        "rmat15": K.gen_rmat(15, 16 << 15, 9),
    }
    return out


@pytest.fixture(scope="module")
def G(rgg15, delaunay15, weighted132):
    return graphs(rgg15, delaunay15, weighted132)


CLUSTER_CASES = [
    ("manual", 3, 0), ("er200w", 10, 1), ("er500", 40, 2), ("rgg15", 938, 0), ("rgg15", 16, 3),
    ("delaunay15", 938, 1), ("weighted132", 40, 4), ("star1500", 200, 5), ("star700w", 2000, 6),
    ("rmat12", 3751, 7), ("rmat12", 30, 8), ("star10k", 300, 9), ("star10k", 20000, 10),
]


@pytest.mark.parametrize("name,bound,seed", CLUSTER_CASES)
@pytest.mark.parametrize("subrounds", [1, 4])
def test_lp_cluster_bit_exact(G, name, bound, seed, subrounds):
    x, a, vw, ew = G[name]
    got, nc, moves = K.lp_cluster(x, a, vw, ew, bound, 10, subrounds, seed)
    want, wnc, wmoves = O.lp_cluster_sync(x, a, vw, ew, bound, 10, subrounds, seed)
    assert nc == wnc
    assert np.array_equal(moves, wmoves)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", ["manual", "er500", "rgg15", "star1500", "rmat12"])
def test_remap_bit_exact(G, name):
    n = len(G[name][0]) - 1
    rng = np.random.default_rng(n)
    for labels in (rng.integers(0, max(n // 3, 1), n), np.arange(n)[::-1].copy(),
                   np.zeros(n, np.int64)):
        got, nc = K.remap(labels)
        want, wnc = O.remap_first_appearance(labels)
        assert nc == wnc and np.array_equal(got, want)


@pytest.mark.parametrize("name,bound,seed", CLUSTER_CASES)
def test_contract_bit_exact(G, name, bound, seed):
    x, a, vw, ew = G[name]
    cl, nc, _ = O.lp_cluster_sync(x, a, vw, ew, bound, 10, 4, seed)
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)


def test_contract_wide_rows(G):
    """Coarse rows with more than 512 distinct neighbours take the global-memory path."""
    x, a, vw, ew = G["star1500"]
    n = len(x) - 1
    cl = np.arange(n, dtype=np.int32)
    cl[1:4] = 0  # hub cluster with three leaves: ~1500 distinct coarse neighbours
    cl, nc = O.remap_first_appearance(cl)
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)
    x, a, vw, ew = G["star10k"]
    n = len(x) - 1
    cl = np.arange(n, dtype=np.int32)
    cl[1:4] = 0  # ~10000 distinct coarse neighbours: beyond the LDS tiers
    cl, nc = O.remap_first_appearance(cl)
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)
    x, a, vw, ew = G["rmat12"]
    n = len(x) - 1
    rng = np.random.default_rng(3)
    cl, nc = O.remap_first_appearance(rng.integers(0, 40, n))  # a few huge clusters
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)


@pytest.mark.parametrize("case", ["mod3", "hub_groups", "big_hub_cluster", "many_buckets"])
def test_contract_chunked_tier(case):
    """Rows with more than 32768 member edges go straight to the chunked, bucketed tier; rows
    with many distinct neighbours fall into it after the LDS tiers."""
    if case in ("mod3", "hub_groups", "big_hub_cluster"):
        x, a, vw, ew = K.gen_rmat(16, 16 << 16, 11)
    else:
        x, a, vw, ew = K.gen_rmat(18, 16 << 18, 12)
    n = len(x) - 1
    if case == "mod3":
        cl = (np.arange(n) % 3).astype(np.int32)  # three huge rows, 3 distinct neighbours each
    elif case == "hub_groups":
        cl = (np.arange(n) >> 4).astype(np.int32)  # 16-node groups
    elif case == "big_hub_cluster":
        deg = np.diff(x)
        cl = np.arange(n, dtype=np.int32)
        cl[np.argsort(-deg, kind="stable")[:64]] = int(np.argmax(deg))  # top hubs together
    else:
        rng = np.random.default_rng(5)
        cl = np.arange(n, dtype=np.int32)
        cl[rng.random(n) < 0.7] = 0  # one cluster of 70 % of the nodes: > 1024 chunks, capped buckets
    cl, nc = O.remap_first_appearance(cl)
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)


def sparse_random_graph(n, m, seed):
    """Random simple graph with about m undirected edges, weighted, CSR with sorted rows."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, 2 * m)
    v = rng.integers(0, n, 2 * m)
    keep = u != v
    u, v = np.minimum(u[keep], v[keep]), np.maximum(u[keep], v[keep])
    key = np.unique(u.astype(np.int64) * n + v)[:m]
    u, v = (key // n).astype(np.int64), (key % n).astype(np.int64)
    w = rng.integers(1, 9, len(u))
    uu, vv, ww = np.concatenate([u, v]), np.concatenate([v, u]), np.concatenate([w, w])
    order = np.lexsort((vv, uu))
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, uu[order] + 1, 1)
    return (np.cumsum(xadj).astype(np.int32), vv[order].astype(np.int32),
            rng.integers(1, 4, n).astype(np.int32), ww[order].astype(np.int32))


def test_contract_lds_tiers():
    """Coarse rows of many low-degree members whose distinct neighbours exceed the wave tier:
    <= 1024 distinct (256-thread tier), <= 4096 (1024-thread tier, several member batches) and
    beyond (global tier), all in one clustering."""
    x, a, vw, ew = sparse_random_graph(40000, 160000, 21)  # average degree 8
    n = len(x) - 1
    rng = np.random.default_rng(22)
    cl = np.arange(n, dtype=np.int32)
    perm = rng.permutation(n)
    pos = 0
    for size in (40, 90, 120, 250, 400, 520, 700, 2000):  # summed degrees ~8 x size
        cl[perm[pos:pos + size]] = perm[pos]
        pos += size
    cl, nc = O.remap_first_appearance(cl)
    got = K.contract(x, a, vw, ew, cl, nc)
    want = O.contract(x, a, vw, ew, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)
    # one row of 1500 members of degree 2 into 600 targets: ~600 distinct, 3000 member edges,
    # more members than the 256-thread tier stages at once (two batches)
    rng = np.random.default_rng(23)
    t = 1500 + rng.integers(0, 600, (1500, 2))
    t[:, 1] = np.where(t[:, 1] == t[:, 0], 1500 + (t[:, 0] - 1500 + 1) % 600, t[:, 1])
    u = np.repeat(np.arange(1500), 2)
    v = t.reshape(-1)
    uu, vv = np.concatenate([u, v]), np.concatenate([v, u])
    order = np.lexsort((vv, uu))
    xs = np.zeros(2101, np.int64)
    np.add.at(xs, uu[order] + 1, 1)
    xs = np.cumsum(xs).astype(np.int32)
    as_ = vv[order].astype(np.int32)
    cl = np.arange(2100, dtype=np.int32)
    cl[:1500] = 0
    cl, nc = O.remap_first_appearance(cl)
    got = K.contract(xs, as_, None, None, cl, nc)
    want = O.contract(xs, as_, None, None, cl, nc)
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)


REFINE_CASES = [("er200w", 2, 0), ("rgg15", 2, 1), ("rgg15", 64, 2), ("delaunay15", 8, 3),
                ("weighted132", 4, 4), ("star1500", 4, 5), ("rmat12", 16, 6), ("star10k", 4, 7),
                ("star10k", 5000, 8), ("rmat12", 3000, 9), ("rmat15", 64, 10), ("rmat15", 8, 11)]


@pytest.mark.parametrize("name,k,seed", REFINE_CASES)
@pytest.mark.parametrize("start", ["random", "blocks", "overloaded"])
def test_lp_refine_bit_exact(G, name, k, seed, start):
    x, a, vw, ew = G[name]
    n = len(x) - 1
    W = n if vw is None else int(vw.sum())
    U = O.upper_bound(W, k, 0.03)
    rng = np.random.default_rng(seed)
    if start == "random":
        part = rng.integers(0, k, n).astype(np.int32)
    elif start == "blocks":
        part = (np.arange(n) * k // max(n, 1)).astype(np.int32)
    else:
        part = np.where(rng.random(n) < 0.7, 0, rng.integers(0, k, n)).astype(np.int32)
    got, moves = K.lp_refine(x, a, vw, ew, k, U, 25, 4, seed, part)
    want, wmoves = O.lp_refine_sync(x, a, vw, ew, k, U, 25, 4, seed, part)
    assert np.array_equal(moves, wmoves)
    assert np.array_equal(got, want)
    if start != "overloaded":
        assert O.edge_cut(x, a, ew, got) <= O.edge_cut(x, a, ew, part)


@pytest.mark.parametrize("name", ["manual", "er200w", "rgg15", "weighted132", "rmat12"])
def test_edge_cut_block_weights_project(G, name):
    x, a, vw, ew = G[name]
    n = len(x) - 1
    rng = np.random.default_rng(11)
    for k in (1, 2, 7, 64):
        part = rng.integers(0, k, n).astype(np.int32)
        assert K.edge_cut(x, a, ew, part) == O.edge_cut(x, a, ew, part)
        assert np.array_equal(K.block_weights(vw, k, part), O.block_weights(vw, k, part))
    mapping = rng.integers(0, max(n // 4, 1), n).astype(np.int32)
    cpart = rng.integers(0, 5, max(n // 4, 1)).astype(np.int32)
    assert np.array_equal(K.project(mapping, cpart), O.project(mapping, cpart))


@pytest.mark.parametrize("name,k,seed", [("rgg15", 2, 0), ("rgg15", 8, 1), ("delaunay15", 4, 2),
                                         ("weighted132", 2, 3), ("rmat12", 8, 4),
                                         ("manual", 2, 0)])
def test_pipeline_matches_mirror(G, name, k, seed):
    x, a, vw, ew = G[name]
    part, st = K.partition(x, a, vw, ew, k=k, imbalance=0.03, seed=seed)
    want, info = mirror_partition(x, a, vw, ew, k, 0.03, seed)
    assert st["levels"] == info["levels"]
    assert st["level_n"] == info["level_n"]
    assert np.array_equal(part, want)
    assert st["cut"] == O.edge_cut(x, a, ew, want)
    assert st["upper_bound"] == info["U"]


@pytest.mark.parametrize("name,k,seed,reorder", [("rmat12", 8, 4, 0), ("rgg15", 2, 0, 1),
                                                  ("star1500", 4, 1, -1), ("er200w", 2, 3, 1)])
def test_pipeline_reorder_modes_match_mirror(G, name, k, seed, reorder):
    x, a, vw, ew = G[name]
    part, st = K.partition(x, a, vw, ew, k=k, imbalance=0.03, seed=seed, reorder=reorder)
    want, info = mirror_partition(x, a, vw, ew, k, 0.03, seed, reorder=reorder)
    assert st["reordered"] == info["reordered"]
    assert st["level_n"] == info["level_n"]
    assert np.array_equal(part, want)
    assert st["cut"] == O.edge_cut(x, a, ew, want)


@pytest.mark.parametrize("name,k,seed,ens", [("rgg15", 4, 1, 3), ("rmat12", 8, 2, 2),
                                              ("weighted132", 2, 0, 4)])
def test_pipeline_ensemble_matches_mirror(G, name, k, seed, ens):
    x, a, vw, ew = G[name]
    part, st = K.partition(x, a, vw, ew, k=k, imbalance=0.03, seed=seed, ensemble=ens)
    want, info = mirror_partition(x, a, vw, ew, k, 0.03, seed, ensemble=ens)
    assert st["level_n"] == info["level_n"]
    assert np.array_equal(part, want)


def test_kaffpa_strongsocial_uses_ensembles(G):
    x, a, vw, ew = G["rgg15"]
    cut, blocks = K.kaffpa(vw, x, ew, a, 16, 0.03, True, 3, K.STRONGSOCIAL)
    part, st = K.partition(x, a, vw, ew, k=16, imbalance=0.03, seed=3, ensemble=7,
                           fm_rounds=K.PRESET_FM_ROUNDS)
    assert np.array_equal(np.array(blocks), part) and cut == st["cut"]


def test_kaffpa_eco_adds_pairwise_fm(G):
    x, a, vw, ew = G["delaunay15"]
    cut, blocks = K.kaffpa(vw, x, ew, a, 8, 0.03, True, 2, K.ECO)
    part, st = K.partition(x, a, vw, ew, k=8, imbalance=0.03, seed=2,
                           fm_rounds=K.PRESET_FM_ROUNDS)
    assert np.array_equal(np.array(blocks), part) and cut == st["cut"]
    assert cut == O.edge_cut(x, a, ew, part)
    # the FM runs on every level (its gains there carry to the finer levels, where the LP
    # refinement continues from a different partition): gains are non-negative and it costs time
    assert st["fm_gain"] > 0 and st["ms_fm"] > 0
    _, plain = K.partition(x, a, vw, ew, k=8, imbalance=0.03, seed=2)
    assert plain["fm_gain"] == 0 and plain["ms_fm"] == 0


@pytest.mark.parametrize("name,k,seed", [("delaunay15", 8, 1), ("rgg15", 2, 2), ("rmat12", 16, 3),
                                         ("weighted132", 4, 4)])
def test_pipeline_fm_matches_mirror(G, name, k, seed):
    """The finest-level pairwise FM (in the renumbered ids for skewed inputs) is bit-identical
    to the mirror's call of the same CPU routine."""
    x, a, vw, ew = G[name]
    got, st = K.partition(x, a, vw, ew, k=k, imbalance=0.03, seed=seed, fm_rounds=3)
    want, info = mirror_partition(x, a, vw, ew, k, 0.03, seed, fm_rounds=3)
    assert np.array_equal(got, want)
    assert st["cut"] == O.edge_cut(x, a, ew, want)
    assert O.block_weights(vw, k, got).max() <= info["U"]


def test_graph_capture_options_are_rejected(G):
    x, a, vw, ew = G["manual"]
    with pytest.raises(K.KahipError, match="disabled"):
        K.partition(x, a, vw, ew, k=2, graph_refine=1)


def test_pipeline_deterministic(G):
    x, a, vw, ew = G["rmat12"]
    p1, s1 = K.partition(x, a, vw, ew, k=16, seed=5)
    p2, s2 = K.partition(x, a, vw, ew, k=16, seed=5)
    assert np.array_equal(p1, p2) and s1["cut"] == s2["cut"]


def test_cached_workspaces_are_reused_across_graphs(G):
    """The partition entry points keep each device's engine (stream and device workspaces)
    between calls: a partition equals the mirror's before and after larger and smaller graphs ran
    on the same workspaces, after a failed call (which drops the cached engine) and after
    release_workspaces."""
    x, a, vw, ew = G["rmat12"]
    want, _ = mirror_partition(x, a, vw, ew, 16, 0.03, 5)

    def check():
        got, st = K.partition(x, a, vw, ew, k=16, imbalance=0.03, seed=5)
        assert np.array_equal(got, want)
        assert st["cut"] == O.edge_cut(x, a, ew, want)

    check()
    for name, k in (("rgg15", 64), ("manual", 2), ("delaunay15", 8)):
        gx, ga, gvw, gew = G[name]
        K.partition(gx, ga, gvw, gew, k=k, imbalance=0.03, seed=1)
        check()
    with pytest.raises(K.KahipError):
        K.partition(x, a, vw, ew, k=0)
    check()
    K.release_workspaces()
    check()
    check()


def test_concurrent_partitions_use_separate_engines(G):
    """Two threads partitioning at once: one holds the device's cached engine, the other runs on
    an engine of its own (the call releases the GIL); both give the mirror's partition."""
    import threading
    x1, a1, vw1, ew1 = G["rmat12"]
    x2, a2, vw2, ew2 = G["rgg15"]
    want1, _ = mirror_partition(x1, a1, vw1, ew1, 16, 0.03, 5)
    want2, _ = mirror_partition(x2, a2, vw2, ew2, 8, 0.03, 1)
    got = {}

    def work(key, args, k, seed, reps=4):
        for i in range(reps):
            got[(key, i)] = K.partition(*args, k=k, imbalance=0.03, seed=seed)[0]

    t1 = threading.Thread(target=work, args=(1, (x1, a1, vw1, ew1), 16, 5))
    t2 = threading.Thread(target=work, args=(2, (x2, a2, vw2, ew2), 8, 1))
    t1.start()
    t2.start()
    t1.join()
    t2.join()
    for i in range(4):
        assert np.array_equal(got[(1, i)], want1)
        assert np.array_equal(got[(2, i)], want2)


def test_profile_modes(G):
    """profile=1 times every launch; profile=2 only the phases and the clustering proposal
    kernels (same partition either way); other values are rejected."""
    x, a, vw, ew = G["rmat12"]
    want, _ = K.partition(x, a, vw, ew, k=16, seed=5)
    full, _ = K.partition(x, a, vw, ew, k=16, seed=5, profile=1)
    full_names = set(K.profile())
    light, _ = K.partition(x, a, vw, ew, k=16, seed=5, profile=2)
    light_names = set(K.profile())
    assert np.array_equal(full, want) and np.array_equal(light, want)
    allowed = {"lp_cluster_total", "contract", "remap", "reorder", "edge_cut"}
    proposals = {n for n in light_names if n.startswith("lp_cluster_propose_")}
    assert {"lp_cluster_total", "contract", "edge_cut"} <= light_names
    assert light_names - proposals <= allowed and "lp_cluster_propose_hub_merge" not in proposals
    assert proposals == {n for n in full_names if n.startswith("lp_cluster_propose_")} - {
        "lp_cluster_propose_hub_merge"}
    assert light_names < full_names
    with pytest.raises(K.KahipError, match="profile"):
        K.partition(x, a, vw, ew, k=16, seed=5, profile=3)


GRAPH_OFF = ("hipGraph-captured LP rounds are disabled: that path hit a host segmentation fault "
             "under rocprofv3 whose cause was not found")


@pytest.mark.skip(reason=GRAPH_OFF)
@pytest.mark.parametrize("name,k,seed", [("rgg15", 8, 1), ("rmat12", 16, 3), ("star1500", 4, 0)])
def test_graph_captured_refinement_matches_mirror(G, name, k, seed):
    """graph_refine = 1 replays each level's refinement rounds from a captured hipGraph; the
    partition is the mirror's, as with direct launches, and stays so on a second call."""
    x, a, vw, ew = G[name]
    want, _ = mirror_partition(x, a, vw, ew, k, 0.03, seed)
    for _ in range(2):
        part, _ = K.partition(x, a, vw, ew, k=k, imbalance=0.03, seed=seed, graph_refine=1)
        assert np.array_equal(part, want)


@pytest.mark.parametrize("scale,samples,seed", [(10, 16 << 10, 1), (12, 8 << 12, 2)])
def test_rmat_generator_matches_numpy(scale, samples, seed):
    g = K.Graph.rmat(scale, samples, seed)
    x, a, vw, ew = g.to_host()
    wx, wa, _, ww = K.gen_rmat(scale, samples, seed)
    assert np.array_equal(x, wx) and np.array_equal(a, wa) and np.array_equal(ew, ww)
    assert np.all(vw == 1)
    g.free()


@pytest.mark.parametrize("n,seed", [(2, 0), (1000, 1), (32768, 2), (100000, 3)])
def test_rgg_generator_matches_numpy(n, seed):
    g = K.Graph.rgg(n, seed)
    x, a, vw, ew = g.to_host()
    wx, wa, _, _ = K.gen_rgg(n, seed)
    assert np.array_equal(x, wx) and np.array_equal(a, wa)
    assert np.all(vw == 1) and np.all(ew == 1)
    assert g.info()["max_degree"] == int(np.diff(wx).max())
    g.free()


def test_device_graph_partition_roundtrip(G):
    x, a, vw, ew = G["rgg15"]
    g = K.Graph.from_host(x, a, vw, ew)
    part, st = g.partition(k=4, seed=3)
    want, _ = K.partition(x, a, vw, ew, k=4, seed=3)
    assert np.array_equal(part, want)
    g.free()


def test_kaffpa_dropin_manual_graph():
    x, a, _, _ = manual_graph()
    cut, blocks = K.kaffpa(None, x, None, a, 2, 0.03, True, 0, K.FASTSOCIAL)
    assert cut == O.edge_cut(x, a, None, np.array(blocks, np.int32))
    assert sorted(set(blocks)) == [0, 1]


def test_kaffpa_void_symbol(rgg15):
    import ctypes as C
    x, a, _, _ = rgg15
    n = C.c_int(len(x) - 1)
    k = C.c_int(2)
    imb = C.c_double(0.03)
    cut = C.c_int(0)
    part = np.zeros(len(x) - 1, np.int32)
    ip = C.POINTER(C.c_int)
    K.lib().kaffpa(C.byref(n), None, x.ctypes.data_as(ip), None, a.ctypes.data_as(ip),
                   C.byref(k), C.byref(imb), True, 0, K.FASTSOCIAL, C.byref(cut),
                   part.ctypes.data_as(ip))
    assert cut.value == O.edge_cut(x, a, None, part)
    assert O.balance(None, 2, part) <= 1.03 + 1e-9


@pytest.mark.parametrize("case", ["empty", "edgeless", "k1", "k_gt_n"])
def test_edge_cases(case):
    if case == "empty":
        x, a = np.zeros(1, np.int32), np.zeros(0, np.int32)
        part, st = K.partition(x, a, k=2)
        assert len(part) == 0 and st["cut"] == 0
    elif case == "edgeless":
        x, a = np.zeros(11, np.int32), np.zeros(0, np.int32)
        part, st = K.partition(x, a, k=3)
        assert st["cut"] == 0 and np.bincount(part, minlength=3).max() <= 4
        cl, nc, _ = K.lp_cluster(x, a, None, None, 5, 10, 4, 0)
        assert nc == 10 and np.array_equal(cl, np.arange(10))
    elif case == "k1":
        x, a, _, _ = manual_graph()
        part, st = K.partition(x, a, k=1)
        assert st["cut"] == 0 and np.all(part == 0)
    else:
        x, a, _, _ = manual_graph()
        part, st = K.partition(x, a, k=8)
        assert part.min() >= 0 and part.max() < 8
        assert st["cut"] == O.edge_cut(x, a, None, part)
