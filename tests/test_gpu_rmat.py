"""Parity and quality on the skewed R-MAT workloads at the sizes the skewed-degree kernel tiers run
at, through the C ABI:

- the full R-MAT scale-24 graph (2^28 samples, (0.57, 0.19, 0.19, 0.05), k = 256, eps = 3 %):
  the partition's cut recomputed by the oracle, block weights within U, labels in range, and two
  calls bit-identical (the whole graph is too large for the CPU mirror; every step it runs is
  bit-exact against the oracle at the scales below);
- R-MAT scale 20 (k = 256 bounds) after the degree renumbering: level-0 clustering and
  contraction bit-exact against the oracle, with multi-chunk hubs whose bucket merges use the
  256- and 512-thread tiers;
- two giant stars (6 M and 18 M leaves): hubs whose first-round buckets exceed 2048 and 8192
  entries, i.e. the 1024-thread merge tier and the per-hub global table, bit-exact;
- cut quality against the reference's own kaffpa(FASTSOCIAL) on RGG 2^20 (k = 64) and R-MAT
  scales 16-18 (k = 256), 10 seeds each (golden cuts of the reference build).
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from conftest import GOLDEN
from pipeline_mirror import reorder_by_degree

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def rmat24():
    _log("generating R-MAT scale 24, 2^28 samples, seed 1 on the GPU")
    g = K.Graph.rmat(24, 1 << 28, 1)
    host = g.to_host()
    yield g, host
    g.free()


def test_rmat24_k256_partition_properties(rmat24):
    g, (x, a, vw, ew) = rmat24
    n = len(x) - 1
    k = 256
    U = O.upper_bound(n, k, 0.03)
    _log("partitioning R-MAT 2^24, k = 256, seed 0 (first call)")
    p1, s1 = g.partition(k=k, seed=0)
    assert p1.min() >= 0 and p1.max() < k
    assert s1["reordered"]
    assert s1["cut"] == O.edge_cut(x, a, ew, p1)
    w = O.block_weights(None, k, p1)
    assert w.max() == s1["max_block_weight"] <= U
    assert s1["balance"] <= 1.03 + 1e-9
    _log(f"cut {s1['cut']} balance {s1['balance']:.4f}; second call")
    p2, s2 = g.partition(k=k, seed=0)
    assert s2["cut"] == s1["cut"]
    assert np.array_equal(p1, p2)


def test_rmat20_level0_cluster_and_contract_bit_exact():
    _log("R-MAT scale 20, 2^24 samples: level-0 clustering + contraction")
    g = K.Graph.rmat(20, 1 << 24, 3)
    x, a, vw, ew = g.to_host()
    g.free()
    (x, a, vw, ew), _ = reorder_by_degree(x, a, None, ew)
    n = len(x) - 1
    deg = np.diff(x.astype(np.int64))
    assert (deg > 4096).sum() >= 8  # multi-chunk hubs: bucketed partial merges
    U = O.upper_bound(n, 256, 0.03)
    B = O.cluster_bound(U, 18)
    seed = O.hash3(3, 0xC1, 0)
    cl, nc, moves = K.lp_cluster(x, a, None, ew, B, 10, 4, seed)
    want, wnc, wmoves = O.lp_cluster_sync(x, a, None, ew, B, 10, 4, seed)
    assert nc == wnc
    assert np.array_equal(moves, wmoves)
    assert np.array_equal(cl, want)
    got = K.contract(x, a, None, ew, cl, nc)
    exp = O.contract(x, a, None, ew, want, wnc)
    for g_, e_ in zip(got, exp):
        assert np.array_equal(g_, e_)


def two_stars(leaves_a, leaves_b):
    """Hubs 0 and 1 with leaves_a and leaves_b leaves (ids 2.. in that order) and a path through
    all leaves; rows sorted by neighbour id."""
    L = leaves_a + leaves_b
    n = L + 2
    leaf = np.arange(2, n, dtype=np.int64)
    ldeg = 1 + (leaf > 2) + (leaf < n - 1)
    deg = np.concatenate([[leaves_a, leaves_b], ldeg])
    x = np.zeros(n + 1, np.int64)
    x[1:] = np.cumsum(deg)
    a = np.empty(int(x[-1]), np.int32)
    a[:leaves_a] = np.arange(2, 2 + leaves_a)
    a[leaves_a:L] = np.arange(2 + leaves_a, n)
    base = x[2:-1]
    a[base] = np.where(leaf < 2 + leaves_a, 0, 1)
    has_prev = leaf > 2
    a[base[has_prev] + 1] = leaf[has_prev] - 1
    nxt = base + 1 + has_prev
    has_next = leaf < n - 1
    a[nxt[has_next]] = leaf[has_next] + 1
    return x.astype(np.int32), a, None, None


def test_giant_hub_merge_tiers_bit_exact():
    # first-round buckets of 2048 buckets per hub: 6M / 2048 ~ 2930 entries (1024-thread tier),
    # 18M / 2048 ~ 8790 entries (past the 8192-entry LDS limit: the per-hub global table)
    _log("two stars (6M + 18M leaves): clustering + contraction")
    x, a, vw, ew = two_stars(6_000_000, 18_000_000)
    # one sub-round: the hubs see every leaf's own label in the first round
    for bound, subrounds, seed in ((1000, 1, 11), (40_000, 4, 12)):
        cl, nc, moves = K.lp_cluster(x, a, None, None, bound, 3, subrounds, seed)
        want, wnc, wmoves = O.lp_cluster_sync(x, a, None, None, bound, 3, subrounds, seed)
        assert nc == wnc
        assert np.array_equal(moves, wmoves)
        assert np.array_equal(cl, want)
    got = K.contract(x, a, None, None, cl, nc)
    exp = O.contract(x, a, None, None, want, wnc)
    for g_, e_ in zip(got, exp):
        assert np.array_equal(g_, e_)


LARGE_PATH = os.path.join(GOLDEN, "ref_kaffpa_large.json")  # its presence is a CPU test
LARGE = json.load(open(LARGE_PATH)) if os.path.exists(LARGE_PATH) else {"graphs": [], "kaffpa": []}
LARGE_CASES = sorted({(r["kind"], r["size"], r["k"]) for r in LARGE["kaffpa"]})


def _graph_sha(x, a, ew):
    h = hashlib.sha256()
    for arr in (x, a) + ((ew,) if ew is not None else ()):
        h.update(np.ascontiguousarray(arr, np.int32).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("kind,size,k", LARGE_CASES)
def test_large_geomean_cut_within_3_percent_of_reference(kind, size, k):
    meta = next(m for m in LARGE["graphs"] if m["kind"] == kind and m["size"] == size)
    if kind == "rgg":
        g = K.Graph.rgg(1 << size, meta["graph_seed"])
    else:
        g = K.Graph.rmat(size, 16 << size, meta["graph_seed"])
    x, a, vw, ew = g.to_host()
    if kind == "rgg":
        ew = None
    assert _graph_sha(x, a, ew) == meta["sha256"]
    U = O.upper_bound(len(x) - 1, k, 0.03)
    ratios = []
    for rec in LARGE["kaffpa"]:
        if (rec["kind"], rec["size"], rec["k"]) != (kind, size, k):
            continue
        part, st = g.partition(k=k, seed=rec["seed"])
        assert st["cut"] == O.edge_cut(x, a, ew, part)
        assert O.block_weights(None, k, part).max() <= U
        ratios.append(st["cut"] / rec["cut"])
    g.free()
    gm = math.exp(np.mean(np.log(ratios)))
    _log(f"{kind} {size} k={k}: geomean cut ratio vs reference FASTSOCIAL = {gm:.4f} "
         f"over {len(ratios)} seeds")
    assert len(ratios) == 10
    assert gm <= 1.03
