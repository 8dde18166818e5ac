"""Host-side checks that need no GPU: the C ABI library loads and exports every symbol the
headers declare, the CPU initial partitioner, file formats and the synthetic generators."""
import os
import re

import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from conftest import ROOT, example, manual_graph, random_graph


def declared_symbols():
    names = set()
    for h in ("kaHIP_interface.h", "kahip_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-zA-Z_][a-zA-Z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("kaffpa", "kahip_")):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = K.lib()
    decl = declared_symbols()
    assert {"kaffpa", "kaffpa_balance", "kaffpa_balance_NE", "kaffpa_gpu_ex"} <= decl
    missing = [s for s in sorted(decl) if not hasattr(lib, s)]
    assert missing == []
    assert decl == set(K.EXPORTS)


def test_default_params_are_the_fastsocial_values():
    p = K.default_params()
    assert (p.cluster_rounds, p.refine_rounds, p.cluster_factor, p.stop_factor) == (10, 25, 18,
                                                                                     5000)
    assert list(p.bins) == [8, 16, 32, 64, 512, 0x7FFFFFFF]


@pytest.mark.parametrize("k", [2, 3, 8, 64])
def test_initial_partition_is_valid_and_balanced(k):
    x, a, vw, ew = example("rgg_n_2_15_s0")
    n = len(x) - 1
    U = O.upper_bound(n, k, 0.03)
    p = K.initial_partition(x, a, vw, ew, k, U, -1, 3)
    assert p.min() >= 0 and p.max() < k
    assert O.block_weights(vw, k, p).max() <= U
    # a sane cut: far below a random assignment
    rnd = np.random.default_rng(0).integers(0, k, n).astype(np.int32)
    assert O.edge_cut(x, a, ew, p) < 0.2 * O.edge_cut(x, a, ew, rnd)


def test_initial_partition_deterministic_and_handles_isolated_nodes():
    x, a, vw, ew = random_graph(200, 0.02, 9, weighted=True)
    n = len(x) - 1
    U = O.upper_bound(int(vw.sum()), 4, 0.03)
    p1 = K.initial_partition(x, a, vw, ew, 4, U, 3, 17)
    p2 = K.initial_partition(x, a, vw, ew, 4, U, 3, 17)
    assert np.array_equal(p1, p2)
    assert O.block_weights(vw, 4, p1).max() <= U
    deg = np.diff(x)
    assert np.any(deg == 0)  # the graph has isolated nodes


def test_initial_partition_quality_on_reference_example():
    """k = 2 on the reference's RGG example: within 5 % of the reference's full FASTSOCIAL."""
    x, a, vw, ew = example("rgg_n_2_15_s0")
    n = len(x) - 1
    U = O.upper_bound(n, 2, 0.03)
    cuts = [O.edge_cut(x, a, ew, K.initial_partition(x, a, vw, ew, 2, U, -1, s))
            for s in range(3)]
    assert min(cuts) <= 1.05 * 190


def test_metis_roundtrip(tmp_path):
    x, a, vw, ew = random_graph(50, 0.1, 3, weighted=True)
    f = tmp_path / "g.graph"
    K.write_metis(str(f), x, a, vw, ew)
    x2, a2, vw2, ew2 = K.read_metis(str(f))
    assert all(np.array_equal(p, q) for p, q in ((x, x2), (a, a2), (vw, vw2), (ew, ew2)))
    part = np.arange(50) % 3
    K.write_partition(str(tmp_path / "p"), part)
    assert np.array_equal(K.read_partition(str(tmp_path / "p"), 50), part)


def test_reference_examples_parse_and_are_valid():
    for name, (n, m) in (("rgg_n_2_15_s0", (32768, 160240)), ("delaunay_n15", (32768, 98274)),
                         ("example_weighted", (132, 328))):
        x, a, vw, ew = example(name)
        assert len(x) - 1 == n and len(a) == 2 * m
        K.check_graph(x, a, vw, ew)


def test_check_graph_rejects_malformed_input():
    x, a, _, _ = manual_graph()
    bad = a.copy()
    bad[0] = 7
    with pytest.raises(ValueError):
        K.check_graph(x, bad)
    asym = a.copy()
    asym[0] = 2  # 0->2 without 2->0
    with pytest.raises(ValueError):
        K.check_graph(x, asym)


def test_rgg_generator_radius_rule():
    x, a, _, _ = K.gen_rgg(4096, 3)
    K.check_graph(x, a)
    r = K.graphs.rgg_radius(4096)
    assert abs(r - 0.55 * np.sqrt(np.log(4096) / 4096)) < 1e-15
    avg = len(a) / 4096
    assert 7 < avg < 12  # pi r^2 n minus border effects


def test_rmat_generator_properties():
    x, a, _, w = K.gen_rmat(10, 16 << 10, 5)
    K.check_graph(x, a, None, w)
    u, v = K.graphs.rmat_samples(10, 16 << 10, 5)
    keep = u != v
    assert int(w.sum()) == 2 * int(keep.sum())  # weight = multiplicity, both directions
    deg = np.diff(x)
    assert deg[0] == deg.max()  # a = 0.57 concentrates edges on low ids


def test_bgf_writer_matches_reference_file_and_round_trips(tmp_path, rgg15):
    """The reference ships examples/rgg_n_2_15_s0.bgf; its SHA-256 is
    f5d9828257b35d428d5b8c182e8b58c1c78d51d8b7664a76e0aa164a12cdf81d. Writing the same graph
    (the METIS fixture) must produce that file byte for byte, and reading it back gives the CSR."""
    import hashlib
    from kahip_amd import graphs
    x, a, _, _ = rgg15
    path = tmp_path / "rgg.bgf"
    graphs.write_bgf(str(path), x, a)
    digest = hashlib.sha256(path.read_bytes()).hexdigest()
    assert digest == "f5d9828257b35d428d5b8c182e8b58c1c78d51d8b7664a76e0aa164a12cdf81d"
    rx, ra, _, _ = graphs.read_bgf(str(path))
    assert np.array_equal(rx, x) and np.array_equal(ra, a)


def test_tile_shape_parameters_are_validated(rgg15):
    """group_items and wave_small take only the instantiated tile shapes; anything else is an
    error before any GPU work."""
    x, a, _, _ = rgg15
    assert K.default_params().group_items == 2 and K.default_params().wave_small == 128
    with pytest.raises(K.KahipError, match="group_items"):
        K.partition(x, a, k=2, group_items=3)
    with pytest.raises(K.KahipError, match="wave_small"):
        K.partition(x, a, k=2, wave_small=100)
