"""The oracle's R1 restatement against the reference's golden vectors (tests/golden/, written by
make_golden.py from the reference compiled from its own sources). Runs without the reference."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, example

with open(os.path.join(GOLDEN, "ref_vectors.json")) as f:
    REF = json.load(f)
ARR = np.load(os.path.join(GOLDEN, "ref_arrays.npz"))

_graphs = {}


def graph(name):
    if name not in _graphs:
        _graphs[name] = example(name)
    return _graphs[name]


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


@pytest.mark.parametrize("rec", REF["lp_cluster"],
                         ids=lambda r: f"{r['graph']}-B{r['bound']}-s{r['seed']}")
def test_lp_cluster_async_matches_reference(rec):
    x, a, vw, ew = graph(rec["graph"])
    cl, nc = O.lp_cluster_async(x, a, vw, ew, rec["bound"], rec["iterations"], rec["seed"])
    assert nc == rec["clusters"]
    assert O.edge_cut(x, a, ew, cl) == rec["cut"]
    assert h(cl) == rec["sha256"]


def test_lp_cluster_full_arrays():
    x, a, vw, ew = graph("rgg_n_2_15_s0")
    for seed in (0, 1):
        cl, _ = O.lp_cluster_async(x, a, vw, ew, 938, 10, seed)
        assert np.array_equal(cl, ARR[f"cluster_rgg_n_2_15_s0_938_{seed}"])


@pytest.mark.parametrize("rec", REF["contract"],
                         ids=lambda r: f"{r['graph']}-B{r['bound']}-s{r['seed']}")
def test_contract_matches_reference(rec):
    x, a, vw, ew = graph(rec["graph"])
    cl, nc = O.lp_cluster_async(x, a, vw, ew, rec["bound"], 10, rec["seed"])
    assert nc == rec["nc"]
    cx, ca, cw, cv = O.contract(x, a, vw, ew, cl, nc)
    assert int(cx[-1]) == rec["mc"]
    assert (h(cx), h(ca), h(cw), h(cv)) == (rec["xadj"], rec["adjncy"], rec["adjwgt"], rec["vwgt"])


def test_contract_weighted_full_arrays():
    x, a, vw, ew = graph("example_weighted")
    cl = ARR["contract_weighted_cluster"]
    nc = int(cl.max()) + 1
    got = O.contract(x, a, vw, ew, cl, nc)
    for g_, key in zip(got, ("xadj", "adjncy", "adjwgt", "vwgt")):
        assert np.array_equal(g_, ARR[f"contract_weighted_{key}"])


@pytest.mark.parametrize("rec", REF["lp_refine"],
                         ids=lambda r: f"{r['graph']}-k{r['k']}-s{r['seed']}")
def test_lp_refine_async_matches_reference(rec):
    x, a, vw, ew = graph(rec["graph"])
    n = len(x) - 1
    k = rec["k"]
    start = (np.arange(n) * k // n).astype(np.int32)
    p = O.lp_refine_async(x, a, vw, ew, k, rec["bound"], rec["rounds"], rec["seed"], start)
    assert O.edge_cut(x, a, ew, p) == rec["cut"]
    assert h(p) == rec["sha256"]


@pytest.mark.parametrize("rec", REF["metrics"], ids=lambda r: f"{r['graph']}-k{r['k']}")
def test_cut_and_balance_match_reference(rec):
    x, a, vw, ew = graph(rec["graph"])
    n = len(x) - 1
    k = rec["k"]
    part = np.random.default_rng(k).integers(0, k, n).astype(np.int32)
    assert O.edge_cut(x, a, ew, part) == rec["cut"]
    assert O.balance(vw, k, part) == rec["balance"]


@pytest.mark.parametrize("rec", REF["upper_bound"], ids=lambda r: f"n{r['n']}-k{r['k']}")
def test_upper_bound_matches_reference(rec):
    assert O.upper_bound(rec["W"], rec["k"], rec["imbalance"]) == rec["U"]


def test_documented_bounds_table():
    """U and B of the benchmark configurations (app/balance_configuration.h:40-46,
    size_constraint_label_propagation.cpp:57) as tabulated for the workloads."""
    for n, k, U, B in ((32768, 2, 16875, 938), (4194304, 64, 67502, 3751),
                       (16777216, 256, 67502, 3751), (65536, 8, 8437, 469), (1024, 8, 131, 8)):
        assert O.upper_bound(n, k, 0.03) == U
        assert O.cluster_bound(U, 18) == B


def test_golden_kaffpa_records_are_balanced():
    """The reference's own FASTSOCIAL results respect the block bound; they are the quality
    targets of the GPU pipeline (tests/test_quality.py)."""
    for rec in REF["kaffpa"]:
        if rec["mode"] == "FASTSOCIAL":
            assert rec["balance"] <= 1.03 + 1e-9


def test_large_golden_cuts_present_and_graphs_reproducible():
    """tests/golden/ref_kaffpa_large.json (reference kaffpa FASTSOCIAL on RGG 2^20 k = 64 and
    R-MAT 2^16..2^18 k = 256, seeds 0-9) exists, holds 10 balanced records per case, and its
    smallest graph regenerates to the recorded CSR hash here (the GPU test checks the rest)."""
    import hashlib
    import json
    import os
    from conftest import GOLDEN
    from kahip_amd import graphs
    d = json.load(open(os.path.join(GOLDEN, "ref_kaffpa_large.json")))
    cases = {}
    for r in d["kaffpa"]:
        cases.setdefault((r["kind"], r["size"], r["k"]), []).append(r["seed"])
        assert r["mode"] == "FASTSOCIAL" and r["balance"] <= 1.03 + 1e-9
    assert set(cases) == {("rgg", 20, 64), ("rmat", 16, 256), ("rmat", 17, 256),
                          ("rmat", 18, 256)}
    assert all(sorted(s) == list(range(10)) for s in cases.values())
    meta = next(m for m in d["graphs"] if (m["kind"], m["size"]) == ("rmat", 16))
    x, a, _, ew = graphs.gen_rmat(16, 16 << 16, meta["graph_seed"])
    h = hashlib.sha256()
    for arr in (x, a, ew):
        h.update(np.ascontiguousarray(arr, np.int32).tobytes())
    assert h.hexdigest() == meta["sha256"] and len(x) - 1 == meta["n"] and x[-1] == meta["m2"]
