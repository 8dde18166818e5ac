"""The oracle's R1 restatement against the reference compiled from its own sources
(oracle/_ref/libkahip_ref.so), on seeded random graphs beyond the golden fixtures. Skipped where
that build is absent."""
import numpy as np
import pytest

import oracle as O
from conftest import random_graph

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference build not present")

CASES = [(80, 0.08, 1, False), (200, 0.03, 2, True), (500, 0.01, 3, True), (64, 0.3, 4, False)]


@pytest.mark.parametrize("n,p,seed,weighted", CASES)
def test_lp_cluster_contract_refine_cut_balance(n, p, seed, weighted):
    x, a, vw, ew = random_graph(n, p, seed, weighted)
    for bound in (3, 10, 1000):
        for s in range(3):
            cl, nc = O.lp_cluster_async(x, a, vw, ew, bound, 10, s)
            rc, rnc = O.ref_lp_cluster(x, a, vw, ew, bound, 10, s)
            assert nc == rnc and np.array_equal(cl, rc)
            got = O.contract(x, a, vw, ew, cl, nc)
            want = O.ref_contract(x, a, vw, ew, cl, nc)
            for g_, w_ in zip(got, want):
                assert np.array_equal(g_, w_)
    W = n if vw is None else int(vw.sum())
    for k in (2, 5, 16):
        U = O.upper_bound(W, k, 0.03)
        assert U == O.ref_upper_bound(n, vw, k, 0.03)
        part = np.random.default_rng(seed + k).integers(0, k, n).astype(np.int32)
        assert O.edge_cut(x, a, ew, part) == O.ref_edge_cut(x, a, ew, part)
        assert O.balance(vw, k, part) == O.ref_balance(x, a, vw, k, part)
        for s in range(2):
            assert np.array_equal(O.lp_refine_async(x, a, vw, ew, k, U, 25, s, part),
                                  O.ref_lp_refine(x, a, vw, ew, k, U, 25, s, part))


def test_coarsening_level_matches_reference():
    """One level of coarsening.cpp:58-80 (LP with B = ceil(U/18) via match(), then contract())
    equals the oracle's LP + contraction with the same bound and seed."""
    from conftest import example
    x, a, vw, ew = example("rgg_n_2_15_s0")
    for k, seed in ((2, 0), (8, 3)):
        mapping, (cx, ca, cw, cv) = O.ref_coarsen_one_level(x, a, vw, ew, k, 0.03, seed)
        U = O.upper_bound(len(x) - 1, k, 0.03)
        cl, nc = O.lp_cluster_async(x, a, vw, ew, O.cluster_bound(U, 18), 10, seed)
        assert np.array_equal(cl, mapping)
        got = O.contract(x, a, vw, ew, cl, nc)
        for g_, w_ in zip(got, (cx, ca, cw, cv)):
            assert np.array_equal(g_, w_)
