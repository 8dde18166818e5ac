"""Full-size checks on the 2^22-node random geometric graph (k = 64): the GPU clustering and
contraction of level 0 against the oracle (bit for bit), and the final partition's cut and
balance recomputed by the oracle."""
import numpy as np
import pytest

import kahip_amd as K
import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N = 1 << 22


@pytest.fixture(scope="module")
def rgg22():
    g = K.Graph.rgg(N, 1)
    host = g.to_host()
    yield g, host
    g.free()


def test_rgg22_level0_cluster_and_contract_bit_exact(rgg22):
    _, (x, a, vw, ew) = rgg22
    U = O.upper_bound(N, 64, 0.03)
    B = O.cluster_bound(U, 18)
    seed = O.hash3(1, 0xC1, 0)
    cl, nc, _ = K.lp_cluster(x, a, None, None, B, 10, 4, seed)
    want, wnc, _ = O.lp_cluster_sync(x, a, None, None, B, 10, 4, seed)
    assert nc == wnc and np.array_equal(cl, want)
    got = K.contract(x, a, None, None, cl, nc)
    exp = O.contract(x, a, None, None, want, wnc)
    for g_, e_ in zip(got, exp):
        assert np.array_equal(g_, e_)


def test_rgg22_partition_cut_and_balance(rgg22):
    g, (x, a, vw, ew) = rgg22
    part, st = g.partition(k=64, seed=1)
    assert st["cut"] == O.edge_cut(x, a, None, part)
    w = O.block_weights(None, 64, part)
    assert w.max() == st["max_block_weight"] <= O.upper_bound(N, 64, 0.03)
    assert st["balance"] <= 1.03 + 1e-9
    assert not st["reordered"]
