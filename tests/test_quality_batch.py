"""Cut quality of the batched mode against the reference's own kaffpa(FASTSOCIAL) on the first 40
graphs of the batched workload (k = 8, eps = 3 %, golden cuts of the reference build): the CPU
mirror of the batched schedule (bit-identical to the GPU batch, tests/test_gpu_batch.py) with the
initial partitioner's default 32 BFS-grown starts per bisection and with the 8 the batched
benchmark uses (2.7x less initial-partitioning time on one core per graph)."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN
from kahip_amd import graphs as GR
from pipeline_mirror import mirror_batch

REF = json.load(open(os.path.join(GOLDEN, "ref_batch_sample.json")))


@pytest.mark.slow
@pytest.mark.parametrize("tries", [-1, 8])
def test_batch_sample_geomean_within_3_percent_of_reference(tries):
    recs = REF["kaffpa"]
    graphs = [GR.batch_graph(REF["batch_seed"], r["graph"]) for r in recs]
    for g, r in zip(graphs, recs):
        assert len(g[0]) - 1 == r["n"] and int(g[0][-1]) == r["m2"]
    parts, info = mirror_batch(graphs, 8, 0.03, 0, ip_tries=tries)
    ratios = []
    for (x, a, vw, ew), p, r, U in zip(graphs, parts, recs, info["U"]):
        assert O.block_weights(vw, 8, p).max() <= U
        ratios.append(O.edge_cut(x, a, ew, p) / r["cut"])
    gm = math.exp(np.mean(np.log(ratios)))
    print(f"batch sample, ip tries {tries}: geomean cut vs reference = {gm:.4f} "
          f"(max {max(ratios):.3f})")
    assert gm <= 1.03
