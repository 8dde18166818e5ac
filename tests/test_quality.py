"""Cut quality of the pipeline against the reference's own kaffpa(FASTSOCIAL) cuts (golden
vectors, 10 seeds per graph and k): the geometric mean of cut(ours) / cut(reference) must stay
within 3 % at eps = 3 %. Ours is the CPU mirror of the GPU pipeline (bit-identical to it, which
the GPU parity tests check), so this runs without a GPU."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, example
from pipeline_mirror import mirror_partition

REF = json.load(open(os.path.join(GOLDEN, "ref_vectors.json")))
CASES = sorted({(r["graph"], r["k"]) for r in REF["kaffpa"] if r["mode"] == "FASTSOCIAL"})


@pytest.mark.slow
@pytest.mark.parametrize("graph,k", CASES)
def test_geomean_cut_within_3_percent_of_reference(graph, k):
    x, a, vw, ew = example(graph)
    ratios = []
    for rec in REF["kaffpa"]:
        if rec["graph"] != graph or rec["k"] != k or rec["mode"] != "FASTSOCIAL":
            continue
        part, info = mirror_partition(x, a, vw, ew, k, 0.03, rec["seed"])
        cut = O.edge_cut(x, a, ew, part)
        assert O.block_weights(vw, k, part).max() <= info["U"]
        ratios.append(cut / rec["cut"])
    gm = math.exp(np.mean(np.log(ratios)))
    print(f"{graph} k={k}: geomean cut ratio vs reference FASTSOCIAL = {gm:.4f} over {len(ratios)} seeds")
    assert len(ratios) == 10
    assert gm <= 1.03


ECO_CASES = sorted({(r["graph"], r["k"]) for r in REF["kaffpa"] if r["mode"] == "ECO"})


@pytest.mark.slow
def test_eco_preset_against_reference_eco():
    """kaffpa(mode=ECO) here is the LP pipeline plus the pairwise FM on the finest level; the
    reference's ECO coarsens by matching and refines by FM. Cuts vs the reference's ECO golden
    cuts (3 seeds per graph and k): 0.75-0.91x on the RGG fixture, 1.05x on delaunay."""
    import kahip_amd as K
    per_case = {}
    for graph, k in ECO_CASES:
        x, a, vw, ew = example(graph)
        ratios = []
        for rec in REF["kaffpa"]:
            if rec["graph"] != graph or rec["k"] != k or rec["mode"] != "ECO":
                continue
            part, info = mirror_partition(x, a, vw, ew, k, 0.03, rec["seed"],
                                          fm_rounds=K.PRESET_FM_ROUNDS)
            assert O.block_weights(vw, k, part).max() <= info["U"]
            ratios.append(O.edge_cut(x, a, ew, part) / rec["cut"])
        per_case[(graph, k)] = math.exp(np.mean(np.log(ratios)))
        print(f"{graph} k={k}: geomean cut ratio vs reference ECO = {per_case[(graph, k)]:.4f}")
    assert max(per_case.values()) <= 1.06
    assert math.exp(np.mean(np.log(list(per_case.values())))) <= 1.0
