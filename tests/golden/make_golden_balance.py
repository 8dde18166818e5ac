#!/usr/bin/env python3
"""Writes the reference's kaffpa(FASTSOCIAL) results in its perfectly balanced mode (imbalance 0:
app/balance_configuration.h:41-43, cycle_refinement post-pass interface/kaHIP_interface.cpp:
123-132) into tests/golden/ref_perfect_balance.json: cut and maximum block weight for the
reference's example graphs rgg_n_2_15_s0 and delaunay_n15, k in {2, 7, 32}, seeds 0-4.

The reference is KaHIP v3.21 compiled from its own sources by oracle/Makefile (run where
/root/reference exists: `python tests/golden/make_golden_balance.py`).
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from kahip_amd import graphs  # noqa: E402


def main():
    O.build(with_reference=True)
    recs = []
    for name in ("rgg_n_2_15_s0", "delaunay_n15"):
        x, a, vw, ew = graphs.read_metis(os.path.join(HERE, name + ".graph.gz"))
        for k in (2, 7, 32):
            for seed in range(5):
                cut, part = O.ref_kaffpa(x, a, vw, ew, k, 0.0, seed, 3)
                recs.append({"graph": name, "k": k, "seed": seed, "mode": "FASTSOCIAL",
                             "imbalance": 0.0, "cut": int(cut),
                             "max_block_weight": int(O.block_weights(vw, k, part).max())})
    out = {"reference": "KaHIP v3.21 (README.md:1) compiled from source by oracle/Makefile",
           "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                      text=True).stdout.splitlines()[0],
           "kaffpa": recs}
    with open(os.path.join(HERE, "ref_perfect_balance.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(recs), "records")


if __name__ == "__main__":
    main()
