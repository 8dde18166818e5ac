#!/usr/bin/env python3
"""Writes the reference's kaffpa(FASTSOCIAL) cuts for the first 40 graphs of the batched
workload (kahip_amd.graphs.batch_graph(1, i): RGG for even i, R-MAT for odd i, n log-uniform in
[2^10, 2^16]), k = 8, eps = 3 %, seed 0, into tests/golden/ref_batch_sample.json.

The reference is KaHIP v3.21 compiled from its own sources by oracle/Makefile (run where
/root/reference exists: `python tests/golden/make_golden_batch.py`). Its RNG is process-global, so
every call runs in a process of its own.
"""
import json
import multiprocessing as mp
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from kahip_amd import graphs  # noqa: E402

GRAPHS = 40


def run_one(i):
    x, a, vw, ew = graphs.batch_graph(1, i)
    cut, part = O.ref_kaffpa(x, a, vw, ew, 8, 0.03, 0, 3)
    return {"graph": i, "n": len(x) - 1, "m2": int(x[-1]), "k": 8, "seed": 0,
            "mode": "FASTSOCIAL", "cut": int(cut), "balance": O.ref_balance(x, a, vw, 8, part)}


def main():
    O.build(with_reference=True)
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1), maxtasksperchild=1) as pool:
        recs = pool.map(run_one, range(GRAPHS), chunksize=1)
    out = {"reference": "KaHIP v3.21 (README.md:1) compiled from source by oracle/Makefile",
           "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                      text=True).stdout.splitlines()[0],
           "batch_seed": 1, "kaffpa": recs}
    with open(os.path.join(HERE, "ref_batch_sample.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(recs), "records")


if __name__ == "__main__":
    main()
