#!/usr/bin/env python3
"""Writes the reference's kaffpa(FASTSOCIAL) cuts on generated graphs larger than the example
fixtures into tests/golden/ref_kaffpa_large.json.

The reference is KaHIP v3.21 compiled from its own sources by oracle/Makefile into
oracle/_ref/libkahip_ref.so (run in a container where /root/reference exists:
`python tests/golden/make_golden_large.py`). The graphs are regenerated from their seeds by
kahip_amd.graphs (gen_rgg / gen_rmat), so only the cuts are stored, together with a SHA-256 of
each graph's CSR arrays that the tests check before they compare cuts.

Cases (eps = 3 %, seeds 0-9): RGG n = 2^20 (graph seed 7), k = 64; R-MAT scales 16, 17, 18 with
16 samples per node (graph seed 7), k = 256. The reference keeps its RNG in process-global state,
so every kaffpa call runs in a process of its own.
"""
import hashlib
import json
import multiprocessing as mp
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from kahip_amd import graphs  # noqa: E402

GRAPH_SEED = 7
CASES = [("rgg", 20, 64), ("rmat", 16, 256), ("rmat", 17, 256), ("rmat", 18, 256)]
SEEDS = range(10)


def make_graph(kind, size):
    if kind == "rgg":
        return graphs.gen_rgg(1 << size, GRAPH_SEED)
    return graphs.gen_rmat(size, 16 << size, GRAPH_SEED)


def graph_sha(x, a, ew):
    h = hashlib.sha256()
    for arr in (x, a) + ((ew,) if ew is not None else ()):
        h.update(np.ascontiguousarray(arr, np.int32).tobytes())
    return h.hexdigest()


def run_one(job):
    kind, size, k, seed = job
    x, a, vw, ew = make_graph(kind, size)
    t = time.time()
    cut, part = O.ref_kaffpa(x, a, vw, ew, k, 0.03, seed, 3)
    dt = time.time() - t
    return {"kind": kind, "size": size, "k": k, "seed": seed, "mode": "FASTSOCIAL", "cut": int(cut),
            "balance": O.ref_balance(x, a, vw, k, part), "seconds": round(dt, 2)}


def main():
    O.build(with_reference=True)
    meta = []
    for kind, size, k in CASES:
        x, a, vw, ew = make_graph(kind, size)
        meta.append({"kind": kind, "size": size, "graph_seed": GRAPH_SEED, "n": len(x) - 1,
                     "m2": int(x[-1]), "sha256": graph_sha(x, a, ew)})
    jobs = [(kind, size, k, s) for kind, size, k in CASES for s in SEEDS]
    ctx = mp.get_context("spawn")
    with ctx.Pool(min(8, os.cpu_count() or 1), maxtasksperchild=1) as pool:
        recs = pool.map(run_one, jobs, chunksize=1)
    out = {"reference": "KaHIP v3.21 (README.md:1) compiled from source by oracle/Makefile",
           "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                      text=True).stdout.splitlines()[0],
           "imbalance": 0.03, "graphs": meta, "kaffpa": recs}
    with open(os.path.join(HERE, "ref_kaffpa_large.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(recs), "kaffpa runs")


if __name__ == "__main__":
    main()
