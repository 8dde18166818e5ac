#!/usr/bin/env python3
"""Writes the golden vectors of the reference (KaHIP v3.21 compiled from its own sources by
oracle/Makefile into oracle/_ref/libkahip_ref.so) into tests/golden/ref_vectors.json and
tests/golden/ref_arrays.npz.

Run in a container where /root/reference exists: `python tests/golden/make_golden.py`.
Inputs are the reference's own example graphs (examples/*.graph, kept here gzipped) and
partitions derived from them; outputs are recorded as SHA-256 of the int32 little-endian array
bytes plus summary numbers, and a few full arrays.
"""
import hashlib
import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from kahip_amd import graphs  # noqa: E402

GRAPHS = ["rgg_n_2_15_s0", "delaunay_n15", "example_weighted"]


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


def load(name):
    return graphs.read_metis(os.path.join(HERE, name + ".graph.gz"))


def main():
    O.build(with_reference=True)
    out = {"reference": "KaHIP v3.21 (README.md:1) compiled from source by oracle/Makefile",
           "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                      text=True).stdout.splitlines()[0],
           "platform": platform.platform(),
           "lp_cluster": [], "contract": [], "lp_refine": [], "metrics": [], "upper_bound": [],
           "kaffpa": []}
    arrays = {}
    for name in GRAPHS:
        x, a, vw, ew = load(name)
        n = len(x) - 1
        for bound in (938, 3751, 16):
            for seed in range(10):
                cl, nc = O.ref_lp_cluster(x, a, vw, ew, bound, 10, seed)
                out["lp_cluster"].append({"graph": name, "bound": bound, "iterations": 10,
                                          "seed": seed, "clusters": nc, "sha256": h(cl),
                                          "cut": O.ref_edge_cut(x, a, ew, cl)})
                if name == "rgg_n_2_15_s0" and bound == 938 and seed < 2:
                    arrays[f"cluster_{name}_{bound}_{seed}"] = cl
                if seed < 3:
                    cx, ca, cw, cv = O.ref_contract(x, a, vw, ew, cl, nc)
                    out["contract"].append({"graph": name, "bound": bound, "seed": seed,
                                            "nc": nc, "mc": int(cx[-1]), "xadj": h(cx),
                                            "adjncy": h(ca), "adjwgt": h(cw), "vwgt": h(cv)})
                    if name == "example_weighted" and bound == 16 and seed == 0:
                        arrays["contract_weighted_cluster"] = cl
                        arrays["contract_weighted_xadj"] = cx
                        arrays["contract_weighted_adjncy"] = ca
                        arrays["contract_weighted_adjwgt"] = cw
                        arrays["contract_weighted_vwgt"] = cv
        W = n if vw is None else int(vw.sum())
        for k in (2, 8, 64):
            U = O.ref_upper_bound(n, vw, k, 0.03)
            for seed in range(3):
                start = (np.arange(n) * k // n).astype(np.int32)
                p = O.ref_lp_refine(x, a, vw, ew, k, U, 25, seed, start)
                out["lp_refine"].append({"graph": name, "k": k, "bound": U, "rounds": 25,
                                         "seed": seed, "start": "blocks", "sha256": h(p),
                                         "cut": O.ref_edge_cut(x, a, ew, p)})
            rng = np.random.default_rng(k)
            part = rng.integers(0, k, n).astype(np.int32)
            out["metrics"].append({"graph": name, "k": k, "partition": "rng(k).integers(0,k,n)",
                                   "cut": O.ref_edge_cut(x, a, ew, part),
                                   "balance": O.ref_balance(x, a, vw, k, part)})
            out["upper_bound"].append({"n": n, "W": W, "k": k, "imbalance": 0.03, "U": U})
    for n, k in ((32768, 2), (4194304, 64), (16777216, 256), (1024, 8), (65536, 8)):
        out["upper_bound"].append({"n": n, "W": n, "k": k, "imbalance": 0.03,
                                   "U": O.ref_upper_bound(n, None, k, 0.03)})
    for name, ks in (("rgg_n_2_15_s0", (2, 8, 64)), ("delaunay_n15", (2, 8))):
        x, a, vw, ew = load(name)
        for k in ks:
            for seed in range(10):
                for mode, mname in ((3, "FASTSOCIAL"), (0, "FAST"), (1, "ECO")):
                    if mname != "FASTSOCIAL" and seed >= 3:
                        continue
                    cut, part = O.ref_kaffpa(x, a, vw, ew, k, 0.03, seed, mode)
                    out["kaffpa"].append({"graph": name, "k": k, "seed": seed, "mode": mname,
                                          "cut": cut, "balance": O.ref_balance(x, a, vw, k, part),
                                          "sha256": h(part)})
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_arrays.npz"), **arrays)
    print("wrote", len(out["lp_cluster"]), "clusterings,", len(out["contract"]), "contractions,",
          len(out["lp_refine"]), "refinements,", len(out["kaffpa"]), "kaffpa runs")


if __name__ == "__main__":
    main()
