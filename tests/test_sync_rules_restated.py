"""An independent restatement of the synchronous LP schedule's rules, written in plain Python from
their prose description (not from the oracle's C++), checked bit for bit against the oracle's
synchronous clustering and refinement on many tiny graphs (n <= 12). A wrong tie direction, a
sub-round assignment, an admission or departure rule that the oracle and the HIP kernels shared
would show here.

The rules (one round = S sub-rounds; node v belongs to sub-round hash3(seed ^ 0x68bc21eb, 0, v)
mod S; nodes of degree 0 never propose):
- Within a sub-round every member proposes against the labels and label weights at its start.
  Candidates are the labels c != own of its neighbours whose weight satisfies
  size(c) + w(v) < bound(c); conn(c) sums the weights of the edges to neighbours labelled c.
  The best candidate maximises (conn, hash3(tie seed of the round, v, c)) and then minimises c;
  the round's tie seed is seed ^ 0x2f6b7d93 ^ (round * 0x9e3779b9) mod 2^32.
- Clustering proposes iff conn(c) > conn(own), or they tie and size(c) >= size(own).
  Refinement proposes iff conn(c) > conn(own), or own is overloaded (size(own) > bound(own)),
  or they tie and size(c) + 2 w(v) <= size(own).
- Priority of a proposal: higher gain conn(c) - conn(own) first, then lower v.
- Refinement: a proposal out of an overloaded block is a departure. Departures from block b all
  pass if size(b) - (their summed weight) >= bound(b); otherwise only b's top-priority departure
  passes. Passing departures then count as proposals into their targets.
- Admission into t: all proposals into t are admitted if size(t) + (their summed weight) <
  bound(t); otherwise only t's top-priority proposal, and only if size(t) + w(v) < bound(t).
- Admitted moves are applied together (labels and label weights), then the next sub-round runs.
- Clustering runs every round over all nodes. Refinement rounds run over a frontier: all nodes
  in round 0, afterwards the neighbours of the nodes that moved in the previous round.
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_graph

M32 = 0xFFFFFFFF


def mix32(x):
    """Wellons' lowbias32 integer hash."""
    x &= M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def hash3(s, a, b):
    return mix32(mix32(mix32((s ^ 0x85EBCA6B) & M32) ^ (a & M32)) ^ (b & M32))


def subround(seed, v, S):
    return hash3(seed ^ 0x68BC21EB, 0, v) % S


def tie_seed(seed, r):
    return (seed ^ 0x2F6B7D93 ^ ((r * 0x9E3779B9) & M32)) & M32


def run_rules(x, a, vw, ew, label, size, bound, rounds, S, seed, refine):
    """label / size are lists (modified in place); bound is a function of the label.
    Returns the moves of every round."""
    n = len(x) - 1
    vw = [1] * n if vw is None else [int(t) for t in vw]
    ew = [1] * int(x[-1]) if ew is None else [int(t) for t in ew]
    nbrs = [[(int(a[e]), ew[e]) for e in range(x[v], x[v + 1])] for v in range(n)]
    active = [True] * n
    moves_per_round = []
    for r in range(rounds):
        ts = tie_seed(seed, r)
        nxt = [False] * n
        moved = 0
        for sr in range(S):
            members = [v for v in range(n) if nbrs[v] and subround(seed, v, S) == sr
                       and (not refine or active[v])]
            props = {}  # v -> (target, gain, departure)
            for v in members:
                own = label[v]
                conn = {}
                for u, w in nbrs[v]:
                    conn[label[u]] = conn.get(label[u], 0) + w
                conn_own = conn.get(own, 0)
                best = None
                for c, cc in conn.items():
                    if c == own or size[c] + vw[v] >= bound(c):
                        continue
                    key = (cc, hash3(ts, v, c), -c)
                    if best is None or key > best[0]:
                        best = (key, c)
                if best is None:
                    continue
                c = best[1]
                cc = conn[c]
                if refine:
                    over = size[own] > bound(own)
                    go = cc > conn_own or over or (cc == conn_own and size[c] + 2 * vw[v] <= size[own])
                else:
                    over = False
                    go = cc > conn_own or (cc == conn_own and size[c] >= size[own])
                if go:
                    props[v] = (c, cc - conn_own, over)

            def prio(v):
                return (props[v][1], -v)

            # departures from overloaded blocks (refinement)
            passing = {}
            dep_by_block = {}
            for v, (c, g, dep) in props.items():
                if dep:
                    dep_by_block.setdefault(label[v], []).append(v)
                else:
                    passing[v] = c
            for b, vs in dep_by_block.items():
                if size[b] - sum(vw[v] for v in vs) >= bound(b):
                    for v in vs:
                        passing[v] = props[v][0]
                else:
                    top = max(vs, key=prio)
                    passing[top] = props[top][0]
            by_target = {}
            for v, t in passing.items():
                by_target.setdefault(t, []).append(v)
            admitted = []
            for t, vs in by_target.items():
                if size[t] + sum(vw[v] for v in vs) < bound(t):
                    admitted += vs
                else:
                    top = max(vs, key=prio)
                    if size[t] + vw[top] < bound(t):
                        admitted.append(top)
            for v in admitted:
                t = passing[v]
                size[label[v]] -= vw[v]
                size[t] += vw[v]
                label[v] = t
                for u, _ in nbrs[v]:
                    nxt[u] = True
            moved += len(admitted)
        moves_per_round.append(moved)
        if refine:
            active = nxt
    return moves_per_round


def first_appearance(label):
    ids, out = {}, []
    for c in label:
        if c not in ids:
            ids[c] = len(ids)
        out.append(ids[c])
    return np.array(out, np.int32), len(ids)


def tiny_graphs():
    out = []
    for i in range(40):
        rng = np.random.default_rng(1000 + i)
        n = int(rng.integers(2, 13))
        p = float(rng.uniform(0.15, 0.8))
        out.append((i, random_graph(n, p, 2000 + i, weighted=bool(i % 2))))
    return out


@pytest.mark.parametrize("i,g", tiny_graphs())
def test_clustering_rules_match_oracle(i, g):
    x, a, vw, ew = g
    n = len(x) - 1
    W = n if vw is None else int(np.sum(vw))
    for bound in (2, 3, max(3, W // 3), W + 1):
        for S in (1, 2, 4):
            seed = (i * 7919 + bound * 31 + S) & M32
            label = list(range(n))
            size = [1] * n if vw is None else [int(t) for t in vw]
            moves = run_rules(x, a, vw, ew, label, size, lambda c: bound, 10, S, seed, False)
            want, wnc, wmoves = O.lp_cluster_sync(x, a, vw, ew, bound, 10, S, seed)
            got, nc = first_appearance(label)
            assert nc == wnc, (bound, S)
            assert np.array_equal(got, want), (bound, S)
            assert list(wmoves) == moves, (bound, S)


@pytest.mark.parametrize("i,g", tiny_graphs())
def test_refinement_rules_match_oracle(i, g):
    x, a, vw, ew = g
    n = len(x) - 1
    w = np.ones(n, np.int64) if vw is None else np.asarray(vw, np.int64)
    rng = np.random.default_rng(500 + i)
    for k in (2, 3):
        W = int(w.sum())
        for U in (max(1, W // k), W // k + 2, W):
            for S in (1, 4):
                start = rng.integers(0, k, n).astype(np.int32)
                seed = (i * 104729 + U * 13 + S * 7 + k) & M32
                label = [int(t) for t in start]
                size = [0] * k
                for v in range(n):
                    size[label[v]] += int(w[v])
                moves = run_rules(x, a, vw, ew, label, size, lambda b: U, 25, S, seed, True)
                want, wmoves = O.lp_refine_sync(x, a, vw, ew, k, U, 25, S, seed, start)
                assert np.array_equal(np.array(label, np.int32), want), (k, U, S)
                assert list(wmoves) == moves, (k, U, S)


def test_hash_restatement_matches_oracle():
    for s, a_, b in ((0, 0, 0), (1, 2, 3), (0xDEADBEEF, 12345, 678), (M32, M32, M32)):
        assert hash3(s, a_, b) == O.hash3(s, a_, b)
