"""Graph and partition file formats and small synthetic generators (host side, numpy).

Formats follow the reference's readers/writers:
  * METIS text graph, formats 0/1/10/11 (lib/io/graph_io.cpp:82-191): header ``n m [fmt]``, one
    line per node, 1-based neighbour ids, optional node weight first (fmt 10/11) and an edge weight
    after each neighbour (fmt 1/11); lines starting with '%' are comments.
  * partition file: one block id per line (lib/io/graph_io.cpp:50-80, 194-203).
  * BGF binary graph (parallel/parallel_src/lib/io/parallel_graph_io.cpp:474-523): little-endian
    uint64 words: version 3, n, number of directed entries; n + 1 byte offsets of the rows
    (the first is the header size, (3 + n + 1) * 8); the neighbour ids.

Graphs are returned as ``(xadj, adjncy, vwgt, adjwgt)`` int32 numpy arrays (0-based CSR, both
directions stored); vwgt / adjwgt are None for unit weights.

The synthetic generators here are exact, seeded and counter-based so that the same graph can be
regenerated anywhere (CPU here, the C++ host generator and the GPU generator in the library):
  * ``rng_u64(seed, stream, i)``: SplitMix64 finaliser of ``seed*K1 + stream*K2 + i``.
  * RGG: points uniform in [0,1)^2, edge iff squared distance < r^2 with
    r = 0.55*sqrt(ln n / n); nodes renumbered in grid-cell order (cells of side >= r, row-major,
    ties by generation index), rows sorted ascending.
  * R-MAT: ``scale`` levels, integer quadrant thresholds floor(p * 2^32); self-loops dropped,
    symmetrised, duplicates merged with weight = multiplicity, rows sorted.
"""
from __future__ import annotations

import gzip
import io
import math
import os

import numpy as np

M64 = (1 << 64) - 1
K1 = 0xD1B54A32D192ED03
K2 = 0x8CB92BA72F3D8DD7


# ----------------------------------------------------------------------------- BGF binary
BGF_VERSION = 3


def write_bgf(path, xadj, adjncy):
    """Write the unweighted structure of a graph as BGF."""
    xadj = np.asarray(xadj, np.uint64)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    head = np.array([BGF_VERSION, n, m2], np.uint64)
    offs = np.uint64((3 + n + 1) * 8) + xadj * np.uint64(8)
    with open(path, "wb") as fh:
        head.astype("<u8").tofile(fh)
        offs.astype("<u8").tofile(fh)
        np.asarray(adjncy, np.uint64).astype("<u8").tofile(fh)


def read_bgf(path):
    """Read a BGF graph; returns (xadj, adjncy, None, None) as int32."""
    head = np.fromfile(path, dtype="<u8", count=3)
    if len(head) != 3 or int(head[0]) != BGF_VERSION:
        raise ValueError(f"{path}: not a BGF file of version {BGF_VERSION}")
    n, m2 = int(head[1]), int(head[2])
    words = np.fromfile(path, dtype="<u8", count=3 + n + 1 + m2)
    if len(words) != 3 + n + 1 + m2:
        raise ValueError(f"{path}: truncated")
    offs = words[3:3 + n + 1].astype(np.int64)
    if offs[0] != (3 + n + 1) * 8 or np.any(np.diff(offs) < 0) or offs[-1] != offs[0] + 8 * m2:
        raise ValueError(f"{path}: inconsistent offsets")
    xadj = ((offs - offs[0]) // 8).astype(np.int32)
    adjncy = words[3 + n + 1:].astype(np.int64)
    if m2 and (adjncy.max() >= n):
        raise ValueError(f"{path}: neighbour id out of range")
    return xadj, adjncy.astype(np.int32), None, None


# ----------------------------------------------------------------------------- METIS IO
def _open_text(path):
    if hasattr(path, "read"):
        return path
    if str(path).endswith(".gz"):
        return io.TextIOWrapper(gzip.open(path, "rb"), encoding="ascii")
    return open(path, "r", encoding="ascii")


def read_metis(path):
    """Read a METIS text graph (fmt 0/1/10/11)."""
    f = _open_text(path)
    try:
        lines = [ln for ln in f.read().splitlines() if not ln.lstrip().startswith("%")]
    finally:
        if not hasattr(path, "read"):
            f.close()
    head = lines[0].split()
    n, m = int(head[0]), int(head[1])
    fmt = int(head[2]) if len(head) > 2 else 0
    has_vw = fmt in (10, 11)
    has_ew = fmt in (1, 11)
    xadj = np.zeros(n + 1, np.int64)
    adj, ew, vw = [], [], []
    for i in range(n):
        tok = lines[1 + i].split() if 1 + i < len(lines) else []
        vals = list(map(int, tok))
        p = 0
        if has_vw:
            vw.append(vals[0])
            p = 1
        rest = vals[p:]
        if has_ew:
            adj.extend(v - 1 for v in rest[0::2])
            ew.extend(rest[1::2])
            xadj[i + 1] = xadj[i] + len(rest) // 2
        else:
            adj.extend(v - 1 for v in rest)
            xadj[i + 1] = xadj[i] + len(rest)
    if xadj[n] != 2 * m:
        raise ValueError(f"edge count mismatch: header says {m} undirected, found {xadj[n]} entries")
    return (xadj.astype(np.int32), np.asarray(adj, np.int32),
            np.asarray(vw, np.int32) if has_vw else None,
            np.asarray(ew, np.int32) if has_ew else None)


def write_metis(path, xadj, adjncy, vwgt=None, adjwgt=None):
    n = len(xadj) - 1
    fmt = (10 if vwgt is not None else 0) + (1 if adjwgt is not None else 0)
    with open(path, "w", encoding="ascii") as f:
        f.write(f"{n} {int(xadj[n]) // 2}" + (f" {fmt}" if fmt else "") + "\n")
        for v in range(n):
            parts = []
            if vwgt is not None:
                parts.append(str(int(vwgt[v])))
            for e in range(int(xadj[v]), int(xadj[v + 1])):
                parts.append(str(int(adjncy[e]) + 1))
                if adjwgt is not None:
                    parts.append(str(int(adjwgt[e])))
            f.write(" ".join(parts) + "\n")


def read_partition(path, n=None):
    with _open_text(path) as f:
        p = np.array([int(x) for x in f.read().split()], np.int32)
    if n is not None and len(p) != n:
        raise ValueError(f"partition has {len(p)} entries, graph has {n} nodes")
    return p


def write_partition(path, part):
    with open(path, "w", encoding="ascii") as f:
        f.write("\n".join(str(int(b)) for b in part) + "\n")


def check_graph(xadj, adjncy, vwgt=None, adjwgt=None):
    """Structural checks in the spirit of app/graphchecker.cpp:91-202. Raises ValueError."""
    n = len(xadj) - 1
    if n < 0 or xadj[0] != 0 or np.any(np.diff(xadj.astype(np.int64)) < 0):
        raise ValueError("xadj must start at 0 and be non-decreasing")
    if len(adjncy) != xadj[n]:
        raise ValueError("len(adjncy) != xadj[n]")
    if len(adjncy) and (adjncy.min() < 0 or adjncy.max() >= n):
        raise ValueError("neighbour id out of range")
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(xadj.astype(np.int64)))
    if np.any(src == adjncy):
        raise ValueError("self-loop")
    w = np.ones(len(adjncy), np.int64) if adjwgt is None else adjwgt.astype(np.int64)
    if adjwgt is not None and np.any(w <= 0):
        raise ValueError("edge weights must be positive")
    fwd = src * n + adjncy
    bwd = adjncy.astype(np.int64) * n + src
    of, ob = np.argsort(fwd, kind="stable"), np.argsort(bwd, kind="stable")
    if not (np.array_equal(fwd[of], bwd[ob]) and np.array_equal(w[of], w[ob])):
        raise ValueError("graph is not symmetric (or parallel edges / weights disagree)")
    if vwgt is not None and np.any(vwgt < 0):
        raise ValueError("negative node weight")


# ----------------------------------------------------------------------------- generators
def _sm64(z):
    """SplitMix64 finaliser on a uint64 numpy array (wrapping arithmetic)."""
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rng_u64(seed, stream, idx):
    """Counter-based 64-bit draws: SplitMix64(seed*K1 + stream*K2 + idx)."""
    base = np.uint64(((seed * K1) + (stream * K2)) & M64)
    with np.errstate(over="ignore"):
        return _sm64(base + np.asarray(idx, dtype=np.uint64))


def rgg_radius(n):
    return 0.55 * math.sqrt(math.log(n) / n)


def _csr_from_pairs(n, u, v, w=None):
    """CSR with rows sorted by neighbour id from directed pairs (already symmetric)."""
    key = u.astype(np.int64) * n + v.astype(np.int64)
    order = np.argsort(key, kind="stable")
    u, v = u[order], v[order]
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, u.astype(np.int64) + 1, 1)
    xadj = np.cumsum(xadj)
    return xadj.astype(np.int32), v.astype(np.int32), (None if w is None else w[order].astype(np.int32))


def gen_rgg(n, seed, radius=None):
    """Random geometric graph (see module docstring). Returns (xadj, adjncy, None, None)."""
    r = rgg_radius(n) if radius is None else radius
    idx = np.arange(2 * n, dtype=np.uint64)
    d = rng_u64(seed, 1, idx)
    xy = (d >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    x, y = xy[0::2].copy(), xy[1::2].copy()
    g = max(1, int(math.floor(1.0 / r)))  # cells of side 1/g >= r
    cx = np.minimum((x * g).astype(np.int64), g - 1)
    cy = np.minimum((y * g).astype(np.int64), g - 1)
    cell = cy * g + cx
    order = np.lexsort((np.arange(n), cell))  # by cell, then generation index
    newid = np.empty(n, np.int64)
    newid[order] = np.arange(n)
    x, y, cell = x[order], y[order], cell[order]
    cx, cy = cx[order], cy[order]
    start = np.searchsorted(cell, np.arange(g * g + 1))
    r2 = r * r
    us, vs = [], []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            ncx, ncy = cx + dx, cy + dy
            ok = (ncx >= 0) & (ncx < g) & (ncy >= 0) & (ncy < g)
            src = np.nonzero(ok)[0]
            nc = ncy[src] * g + ncx[src]
            lo, hi = start[nc], start[nc + 1]
            cnt = hi - lo
            rep_src = np.repeat(src, cnt)
            offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            tgt = np.repeat(lo, cnt) + offs
            ddx = x[rep_src] - x[tgt]
            ddy = y[rep_src] - y[tgt]
            keep = (ddx * ddx + ddy * ddy < r2) & (rep_src != tgt)
            us.append(rep_src[keep])
            vs.append(tgt[keep])
    u = np.concatenate(us)
    v = np.concatenate(vs)
    xadj, adjncy, _ = _csr_from_pairs(n, u, v)
    return xadj, adjncy, None, None


def batch_spec(seed, i, log2_min=10.0, log2_max=16.0):
    """Graph i of a generated batch: (kind, size, graph_seed) with kind 'rgg' (n = size) for
    even i and 'rmat' (scale = size, 16 samples per node) for odd i;
    x = log2_min + u (log2_max - log2_min), u = (d >> 11) 2^-53, d = rng_u64(seed, 3, i);
    RGG n = floor(2^x + 0.5), R-MAT scale = floor(x + 0.5); graph seed = rng_u64(seed, 4, i)."""
    d = int(rng_u64(seed, 3, i))
    u = float(d >> 11) * (1.0 / 9007199254740992.0)
    x = log2_min + u * (log2_max - log2_min)
    gseed = int(rng_u64(seed, 4, i))
    if i % 2 == 0:
        return "rgg", int(math.floor(2.0 ** x + 0.5)), gseed
    return "rmat", int(math.floor(x + 0.5)), gseed


def batch_graph(seed, i, log2_min=10.0, log2_max=16.0):
    kind, size, gseed = batch_spec(seed, i, log2_min, log2_max)
    if kind == "rgg":
        return gen_rgg(size, gseed)
    return gen_rmat(size, 16 << size, gseed)


RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    t1 = int(math.floor(a * 4294967296.0))
    t2 = int(math.floor((a + b) * 4294967296.0))
    t3 = int(math.floor((a + b + c) * 4294967296.0))
    return t1, t2, t3


def rmat_samples(scale, nsamples, seed, a=0.57, b=0.19, c=0.19):
    """The raw R-MAT (u, v) samples: sample j uses draws rng_u64(seed, 2, j*ceil(scale/2)+t),
    each 64-bit draw giving two 32-bit uniforms (low half first)."""
    t1, t2, t3 = rmat_thresholds(a, b, c)
    per = (scale + 1) // 2
    u = np.zeros(nsamples, np.int64)
    v = np.zeros(nsamples, np.int64)
    j = np.arange(nsamples, dtype=np.uint64)
    for lvl in range(scale):
        d = rng_u64(seed, 2, j * np.uint64(per) + np.uint64(lvl // 2))
        r = (d & np.uint64(0xFFFFFFFF)) if lvl % 2 == 0 else (d >> np.uint64(32))
        r = r.astype(np.int64)
        q = (r >= t1).astype(np.int64) + (r >= t2) + (r >= t3)
        u = u * 2 + (q >> 1)
        v = v * 2 + (q & 1)
    return u, v


def gen_rmat(scale, nsamples, seed, a=0.57, b=0.19, c=0.19):
    """R-MAT graph (see module docstring). Returns (xadj, adjncy, None, adjwgt)."""
    n = 1 << scale
    u, v = rmat_samples(scale, nsamples, seed, a, b, c)
    keep = u != v
    u, v = u[keep], v[keep]
    uu = np.concatenate([u, v])
    vv = np.concatenate([v, u])
    key = uu * n + vv
    uk, cnt = np.unique(key, return_counts=True)
    su = (uk // n).astype(np.int64)
    sv = (uk % n).astype(np.int64)
    xadj, adjncy, w = _csr_from_pairs(n, su, sv, cnt.astype(np.int64))
    return xadj, adjncy, None, w
