"""CPU restatement of the edge ratings (test infrastructure only: imported by tests/, never by the
library).

* expansion_star2_scores: the plain rating w(e)^2 / (c(u) c(v)) of
  lib/partition/coarsening/edge_rating/edge_ratings.cpp:171-183, quantised to integers as
  clamp(round(1024 w^2 / (c(u) c(v))), 1, 4096) with round-half-up in exact integer arithmetic.
* edge_features / mlp_forward / mlp_loss_grad / quantize_logit: the learned rating, which has no
  reference counterpart. The MLP is evaluated in float64, optionally with the BF16 roundings the
  GPU applies (inputs, weights, layer-1 activations) so that the remaining difference is the
  GPU's FP32 accumulation only.

Parity pins (tests/test_oracle_rating.py): forward and gradient against torch float64 autograd
and against finite differences of the float64 loss; the BF16 rounding against torch's bfloat16
conversion; features against a per-edge brute force on small graphs; scores by exact integers.
"""
from __future__ import annotations

import numpy as np

from . import hash3, lp_cluster_sync

FEATURES = 8
HIDDEN = 64
SCORE_MAX = 4096
PARAMS = HIDDEN * FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
OFF_W1 = 0
OFF_B1 = OFF_W1 + HIDDEN * FEATURES
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN


def mix32(x):
    """Wellons' lowbias32 on uint32 arrays."""
    x = np.asarray(x, np.uint32).astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def _sources(xadj):
    xadj = np.asarray(xadj, np.int64)
    return np.repeat(np.arange(len(xadj) - 1), np.diff(xadj))


def sketch(xadj, adjncy):
    """sig[u] = OR over neighbours y of 1 << (mix32(y) & 63)."""
    n = len(xadj) - 1
    bits = np.left_shift(np.uint64(1), (mix32(np.asarray(adjncy, np.uint32)) & 63).astype(np.uint64))
    sig = np.zeros(n, np.uint64)
    np.bitwise_or.at(sig, _sources(xadj), bits)
    return sig


def _popcount64(x):
    return np.bitwise_count(np.asarray(x, np.uint64)).astype(np.int64)


def sweep_labels(xadj, adjncy, vwgt, adjwgt, bound, seed):
    """The three one-round synchronous LP clusterings behind feature 6 (one sub-round each)."""
    out = []
    for i in range(3):
        sd = hash3(seed, 0xA6, i)
        lab, _, _ = lp_cluster_sync(np.asarray(xadj, np.int32), np.asarray(adjncy, np.int32),
                                    vwgt, adjwgt, bound, 1, 1, sd)
        out.append(np.asarray(lab))
    return out


def edge_features(xadj, adjncy, vwgt, adjwgt, bound, seed, labels=None):
    """float64 [xadj[n], 8] features of every directed edge (see include/kahip_amd.h).
    `labels` (optional) replaces the three sweep clusterings; the GPU labels them by the raw
    label, this oracle by the relabelled cluster id, and agreement is invariant under relabel."""
    xadj = np.asarray(xadj, np.int64)
    adjncy = np.asarray(adjncy, np.int64)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    vw = np.ones(n, np.int64) if vwgt is None else np.asarray(vwgt, np.int64)
    ew = np.ones(m2, np.int64) if adjwgt is None else np.asarray(adjwgt, np.int64)
    u = _sources(xadj)
    v = adjncy
    deg = np.diff(xadj)
    du, dv = deg[u], deg[v]
    cu, cv = vw[u], vw[v]
    sig = sketch(xadj, adjncy)
    cn = _popcount64(sig[u] & sig[v])
    if labels is None:
        labels = sweep_labels(xadj, adjncy, vwgt, adjwgt, bound, seed)
    agree = sum((l[u] == l[v]).astype(np.int64) for l in labels)
    w = ew.astype(np.float64)
    es2 = w * w / np.maximum((cu * cv).astype(np.float64), 1.0)
    f = np.empty((m2, FEATURES), np.float64)
    f[:, 0] = np.log2(1.0 + np.minimum(du, dv))
    f[:, 1] = np.log2(1.0 + np.maximum(du, dv))
    f[:, 2] = np.log2(1.0 + w)
    f[:, 3] = np.log2(1.0 + np.minimum(cu, cv))
    f[:, 4] = np.log2(1.0 + np.maximum(cu, cv))
    f[:, 5] = cn / 8.0
    f[:, 6] = agree
    f[:, 7] = np.log2(1.0 + es2)
    return f


def score_shift(max_degree):
    s = 0
    while s < 12 and max_degree * (SCORE_MAX >> s) >= 2 ** 31:
        s += 1
    return s


def expansion_star2_scores(xadj, adjncy, vwgt, adjwgt):
    xadj = np.asarray(xadj, np.int64)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    vw = np.ones(n, np.int64) if vwgt is None else np.asarray(vwgt, np.int64)
    ew = np.ones(m2, np.int64) if adjwgt is None else np.asarray(adjwgt, np.int64)
    u = _sources(xadj)
    c = vw[u] * vw[np.asarray(adjncy, np.int64)]
    num = 1024 * ew * ew
    q = np.where(c > 0, (num + c // 2) // np.maximum(c, 1), SCORE_MAX)
    q = np.clip(q, 1, SCORE_MAX)
    sh = score_shift(int(np.diff(xadj).max()) if n else 0)
    return np.maximum(q >> sh, 1).astype(np.int32)


def quantize_logit(logit, shift=0):
    """clamp(rint(128 logit + 2048), 1, 4096) >> shift, at least 1; the arithmetic in float32
    as on the GPU."""
    x = np.asarray(logit, np.float32) * np.float32(128.0) + np.float32(2048.0)
    q = np.where(x >= 1.0, np.where(x >= SCORE_MAX, SCORE_MAX, np.rint(x)), 1)
    q = np.nan_to_num(q, nan=1).astype(np.int64)
    return np.maximum(q >> shift, 1).astype(np.int32)


def bf16_round(x):
    """Round float64/float32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(HIDDEN, FEATURES)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(HIDDEN, HIDDEN)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def _forward(features, params, bf16):
    X = np.asarray(features, np.float64).reshape(-1, FEATURES)
    W1, b1, W2, b2, w3, b3 = unpack(params)
    if bf16:
        X, W1, W2 = bf16_round(X), bf16_round(W1), bf16_round(W2)
    Z1 = X @ W1.T + b1
    H1 = np.maximum(Z1, 0.0)
    if bf16:
        H1 = bf16_round(H1)
    Z2 = H1 @ W2.T + b2
    H2 = np.maximum(Z2, 0.0)
    z = H2 @ w3 + b3
    return X, W1, W2, Z1, H1, Z2, H2, z


def mlp_forward(features, params, bf16=False):
    """Logits in float64; bf16=True applies the GPU's BF16 roundings."""
    return _forward(features, params, bf16)[-1]


def mlp_loss_grad(features, labels, params, bf16=False, bf16_backward=False):
    """Mean binary cross-entropy with logits and its gradient (float64). With bf16=True the
    forward uses the GPU's roundings and the backward treats them as the identity. With
    bf16_backward=True the backward products also take BF16 operands, as the GPU's MFMA training
    kernel does: D2 = dL/dZ2 is rounded before dW2 = D2^T H1 and dH1 = D2 W2, D1 = dL/dZ1 before
    dW1 = D1^T X; the bias and w3 gradients stay unrounded sums."""
    X, W1, W2, Z1, H1, Z2, H2, z = _forward(features, params, bf16)
    y = np.asarray(labels, np.float64)
    N = len(y)
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / N
    _, _, _, _, w3, _ = unpack(params)
    g = np.zeros(PARAMS, np.float64)
    g[OFF_W3:OFF_B3] = H2.T @ dz
    g[OFF_B3] = dz.sum()
    D2 = np.outer(dz, w3) * (Z2 > 0)
    D2p = bf16_round(D2) if bf16_backward else D2
    g[OFF_W2:OFF_B2] = (D2p.T @ H1).ravel()
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    D1 = (D2p @ W2) * (Z1 > 0)
    D1p = bf16_round(D1) if bf16_backward else D1
    g[OFF_W1:OFF_B1] = (D1p.T @ X).ravel()
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    return float(loss), g
