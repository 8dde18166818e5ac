"""Pins of the rating oracle (oracle/rating.py) on the CPU: torch float64 autograd, finite
differences, torch's bfloat16 conversion, per-edge brute force and exact rationals."""
from fractions import Fraction
import math

import numpy as np
import pytest

import oracle as O
import oracle.rating as R
from conftest import example, random_graph


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3 * scale, R.PARAMS)
    return p


def torch_mlp(X, params):
    torch = pytest.importorskip("torch")
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    W1 = p[R.OFF_W1:R.OFF_B1].reshape(R.HIDDEN, R.FEATURES)
    b1 = p[R.OFF_B1:R.OFF_W2]
    W2 = p[R.OFF_W2:R.OFF_B2].reshape(R.HIDDEN, R.HIDDEN)
    b2 = p[R.OFF_B2:R.OFF_W3]
    w3 = p[R.OFF_W3:R.OFF_B3]
    b3 = p[R.OFF_B3]
    x = torch.tensor(X, dtype=torch.float64)
    h = torch.relu(torch.nn.functional.linear(x, W1, b1))
    h = torch.relu(torch.nn.functional.linear(h, W2, b2))
    return p, h @ w3 + b3


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000) * 10, rng.normal(size=5000) * 1e-3,
                        [0.0, -0.0, 1.0, 3.0e38, -3.0e38, 1e-40]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).double().numpy()
    assert np.array_equal(R.bf16_round(x), ref)


def test_forward_and_gradient_match_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    X = rng.normal(size=(300, 8)) * 2
    y = (rng.random(300) < 0.4).astype(np.float64)
    params = rand_params(3)
    p, z = torch_mlp(X, params)
    assert np.allclose(R.mlp_forward(X, params), z.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss_t = torch.nn.functional.binary_cross_entropy_with_logits(z, torch.tensor(y))
    loss_t.backward()
    loss, g = R.mlp_loss_grad(X, y, params)
    assert abs(loss - loss_t.item()) < 1e-12
    assert np.allclose(g, p.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    X = rng.normal(size=(64, 8))
    y = (rng.random(64) < 0.5).astype(np.float64)
    params = rand_params(5)
    _, g = R.mlp_loss_grad(X, y, params)
    idx = list(rng.choice(R.PARAMS, 40, replace=False)) + [R.OFF_B3, R.OFF_W1, R.OFF_W2 + 65,
                                                           R.OFF_W3 + 7, R.OFF_B1 + 3]
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.mlp_loss_grad(X, y, pp)[0] - R.mlp_loss_grad(X, y, pm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_forward_closed_form_selector_network():
    # W1 copies feature f to hidden unit f (others 0), W2 = identity, w3 = (1..8, 0...), biases 0
    # => logit = sum_f (f + 1) * relu(x_f): pins the row/column layout of W1, W2 and w3.
    p = np.zeros(R.PARAMS)
    W1 = p[R.OFF_W1:R.OFF_B1].reshape(64, 8)
    for f in range(8):
        W1[f, f] = 1.0
    p[R.OFF_W2:R.OFF_B2].reshape(64, 64)[:] = np.eye(64)
    p[R.OFF_W3:R.OFF_W3 + 8] = np.arange(1, 9)
    p[R.OFF_B3] = -0.5
    X = np.random.default_rng(6).normal(size=(50, 8))
    expect = (np.maximum(X, 0) * np.arange(1, 9)).sum(axis=1) - 0.5
    assert np.allclose(R.mlp_forward(X, p), expect)
    # BF16 roundings: inputs are rounded, everything else here is exact in BF16
    assert np.allclose(R.mlp_forward(X, p, bf16=True),
                       (np.maximum(R.bf16_round(X), 0) * np.arange(1, 9)).sum(axis=1) - 0.5)


def test_quantize_logit_closed_forms():
    q = R.quantize_logit(np.array([0.0, 1.0, -1.0, 100.0, -100.0, np.nan, 0.5 / 128,
                                   1.5 / 128, -15.99, 15.99]))
    assert q.tolist() == [2048, 2176, 1920, 4096, 1, 1, 2048, 2050, 1, 4095]
    assert R.quantize_logit(np.array([0.0]), shift=3).tolist() == [256]
    assert R.quantize_logit(np.array([-100.0]), shift=3).tolist() == [1]


def brute_features(xadj, adjncy, vwgt, adjwgt, labels):
    n = len(xadj) - 1
    out = []
    sig = []
    for u in range(n):
        s = 0
        for e in range(xadj[u], xadj[u + 1]):
            y = int(adjncy[e])
            h = int(R.mix32(np.array([y], np.uint32))[0])
            s |= 1 << (h & 63)
        sig.append(s)
    for u in range(n):
        for e in range(xadj[u], xadj[u + 1]):
            v = int(adjncy[e])
            du, dv = xadj[u + 1] - xadj[u], xadj[v + 1] - xadj[v]
            w = 1 if adjwgt is None else int(adjwgt[e])
            cu = 1 if vwgt is None else int(vwgt[u])
            cv = 1 if vwgt is None else int(vwgt[v])
            agree = sum(int(l[u] == l[v]) for l in labels)
            out.append([math.log2(1 + min(du, dv)), math.log2(1 + max(du, dv)), math.log2(1 + w),
                        math.log2(1 + min(cu, cv)), math.log2(1 + max(cu, cv)),
                        bin(sig[u] & sig[v]).count("1") / 8, agree,
                        math.log2(1 + w * w / max(cu * cv, 1))])
    return np.array(out).reshape(-1, 8)


@pytest.mark.parametrize("which", ["weighted", "random"])
def test_features_brute_force(which):
    if which == "weighted":
        x, a, vw, ew = example("example_weighted")
    else:
        x, a, vw, ew = random_graph(300, 0.02, seed=9, weighted=True)
    labels = R.sweep_labels(x, a, vw, ew, 40, 77)
    f = R.edge_features(x, a, vw, ew, 40, 77)
    assert np.allclose(f, brute_features(x, a, vw, ew, labels), rtol=0, atol=1e-12)
    # the sweeps are one synchronous LP round each: a label either stays or is a neighbour's
    for lab in labels:
        assert lab.shape == (len(x) - 1,)


def test_mix32_matches_oracle_hash():
    # hash3(s, a, b) = mix32(mix32(mix32(s ^ 0x85ebca6b) ^ a) ^ b)
    for s, a, b in [(0, 0, 0), (1, 2, 3), (0xFFFFFFFF, 7, 0x12345678)]:
        h = R.mix32(np.array([s ^ 0x85EBCA6B], np.uint32))[0]
        h = R.mix32(np.array([int(h) ^ a], np.uint32))[0]
        h = R.mix32(np.array([int(h) ^ b], np.uint32))[0]
        assert int(h) == O.hash3(s, a, b)


def test_expansion_star2_exact_rationals():
    x, a, vw, ew = random_graph(200, 0.03, seed=3, weighted=True)
    vw = np.asarray(vw).copy()
    vw[::7] = 0  # c(u) c(v) = 0 takes the maximum score
    s = R.expansion_star2_scores(x, a, vw, ew)
    k = 0
    for u in range(len(x) - 1):
        for e in range(x[u], x[u + 1]):
            c = int(vw[u]) * int(vw[a[e]])
            if c == 0:
                q = 4096
            else:
                fr = Fraction(1024 * int(ew[e]) ** 2, c)
                q = math.floor(fr + Fraction(1, 2))
            assert s[k] == min(max(q, 1), 4096)
            k += 1


def test_score_shift_bounds():
    assert R.score_shift(0) == 0
    assert R.score_shift(524287) == 0
    assert R.score_shift(524288) == 1
    for d in [1, 1000, 524288, 10 ** 6, 2 ** 30]:
        s = R.score_shift(d)
        assert d * (4096 >> s) < 2 ** 31
        assert s == 0 or d * (4096 >> (s - 1)) >= 2 ** 31


def test_bf16_backward_rounds_only_the_product_operands():
    """bf16_backward changes the gradients that come from the rounded D2 / D1 (dW2, dW1 and,
    through D1 = D2 W2, db1) by the BF16 operand error (relative 2^-9 per operand) and leaves
    db2, dw3, db3 and the loss unchanged."""
    rng = np.random.default_rng(3)
    X = np.abs(rng.normal(size=(200, 8))) * 2
    y = (rng.random(200) < 0.4).astype(np.float64)
    p = rng.normal(size=R.PARAMS) * 0.3
    l0, g0 = R.mlp_loss_grad(X, y, p, bf16=True)
    l1, g1 = R.mlp_loss_grad(X, y, p, bf16=True, bf16_backward=True)
    assert l0 == l1
    for lo, hi in ((R.OFF_B2, R.OFF_W3), (R.OFF_W3, R.PARAMS)):
        assert np.array_equal(g0[lo:hi], g1[lo:hi])
    for lo, hi in ((R.OFF_W1, R.OFF_B1), (R.OFF_B1, R.OFF_W2), (R.OFF_W2, R.OFF_B2)):
        d = np.linalg.norm(g1[lo:hi] - g0[lo:hi]) / np.linalg.norm(g0[lo:hi])
        assert 0 < d < 2.0 ** -8
