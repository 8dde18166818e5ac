"""GPU tests of the edge ratings: features, the MFMA MLP, the training kernels and the rated
clustering pipeline, against oracle/rating.py and the pipeline mirror.

Tolerances (float work only; every integer is compared exactly):
* features: the log2 columns within 2e-6 absolute of float64 (hardware log2 on float32 inputs),
  the common-neighbour and agreement columns exactly.
* MLP forward vs the float64 oracle with the same BF16 roundings: |dz| <= 1e-3 (1 + |z|) for
  every edge (the remaining difference is FP32 accumulation order, plus rare one-ulp flips of a
  BF16-rounded layer-1 activation), mean |dz| <= 1e-5 (1 + mean |z|).
* MLP forward vs plain float64 (no roundings): |dz| <= 0.05 (1 + |z|); the measured error is
  printed.
* gradient vs the float64 oracle with the same roundings: ||g - g_ref|| <= 1e-3 ||g_ref||.
"""
import numpy as np
import pytest

import kahip_amd as K
import oracle as O
import oracle.rating as R
from conftest import manual_graph, random_graph
from pipeline_mirror import mirror_partition

pytestmark = pytest.mark.gpu


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(R.PARAMS)
    p[R.OFF_W1:R.OFF_B1] = rng.normal(0, np.sqrt(2 / 8), 512) * scale
    p[R.OFF_B1:R.OFF_W2] = rng.normal(0, 0.1, 64)
    p[R.OFF_W2:R.OFF_B2] = rng.normal(0, np.sqrt(2 / 64), 4096) * scale
    p[R.OFF_B2:R.OFF_W3] = rng.normal(0, 0.1, 64)
    p[R.OFF_W3:R.OFF_B3] = rng.normal(0, np.sqrt(1 / 64), 64)
    p[R.OFF_B3] = 0.2
    return p.astype(np.float32)


def hub_graph(n=6000):
    """A ring plus two hubs of degree n - 1 and about n / 3: rows far longer than one thread's
    share of the neighbourhood sketch, next to rows of degree 2-4."""
    u = [np.arange(n), np.zeros(n - 1, np.int64), np.ones(n // 3 - 1, np.int64)]
    v = [(np.arange(n) + 1) % n, np.arange(1, n), np.arange(1, n // 3) * 3 + 2]
    u, v = np.concatenate(u), np.concatenate(v)
    pairs = np.unique(np.stack([np.minimum(u, v), np.maximum(u, v)], 1), axis=0)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    uu = np.concatenate([pairs[:, 0], pairs[:, 1]])
    vv = np.concatenate([pairs[:, 1], pairs[:, 0]])
    order = np.lexsort((vv, uu))
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, uu + 1, 1)
    return np.cumsum(xadj).astype(np.int32), vv[order].astype(np.int32), None, None


@pytest.fixture(scope="module")
def graphs(rgg15, weighted132):
    return {
        "manual": manual_graph(),
        "er300w": random_graph(300, 0.03, 11, weighted=True),
        "weighted132": weighted132,
        "rgg15": rgg15,
        "rmat12": K.gen_rmat(12, 16 << 12, 3),
        "hubs": hub_graph(),
    }


@pytest.mark.parametrize("name", ["manual", "er300w", "weighted132", "rgg15", "rmat12", "hubs"])
def test_features_match_oracle(graphs, name):
    x, a, vw, ew = graphs[name]
    bound, seed = 40, 1234
    f = K.edge_features(x, a, vw, ew, bound, seed)
    want = R.edge_features(x, a, vw, ew, bound, seed)
    assert f.shape == want.shape
    assert np.array_equal(f[:, 5], want[:, 5].astype(np.float32))
    assert np.array_equal(f[:, 6], want[:, 6].astype(np.float32))
    cols = [0, 1, 2, 3, 4, 7]
    assert np.max(np.abs(f[:, cols] - want[:, cols])) <= 2e-6


@pytest.mark.parametrize("m", [1, 15, 16, 17, 1000, 65536 + 5])
def test_mlp_forward_matches_oracle(m):
    rng = np.random.default_rng(m)
    X = np.abs(rng.normal(size=(m, 8))) * np.array([3, 5, 2, 1, 2, 2, 3, 3])
    params = rand_params(m)
    z = K.mlp_forward(X.astype(np.float32), params)
    want = R.mlp_forward(X.astype(np.float32), params, bf16=True)
    err = np.abs(z - want)
    assert np.all(err <= 1e-3 * (1 + np.abs(want))), np.max(err)
    assert err.mean() <= 1e-5 * (1 + np.abs(want).mean())
    exact = R.mlp_forward(X.astype(np.float32), params, bf16=False)
    rel = np.abs(z - exact) / (1 + np.abs(exact))
    print(f"m={m}: max |z - fp64| / (1 + |z|) = {rel.max():.3e}, mean {rel.mean():.3e}")
    assert rel.max() <= 0.05


def test_mlp_forward_selector_network_layout():
    # Each MFMA operand slot carries a distinct weight: a mislaid lane or K index changes the sum.
    p = np.zeros(R.PARAMS, np.float32)
    W1 = p[R.OFF_W1:R.OFF_B1].reshape(64, 8)
    W2 = p[R.OFF_W2:R.OFF_B2].reshape(64, 64)
    for h in range(64):
        W1[h, h % 8] = 1.0 + (h // 8) / 8.0
    for h in range(64):
        W2[h, h] = 1.0
        W2[h, (h * 7 + 3) % 64] += 0.5
    p[R.OFF_W3:R.OFF_B3] = (np.arange(64) % 13 + 1) / 16.0
    X = np.random.default_rng(8).integers(0, 8, size=(4096, 8)).astype(np.float32)
    z = K.mlp_forward(X, p)
    want = R.mlp_forward(X, p, bf16=True)
    assert np.max(np.abs(z - want)) <= 1e-3 * (1 + np.abs(want).max())


@pytest.mark.parametrize("count", [1, 64, 100, 1000])
def test_gradient_matches_oracle(count):
    rng = np.random.default_rng(count)
    X = np.abs(rng.normal(size=(count, 8))).astype(np.float32) * 2
    y = (rng.random(count) < 0.3).astype(np.float32)
    params = rand_params(count + 7)
    loss, g = K.mlp_gradient(X, y, params)
    # the MFMA backward takes BF16 operands; against the oracle with the same roundings the
    # difference is fp32 accumulation order plus the rare operand whose fp32 and float64 values
    # round to neighbouring BF16 numbers
    want_loss, want = R.mlp_loss_grad(X, y, params, bf16=True, bf16_backward=True)
    assert abs(loss - want_loss) <= 1e-5 * (1 + abs(want_loss))
    assert np.linalg.norm(g - want) <= 2e-3 * np.linalg.norm(want)
    # and against the exact float64 backward: the BF16 operand error (2^-9 relative per operand)
    _, exact = R.mlp_loss_grad(X, y, params, bf16=True)
    assert np.linalg.norm(g - exact) <= 1e-2 * np.linalg.norm(exact)
    # each weight block on its own (a transposed or mislaid operand shows in its block)
    for lo, hi in ((R.OFF_W1, R.OFF_B1), (R.OFF_W2, R.OFF_B2)):
        assert np.linalg.norm(g[lo:hi] - want[lo:hi]) <= 2e-3 * np.linalg.norm(want[lo:hi]) + 1e-12


def test_training_is_deterministic_and_learns():
    rng = np.random.default_rng(5)
    X = np.abs(rng.normal(size=(20000, 8))).astype(np.float32) * 2
    y = ((X[:, 5] + 0.5 * X[:, 6] - X[:, 1] * 0.3) > 1.0).astype(np.float32)
    p0 = rand_params(9, scale=0.5)
    w1, l1 = K.mlp_train(X, y, p0, batch=1024, epochs=8, lr=5e-3)
    w2, l2 = K.mlp_train(X, y, p0, batch=1024, epochs=8, lr=5e-3)
    assert np.array_equal(w1, w2) and np.array_equal(l1, l2)
    assert l1[-1] < 0.7 * l1[0]
    acc = ((K.mlp_forward(X, w1) > 0) == (y > 0.5)).mean()
    assert acc > 0.9


@pytest.mark.parametrize("name", ["manual", "er300w", "weighted132", "rmat12"])
def test_expansion_star2_scores_exact(graphs, name):
    x, a, vw, ew = graphs[name]
    s = K.edge_scores(x, a, vw, ew, K.RATING_EXPANSION_STAR2)
    assert np.array_equal(s, R.expansion_star2_scores(x, a, vw, ew))


@pytest.mark.parametrize("name", ["er300w", "rgg15", "rmat12"])
def test_learned_scores_are_quantised_logits(graphs, name):
    x, a, vw, ew = graphs[name]
    params = rand_params(21)
    s = K.edge_scores(x, a, vw, ew, K.RATING_LEARNED, params, 40, 99)
    f = K.edge_features(x, a, vw, ew, 40, 99)
    z = K.mlp_forward(f, params)
    assert np.array_equal(s, R.quantize_logit(z, R.score_shift(int(np.diff(x).max()))))
    # against the oracle's float64 chain with the same BF16 roundings: equal but for logits
    # next to a rounding boundary of the 1/128 score grid
    want = R.quantize_logit(R.mlp_forward(R.edge_features(x, a, vw, ew, 40, 99), params,
                                          bf16=True))
    assert np.mean(s == want) >= 0.97
    assert np.max(np.abs(s.astype(np.int64) - want)) <= 2


@pytest.mark.parametrize("name,k,seed", [("rgg15", 2, 0), ("rmat12", 8, 1), ("weighted132", 2, 2)])
def test_rated_pipeline_expansion_matches_mirror(graphs, name, k, seed):
    x, a, vw, ew = graphs[name]
    part, st = K.partition(x, a, vw, ew, k=k, seed=seed, rating=K.RATING_EXPANSION_STAR2)
    want, info = mirror_partition(
        x, a, vw, ew, k, 0.03, seed,
        scores=lambda x_, a_, v_, w_, B, rs: R.expansion_star2_scores(x_, a_, v_, w_))
    assert st["level_n"] == info["level_n"]
    assert np.array_equal(part, want)
    assert st["cut"] == O.edge_cut(x, a, ew, want)


@pytest.mark.parametrize("name,k,seed,levels", [("rgg15", 2, 0, 64), ("rmat12", 8, 1, 64),
                                                ("rgg15", 4, 3, 1)])
def test_rated_pipeline_learned_matches_mirror(graphs, name, k, seed, levels):
    """The clustering choices driven by the integer scores are bit-exact; the scores of each
    level come from the GPU (the float features and MLP are checked above)."""
    x, a, vw, ew = graphs[name]
    params = rand_params(33)
    part, st = K.partition(x, a, vw, ew, k=k, seed=seed, rating=K.RATING_LEARNED,
                           rating_levels=levels, mlp_weights=params)
    want, info = mirror_partition(
        x, a, vw, ew, k, 0.03, seed, rating_levels=levels,
        scores=lambda x_, a_, v_, w_, B, rs: K.edge_scores(x_, a_, v_, w_, K.RATING_LEARNED,
                                                           params, B, rs))
    assert st["level_n"] == info["level_n"]
    assert np.array_equal(part, want)
    assert st["ms_rating"] > 0
