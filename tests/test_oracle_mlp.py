"""Pins of the oracle's rank-model restatement (CPU only): the analytic
gradient against central finite differences, and the product's host-side
model initialization (normalization statistics, Glorot bounds)."""
import numpy as np

import iesolver
import oracle


def _data(n=300, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, oracle.RANK_FEATURES)) * [1, 2, 3, 1, 0.5, 2, 0.1, 1, 0.5] + 4
    y = 15 + 3 * np.tanh(X[:, 0] - 4) + X[:, 2] + rng.normal(size=n)
    return X, y


def _model(seed=0):
    rng = np.random.default_rng(seed)
    m = rng.normal(size=oracle.RANK_MODEL_LEN) * 0.3
    F = oracle.RANK_FEATURES
    m[0:F] = 4.0
    m[F:2 * F] = 1.5
    m[-2], m[-1] = 15.0, 3.0
    return m


def test_layout_matches_the_product():
    assert oracle.RANK_MODEL_LEN == iesolver.RANK_MODEL_LEN
    assert oracle.RANK_PARAMS == iesolver.RANK_PARAMS
    assert oracle.RANK_FEATURES == iesolver.RANK_FEATURES


def test_gradient_matches_central_differences():
    X, y = _data()
    m = _model()
    loss, g = oracle.mlp_loss_grad(m, X, y)
    assert g.shape == (oracle.RANK_PARAMS,)
    off = 2 * oracle.RANK_FEATURES
    rng = np.random.default_rng(1)
    for k in list(rng.choice(oracle.RANK_PARAMS, 60, replace=False)) + [0, oracle.RANK_PARAMS - 1]:
        h = 1e-6
        mp, mm = m.copy(), m.copy()
        mp[off + k] += h
        mm[off + k] -= h
        fd = (oracle.mlp_loss_grad(mp, X, y)[0] - oracle.mlp_loss_grad(mm, X, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, (k, fd, g[k])


def test_predict_is_the_forward_pass():
    X, y = _data(20)
    m = _model()
    p = oracle.mlp_predict(m, X)
    loss, _ = oracle.mlp_loss_grad(m, X, y)
    assert np.isclose(loss, np.mean(((p - y) / m[-1]) ** 2))


def test_host_model_init():
    X, y = _data()
    s = iesolver.RSSolver(device=-1)
    m = s.rank_model_init(X, y, seed=7)
    F, H = oracle.RANK_FEATURES, oracle.RANK_HID
    assert m.shape == (oracle.RANK_MODEL_LEN,)
    assert np.allclose(m[:F], X.mean(0), rtol=1e-5)
    assert np.allclose(m[F:2 * F], X.std(0), rtol=1e-4)
    assert np.isclose(m[-2], y.mean(), rtol=1e-5) and np.isclose(m[-1], y.std(), rtol=1e-4)
    W1 = m[2 * F:2 * F + F * H]
    assert np.abs(W1).max() <= np.sqrt(6 / (F + H)) and np.abs(W1).std() > 0.1
    assert np.array_equal(m, s.rank_model_init(X, y, seed=7))
    assert not np.array_equal(m, s.rank_model_init(X, y, seed=8))
