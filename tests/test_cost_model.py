"""The cost model of the batching form on the CPU: the numpy oracle's
gradients against central finite differences, the feature function, and the
committed timing database and model."""
import os

import numpy as np
import pytest

import kfft_amd as K
from oracle import mlp as M

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "kfft_amd", "data")


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, h, B = 7, 9, 13
    p = rng.normal(0, 0.5, M.param_count(nin, h))
    X = rng.normal(0, 1, (B, nin))
    t = rng.normal(0, 1, B)
    g = M.gradient(p, X, t, nin, h)
    eps = 1e-6
    for i in rng.choice(p.size, 40, replace=False):
        d = np.zeros_like(p)
        d[i] = eps
        fd = (M.loss(p + d, X, t, nin, h) - M.loss(p - d, X, t, nin, h)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_oracle_adam_first_step_is_lr_sign():
    """With bias correction the first Adam step moves every parameter by lr
    against the sign of its gradient (up to eps)."""
    rng = np.random.default_rng(1)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    q, _, _ = M.adam(p, g, np.zeros(50), np.zeros(50), 1, 0.01)
    assert np.allclose(p - q, 0.01 * np.sign(g), atol=1e-6)


def test_features_shape_and_range():
    for n, kind, cnt in [(16, 0, 1), (65537, 0, 64), (1000, 1, 5), (4097, 0, 3), (65536, 1, 256)]:
        for ch in K.tune_candidates():
            f = K.cost_features(n, kind, cnt, ch)
            assert f.shape == (19,) and f.dtype == np.float32
            assert np.all(f >= 0) and np.all(f <= 1.5)


def test_features_name_the_algorithm():
    f = K.cost_features(65537, 0, 1, (16, 0, 0))
    assert f[6] == 1 and f[7] == 0          # Rader
    f = K.cost_features(4099, 0, 1, (16, 0, 0))  # 4098 = 2 * 3 * 683: Bluestein
    assert f[6] == 0 and f[7] == 1
    f = K.cost_features(4096, 0, 1, (16, 0, 0))
    assert f[6] == 0 and f[7] == 0 and f[18] == 1  # one pass
    assert K.cost_features(4096, 0, 1, (16, 1, 0))[18] == 0


@pytest.mark.skipif(not os.path.exists(os.path.join(DATA, "tune_db_mi355x.npz")),
                    reason="timing database not collected")
def test_timing_database_is_consistent():
    db = dict(np.load(os.path.join(DATA, "tune_db_mi355x.npz")))
    assert set(db) >= {"n", "kind", "count", "cls", "cand", "us"}
    assert len(db["us"]) >= 20000 and np.all(db["us"] > 0)
    assert np.all((db["n"] >= 16) & (db["n"] <= 65536))
    # every class has its default choice measured
    for c in np.unique(db["cls"])[:200]:
        assert ((db["cls"] == c) & (db["cand"] == 0)).any()
