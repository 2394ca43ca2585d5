"""The corrector's CPU oracle (oracle/mlp_oracle.py) pinned by mathematics:
its analytic gradient equals central finite differences of its loss, its
weights sum to one (a constant neighbourhood predicts the constant), and the
package draws the same initial parameters."""

import numpy as np

import mlp_oracle as M
from tod_generation_mapbased_beam_amd import corrector


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.N_IN))
    X[:, :M.GEO] = rng.uniform(-0.5, 0.5, (n, M.GEO))
    xh = X[:, M.GEO:]
    xh -= xh.mean(axis=1, keepdims=True)
    xh /= np.sqrt(np.mean(xh ** 2, axis=1, keepdims=True))
    w = rng.uniform(0.0, 1.0, (n, 4))
    w /= w.sum(axis=1, keepdims=True)
    t = rng.standard_normal(n) * 0.3
    return X, w, t


def test_gradient_matches_finite_differences():
    X, w, t = _batch(40, 1)
    p = M.init_params(3).astype(np.float64)
    p += np.random.default_rng(4).standard_normal(p.size) * 0.05  # nonzero biases too
    loss, g, _ = M.loss_grad(p, X, w, t)
    rng = np.random.default_rng(5)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _, _ = M.loss_grad(p + e, X, w, t)
        lm, _, _ = M.loss_grad(p - e, X, w, t)
        fd = (lp - lm) / (2 * h) / X.shape[0]
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_derivative():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_grad(z), fd, atol=1e-8)


def test_weights_sum_to_one_and_constant_neighbourhood():
    X, w, _ = _batch(10, 2)
    p = M.init_params(7)
    O, _ = M.forward(p, X)
    corr = O - O.mean(axis=1, keepdims=True)
    np.testing.assert_allclose(corr.sum(axis=1), 0.0, atol=1e-12)
    X0 = X.copy()
    X0[:, M.GEO:] = 0.0  # constant neighbourhood: xhat = 0 -> prediction = mean
    np.testing.assert_allclose(M.head(M.forward(p, X0)[0], X0, w), 0.0, atol=0)


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0], np.float32)
    r = M.bf16(a)
    # 1 + 2^-8 is halfway between two bf16 values and rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert r[3] == -3.140625 and r[4] == 65536.0


def test_package_init_matches_oracle():
    np.testing.assert_array_equal(corrector.init_params(11), M.init_params(11))
