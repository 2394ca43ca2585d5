"""The MLP stencil corrector's CPU oracle (oracle/smlp_oracle.py), pinned
before it judges the GPU code:

* its loss gradient equals central finite differences of its loss (exact
  float64 model), so the backward formulas (head, GELU layers) are right;
* with W3 = 0 the network emits its output bias for every pixel, so the
  channels must equal the oracle's fixed-stencil channels for the same 32
  coefficients (or_fcorr_channels, pinned in test_fcorr_oracle.py);
* a constant map gives A0 = the constant and A1 = 0 for any parameters; the
  channels scale with the map (the inputs are normalised by the local RMS);
* the dense operator reproduces the oracle's factorised-corrector TOD
  (linearity of or_gather_accum_fcorr in its two channels);
* the package's initial draw is the oracle's.
"""

import numpy as np
import pytest

import oracle as O
import smlp_oracle as SO


def _tiny(seed=0, nside=4, B=6, S=24):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((S, 3)) * np.array([0.12, 0.12, 0.0]) + np.array([0.0, 0.0, 1.0])
    vec = (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)
    bv = rng.uniform(0.2, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    from cases import random_rot_params

    _, params = random_rot_params(B, S, rng)
    m = rng.standard_normal(12 * nside * nside).astype(np.float32)
    return nside, vec, bv, params, m, B, S


def test_dense_operator_reproduces_the_fcorr_tod():
    nside, vec, bv, params, m, B, S = _tiny(1)
    K0, K1 = SO.dense_operator(vec, params, nside, bv, B, S)
    rng = np.random.default_rng(2)
    A0 = rng.standard_normal(m.size).astype(np.float32)
    A1 = rng.standard_normal(m.size).astype(np.float32)
    t = np.zeros((1, B))
    O.gather_accum_fcorr(vec, *params, nside, A0, A1, bv, B, S, t)
    got = K0 @ A0.astype(np.float64) + K1 @ A1.astype(np.float64)
    assert np.max(np.abs(got - t[0])) <= 1e-12 * np.max(np.abs(t[0]))


def test_gradient_matches_finite_differences():
    nside, vec, bv, params, m, B, S = _tiny(3)
    K0, K1 = SO.dense_operator(vec, params, nside, bv, B, S)
    F, z = SO.features(nside, m)
    x = SO.inputs(F, z)
    rng = np.random.default_rng(4)
    flat = SO.init_params(5, rng.normal(0.0, 0.2, 32), w3_scale=0.5).astype(np.float64)
    target = rng.standard_normal(B)
    norm = 3.0
    loss, g = SO.loss_grad(flat, F, x, K0, K1, target, norm)
    idx = np.concatenate([rng.choice(SO.N_PARAMS, 60, replace=False),
                          [7328 - 32 + 3, 7328 - 32 + 20, 1024 + 5, 5]])
    h = 1e-6
    for i in idx:
        if i in (7328 - 32, 7328 - 16):
            continue  # unused outputs 0 and 16
        fp = flat.copy()
        fm = flat.copy()
        fp[i] += h
        fm[i] -= h
        num = (SO.loss_grad(fp, F, x, K0, K1, target, norm)[0] - SO.loss_grad(fm, F, x, K0, K1, target, norm)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])
    # the unused outputs get no gradient
    assert g[7328 - 32] == 0.0 and g[7328 - 16] == 0.0


def test_zero_w3_gives_the_fixed_stencil_channels():
    nside = 16
    rng = np.random.default_rng(6)
    m = rng.standard_normal(12 * nside * nside).astype(np.float32)
    coef = rng.normal(0.0, 0.3, 32)
    coef[0] = 1.0
    coef[16] = 0.0
    flat = SO.init_params(7, coef, w3_scale=0.0)
    A0, A1 = SO.channels(nside, m, flat)
    R0, R1 = O.fcorr_channels(nside, m, coef)
    # b3 is stored as float32 (the package's parameter type)
    c32 = coef.astype(np.float32).astype(np.float64)
    c32[0] = 1.0
    S0, S1 = O.fcorr_channels(nside, m, c32)
    np.testing.assert_allclose(A0, S0, rtol=0, atol=1e-12 * np.max(np.abs(S0)))
    np.testing.assert_allclose(A1, S1, rtol=0, atol=1e-12 * np.max(np.abs(S1)))
    assert np.max(np.abs(A0 - R0)) <= 1e-6 * np.max(np.abs(R0))


@pytest.mark.parametrize("w3_scale", [0.0, 1.0])
def test_constant_map_and_scaling(w3_scale):
    nside = 8
    rng = np.random.default_rng(8)
    flat = SO.init_params(9, rng.normal(0.0, 0.3, 32), w3_scale=w3_scale)
    c = np.full(12 * nside * nside, -1.75, np.float32)
    A0, A1 = SO.channels(nside, c, flat)
    assert np.max(np.abs(A0 + 1.75)) < 1e-12 and np.max(np.abs(A1)) < 1e-12
    m = rng.standard_normal(12 * nside * nside).astype(np.float32)
    B0, B1 = SO.channels(nside, m, flat)
    C0, C1 = SO.channels(nside, (4.0 * m).astype(np.float32), flat)
    np.testing.assert_allclose(C0, 4.0 * B0, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(C1, 4.0 * B1, rtol=1e-12, atol=1e-12)


def test_package_initial_draw_is_the_oracles():
    from tod_generation_mapbased_beam_amd import smlp

    coef = np.linspace(-0.5, 0.5, 32)
    for seed, sc in ((0, 0.0), (11, 0.3)):
        np.testing.assert_array_equal(smlp.init_params(seed, coef, sc), SO.init_params(seed, coef, sc))
    assert smlp.N_PARAMS == SO.N_PARAMS
