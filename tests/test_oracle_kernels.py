"""The reference's kernel-contract tests, ported against the CPU oracle.

tests/test_tod_bilinear.py:709-1170 and tests/test_tod_nearest.py:320-846:
fused vs split pipeline, zero beam, constant map, in-place accumulation,
determinism (and thread-count independence), Q=1 at boresight, T unchanged
by spin-2, Q/U against a toy model built from the bearing construction,
nearest-pixel hot pixel and Q^2+U^2 preservation.
"""

import math

import numpy as np
import pytest

import oracle as O
from cases import random_rot_params, recenter_params, rel_err


def _fused(vec, params, nside, mp, bv, B, S, c_q=-1, c_u=-1, nthreads=0):
    tod = np.zeros((mp.shape[0], B))
    O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, tod, c_q, c_u, nthreads=nthreads)
    return tod


def _nearest(vec, params, nside, mp, bv, B, S, c_q=-1, c_u=-1):
    tod = np.zeros((mp.shape[0], B))
    O.gather_accum_nearest(vec, *params, nside, mp, bv, B, S, tod, c_q, c_u)
    return tod


def _split(vec, params, nside, mp, bv, B, S):
    """Split pipeline: float32 batch rotation -> angles -> interp weights ->
    gather (tests/test_tod_bilinear.py:619-642)."""
    vr = O.rodrigues_batch(vec, *params).reshape(-1, 3).astype(np.float64)
    rxy = np.sqrt(vr[:, 0] ** 2 + vr[:, 1] ** 2)
    theta = np.arctan2(rxy, vr[:, 2])
    phi = np.mod(np.arctan2(vr[:, 1], vr[:, 0]), 2 * np.pi)
    pix, wgt = O.get_interp_weights(nside, theta, phi)
    tod = np.zeros((mp.shape[0], B))
    O.gather_accum(pix, wgt, bv, mp, B, S, tod)
    return tod


@pytest.mark.parametrize("nside", [4, 16, 64])
def test_fused_agrees_with_split_pipeline(nside):
    rng = np.random.default_rng(45)
    B, S = 8, 20
    vec, params = random_rot_params(B, S, rng)
    mp = rng.uniform(0.5, 1.5, (3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    np.testing.assert_allclose(_fused(vec, params, nside, mp, bv, B, S),
                               _split(vec, params, nside, mp, bv, B, S), atol=1e-5)


def test_zero_beam_constant_map_and_accumulation():
    rng = np.random.default_rng(41)
    nside, B, S = 16, 6, 12
    vec, params = random_rot_params(B, S, rng)
    mp = rng.standard_normal((2, 12 * nside * nside)).astype(np.float32)
    assert np.array_equal(_fused(vec, params, nside, mp, np.zeros(S, np.float32), B, S),
                          np.zeros((2, B)))
    const = np.full((2, 12 * nside * nside), 7.5, np.float32)
    bv = np.full(S, 1.0 / S, np.float32)
    np.testing.assert_allclose(_fused(vec, params, nside, const, bv, B, S), 7.5, atol=1e-4)
    tod = np.zeros((2, B))
    O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, tod)
    first = tod.copy()
    O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, tod)
    np.testing.assert_allclose(tod, 2 * first, atol=1e-14)


def test_deterministic_and_thread_count_independent():
    rng = np.random.default_rng(44)
    nside, B, S = 16, 64, 30
    vec, params = random_rot_params(B, S, rng)
    mp = rng.uniform(0, 1, (3, 12 * nside * nside)).astype(np.float32)
    bv = np.full(S, 1.0 / S, np.float32)
    a = _fused(vec, params, nside, mp, bv, B, S, 1, 2, nthreads=1)
    b = _fused(vec, params, nside, mp, bv, B, S, 1, 2, nthreads=4)
    assert np.array_equal(a, b)


def test_q_one_at_boresight():
    """tests/test_tod_bilinear.py:1027-1076"""
    nside = 64
    rng = np.random.default_rng(50)
    B = 4
    ax = rng.standard_normal((B, 3))
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    ax[ax[:, 2] < -0.95, 2] = -0.9
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    params = recenter_params(ax.astype(np.float32))
    vec = np.array([[0.0, 0.0, 1.0]], np.float32)
    mp = np.zeros((3, 12 * nside * nside), np.float32)
    mp[1] = 1.0
    t = _fused(vec, params, nside, mp, np.ones(1, np.float32), B, 1, 1, 2)
    np.testing.assert_allclose(t[0], 0.0, atol=1e-3)
    np.testing.assert_allclose(t[1], 1.0, atol=1e-3)
    np.testing.assert_allclose(t[2], 0.0, atol=1e-3)


def test_t_unchanged_by_spin2():
    """tests/test_tod_bilinear.py:1122-1170"""
    rng = np.random.default_rng(52)
    nside, B, S = 16, 5, 8
    vec, params = random_rot_params(B, S, rng)
    mp = rng.uniform(-1, 1, (3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    a = _fused(vec, params, nside, mp, bv, B, S, 1, 2)
    b = _fused(vec, params, nside, mp, bv, B, S)
    np.testing.assert_allclose(a[0], b[0], atol=1e-12)


def _toy_qu(vec, params, nside, mp, bv, B, S):
    """Q/U toy model (tests/test_tod_bilinear.py:644-707): interpolation
    weights from the theta-driven lookup, spin-2 from the explicit bearing
    construction with atan2, per neighbour pixel."""
    vr = O.rotated_dirs(vec, *params)
    ax = params[3].astype(np.float64)
    tod = np.zeros((3, B))
    for b in range(B):
        v = vr[b]
        theta = np.arctan2(np.sqrt(v[:, 0] ** 2 + v[:, 1] ** 2), v[:, 2])
        phi = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2 * np.pi)
        pix, w = O.get_interp_weights(nside, theta, phi)
        tb = math.atan2(math.hypot(ax[b, 0], ax[b, 1]), ax[b, 2])
        pb = math.atan2(ax[b, 1], ax[b, 0]) % (2 * math.pi)
        for s in range(S):
            for k in range(4):
                p = pix[k, s]
                zc, pc = O.pix2zphi(nside, p)
                tc = math.acos(zc)
                dphi = pb - pc
                cb = max(-1.0, min(1.0, math.sin(tc) * math.sin(tb) * math.cos(dphi) + zc * math.cos(tb)))
                sb = math.sqrt(max(0.0, 1 - cb * cb))
                if sb < 1e-12:
                    c2, s2 = 1.0, 0.0
                else:
                    sa_ = math.sin(tb) * math.sin(dphi) / sb
                    ca_ = (math.sin(tb) * zc * math.cos(dphi) - math.cos(tb) * math.sin(tc)) / sb
                    sg = math.sin(tc) * math.sin(dphi) / sb
                    cg = -(math.sin(tc) * math.cos(tb) * math.cos(dphi) - zc * math.sin(tb)) / sb
                    d = math.atan2(sa_, ca_) - math.atan2(sg, cg)
                    c2, s2 = math.cos(2 * d), math.sin(2 * d)
                q, u = float(mp[1, p]), float(mp[2, p])
                wb = w[k, s] * float(bv[s])
                tod[0, b] += wb * float(mp[0, p])
                tod[1, b] += wb * (q * c2 + u * s2)
                tod[2, b] += wb * (-q * s2 + u * c2)
    return tod


@pytest.mark.parametrize("nside", [16, 64])
def test_qu_agrees_with_toy_model(nside):
    """tests/test_tod_bilinear.py:1078-1120 (atol/rtol 1e-4)"""
    rng = np.random.default_rng(51 + nside)
    B, S = 6, 12
    vec, params = random_rot_params(B, S, rng)
    mp = rng.uniform(-1, 1, (3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    got = _fused(vec, params, nside, mp, bv, B, S, 1, 2)
    np.testing.assert_allclose(got, _toy_qu(vec, params, nside, mp, bv, B, S), atol=1e-4, rtol=1e-4)


# ------------------------------------------------------------------ nearest

def test_nearest_hot_pixel_at_north_pole():
    """tests/test_tod_nearest.py:687-720"""
    nside = 16
    mp = np.zeros((1, 12 * nside * nside), np.float32)
    # pixel nearest to theta=0.05, phi=0 (ring 1 at nside 16 has pixels 0..3)
    z = math.cos(0.05)
    p, _, _ = O.nearest_pix(nside, z, math.sin(0.05), 0.0)
    mp[0, p] = 42.0
    vec = np.array([[0.0, 0.0, 1.0]], np.float32)
    params = recenter_params(np.array([[0.0, 0.0, 1.0]], np.float32))
    t = _nearest(vec, params, nside, mp, np.ones(1, np.float32), 1, 1)
    assert abs(t[0, 0] - 42.0) < 1e-10


def test_nearest_qu_norm_preserved():
    """Spin-2 is a rotation of (Q, U): single-pixel beam, Q^2+U^2 equals the
    pixel's (tests/test_tod_nearest.py:560-604)."""
    rng = np.random.default_rng(7)
    nside, B = 32, 20
    ax = rng.standard_normal((B, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    params = recenter_params(ax.astype(np.float32))
    vec = np.array([[0.0, 0.0, 1.0]], np.float32)
    mp = rng.uniform(-1, 1, (3, 12 * nside * nside)).astype(np.float32)
    t = _nearest(vec, params, nside, mp, np.ones(1, np.float32), B, 1, 1, 2)
    pix = O.nearest_indices(vec, *params, nside, 0, B)[:, 0]
    np.testing.assert_allclose(t[1] ** 2 + t[2] ** 2,
                               mp[1, pix].astype(np.float64) ** 2 + mp[2, pix].astype(np.float64) ** 2,
                               rtol=1e-12)
    np.testing.assert_allclose(t[0], mp[0, pix], rtol=0, atol=0)


def test_nearest_equal_q_u_slots_do_not_crash():
    rng = np.random.default_rng(50)
    nside, B, S = 16, 3, 4
    vec, params = random_rot_params(B, S, rng)
    mp = rng.standard_normal((2, 12 * nside * nside)).astype(np.float32)
    t = _nearest(vec, params, nside, mp, np.full(S, 0.25, np.float32), B, S, 0, 0)
    assert np.all(np.isfinite(t))


def test_rel_err_helper():
    a = np.array([[1.0, -1.0, 2.0]])
    assert rel_err(a, a) == 0.0
