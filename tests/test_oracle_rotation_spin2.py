"""Pins for the oracle's rotations and spin-2 factors.

Rodrigues: closed forms (identity, 90 degrees about z, roll), norm
preservation and agreement with scipy's independent rotation-vector
implementation.  Spin-2: the explicit bearing construction with atan2 that
the reference's tests use as their reference (tests/test_tod_bilinear.py:
259-292), plus its invariants.
"""

import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from tod_generation_mapbased_beam_amd import tod_rotations as R


def _rod(v, k, a, p, psi):
    return np.array(O.rodrigues_apply_one(*v, *k, math.cos(a), math.sin(a), *p,
                                          math.cos(psi), math.sin(psi)))


def test_rodrigues_identity_and_quarter_turns():
    """tests/test_tod_rotations.py:162-231"""
    v = (0.3, -0.2, math.sqrt(1 - 0.13))
    np.testing.assert_allclose(_rod(v, (0, 0, 1), 0.0, (0, 0, 1), 0.0), v, atol=1e-15)
    np.testing.assert_allclose(_rod((1, 0, 0), (0, 0, 1), math.pi / 2, (0, 0, 1), 0.0),
                               (0, 1, 0), atol=1e-15)
    np.testing.assert_allclose(_rod((1, 0, 0), (0, 0, 1), math.pi / 2, (0, 0, 1), math.pi / 2),
                               (-1, 0, 0), atol=1e-15)


def test_rodrigues_matches_scipy_composition():
    rng = np.random.default_rng(0)
    for _ in range(300):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        k = rng.standard_normal(3)
        k /= np.linalg.norm(k)
        p = rng.standard_normal(3)
        p /= np.linalg.norm(p)
        a = rng.uniform(0, math.pi)
        psi = rng.uniform(0, 2 * math.pi)
        ref = Rotation.from_rotvec(p * psi).apply(Rotation.from_rotvec(k * a).apply(v))
        got = _rod(v, k, a, p, psi)
        np.testing.assert_allclose(got, ref, atol=1e-14)
        assert abs(np.linalg.norm(got) - 1) < 1e-14


def test_rodrigues_batch_float32_contract():
    """tests/test_tod_rotations.py:309-342: the mixed float32/float64 batch
    kernel agrees with the binary64 scalar kernel to float32 rounding."""
    rng = np.random.default_rng(1)
    B, S = 6, 9
    v = rng.standard_normal((S, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v.astype(np.float32)
    k = rng.standard_normal((B, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    a = rng.uniform(0, math.pi, B)
    p = rng.standard_normal((B, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    ps = rng.uniform(0, 2 * math.pi, B)
    ops = [k.astype(np.float32), np.cos(a).astype(np.float32), np.sin(a).astype(np.float32),
           p.astype(np.float32), np.cos(ps).astype(np.float32), np.sin(ps).astype(np.float32)]
    out = O.rodrigues_batch(v, *ops)
    assert out.dtype == np.float32 and out.shape == (B, S, 3)
    exact = O.rotated_dirs(v, *ops)
    np.testing.assert_allclose(out, exact, atol=1e-6)


def test_precompute_rotation_vector_batch_properties():
    """tests/test_tod_rotations.py:500-570"""
    N = 201
    ra = np.zeros((N, N))
    dec = np.zeros((N, N))
    rng = np.random.default_rng(3)
    phi = rng.uniform(0, np.pi / 6, 20)
    theta = rng.uniform(np.pi / 3, 2 * np.pi / 3, 20)
    rot, beta = R.precompute_rotation_vector_batch(ra, dec, phi, theta, center_idx=(N // 2, N // 2))
    assert rot.shape == (20, 3) and beta.shape == (20,) and rot.dtype == np.float64
    assert np.all((beta >= 0) & (beta < 2 * np.pi))
    rot, _ = R.precompute_rotation_vector_batch(ra, dec, np.array([0.0]), np.array([np.pi / 2]),
                                                center_idx=(N // 2, N // 2))
    assert np.linalg.norm(rot) < 1e-10
    rot, _ = R.precompute_rotation_vector_batch(ra, dec, np.array([np.pi / 2]), np.array([np.pi / 2]),
                                                center_idx=(N // 2, N // 2))
    assert abs(np.linalg.norm(rot) - np.pi / 2) < 1e-6


def test_rotation_maps_beam_centre_to_boresight():
    """The full host prep + rotation sends the pole-centred beam's centre to
    each boresight and keeps the beam frame's north aligned with psi."""
    from tod_generation_mapbased_beam_amd import inputs

    th, ph, ps = inputs.ring_scan(200, 10.0)
    axes, ca, sa, ax_pts, cp, sp = R.rotation_params_for_pointing(th, ph, ps)
    for b in range(0, 200, 13):
        got = np.array(O.rodrigues_apply_one(0.0, 0.0, 1.0, *axes[b], ca[b], sa[b], *ax_pts[b], cp[b], sp[b]))
        want = np.array([np.sin(th[b]) * np.cos(ph[b]), np.sin(th[b]) * np.sin(ph[b]), np.cos(th[b])])
        np.testing.assert_allclose(got, want, atol=2e-6)


def test_rotation_params_float32_contract():
    """tests/test_tod_rotations.py:101-154"""
    rng = np.random.default_rng(2)
    rv = rng.standard_normal((7, 3)) * 0.3
    rv[0] = 0.0
    out = R._rotation_params(rv, rng.uniform(0, 6, 7), rng.uniform(0.1, 3, 7), rng.uniform(0, 6, 7))
    assert all(o.dtype == np.float32 for o in out)
    axes, cos_a, sin_a, ax_pts, cos_p, sin_p = out
    assert np.all(axes[0] == 0) and cos_a[0] == 1.0
    np.testing.assert_allclose(np.linalg.norm(axes[1:], axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(ax_pts, axis=1), 1.0, atol=1e-6)


# ------------------------------------------------------------------ spin-2

def _atan2_construction(t1, p1, t2, p2):
    """Bearing construction (tests/test_tod_bilinear.py:259-292)."""
    z1, s1 = math.cos(t1), math.sin(t1)
    z2, s2 = math.cos(t2), math.sin(t2)
    d = p2 - p1
    cd, sd = math.cos(d), math.sin(d)
    cb = max(-1.0, min(1.0, s1 * s2 * cd + z1 * z2))
    sb = math.sqrt(max(0.0, 1.0 - cb * cb))
    if sb < 1e-12:
        return None
    alpha = math.atan2(s2 * sd / sb, (s2 * z1 * cd - z2 * s1) / sb)
    gamma = math.atan2(s1 * sd / sb, -(s1 * z2 * cd - z1 * s2) / sb)
    dl = alpha - gamma
    return math.cos(2 * dl), math.sin(2 * dl)


def test_spin2_matches_bearing_construction():
    rng = np.random.default_rng(42)
    n = 0
    for _ in range(400):
        t1, p1 = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
        t2, p2 = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
        ref = _atan2_construction(t1, p1, t2, p2)
        if ref is None:
            continue
        c2, s2 = O.spin2(math.cos(t1), math.sin(t1), p1, math.cos(t2), math.sin(t2), p2)
        assert abs(c2 - ref[0]) < 1e-12 and abs(s2 - ref[1]) < 1e-12
        n += 1
    assert n > 350


def test_spin2_invariants():
    """tests/test_tod_bilinear.py:230-327"""
    z, s = math.cos(0.5), math.sin(0.5)
    assert O.spin2(z, s, 1.2, z, s, 1.2) == (1.0, 0.0)
    assert O.spin2(1.0, 0.0, 0.0, 1.0, 0.0, 0.0) == (1.0, 0.0)
    rng = np.random.default_rng(0)
    for _ in range(200):
        t1, p1 = rng.uniform(0.05, 3.09), rng.uniform(0, 6.28)
        t2, p2 = rng.uniform(0.05, 3.09), rng.uniform(0, 6.28)
        a = (math.cos(t1), math.sin(t1), p1)
        b = (math.cos(t2), math.sin(t2), p2)
        c, sn = O.spin2(*a, *b)
        assert abs(c * c + sn * sn - 1) < 1e-14
        c2, s2 = O.spin2(*b, *a)  # antisymmetry
        assert abs(c - c2) < 1e-12 and abs(sn + s2) < 1e-12
    for t1, t2 in [(0.3, 1.0), (1.0, 2.5), (0.1, math.pi - 0.1)]:  # same meridian
        c, sn = O.spin2(math.cos(t1), math.sin(t1), 0.0, math.cos(t2), math.sin(t2), 0.0)
        assert abs(c - 1) < 1e-12 and abs(sn) < 1e-12
    eps = 1e-7
    c, sn = O.spin2(math.cos(1.0), math.sin(1.0), 1.5, math.cos(1 + eps), math.sin(1 + eps), 1.5 + eps)
    assert math.isfinite(c) and abs(c * c + sn * sn - 1) < 1e-12
