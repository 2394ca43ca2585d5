"""The inequalities behind the nearest-pixel fast path's certification.

The fast path evaluates the cos-distance sin(theta) sin(theta_r) cos(g) +
z z_r only for each ring's nearer candidate (smaller |g|), with a Taylor
cosine through g^8 for |g| <= 0.05, and bounds the nearer candidate's lead
over the farther one from below by 0.4995 sin(theta) sin(theta_r) (g_l^2 -
g_w^2).  Both facts are checked here against numpy's cosine and the exact
product form cos a - cos b = 2 sin((b + a) / 2) sin((b - a) / 2).
"""

import numpy as np

G_MAX = 0.05  # largest longitude gap the fast path accepts
HALF_LOW = 0.4995  # lower-bound factor of the same-ring lead


def _cos_taylor8(x):
    x2 = x * x
    p = x2 * (1.0 / 40320.0) - 1.0 / 720.0
    p = x2 * p + 1.0 / 24.0
    p = x2 * p - 0.5
    return x2 * p + 1.0


def test_taylor_cosine_matches_cos_to_an_ulp():
    x = np.linspace(-G_MAX, G_MAX, 200_001)
    err = np.abs(_cos_taylor8(x) - np.cos(x))
    assert err.max() <= 2.0 ** -52


def test_same_ring_lead_lower_bound():
    rng = np.random.default_rng(3)
    n = 1_000_000
    gl = rng.uniform(-G_MAX, G_MAX, n)
    gw = rng.uniform(0.0, 1.0, n) * gl * rng.choice([-1.0, 1.0], n)
    # the extremes: both gaps at the limit, nearly equal, and one at zero
    gl = np.concatenate([gl, [G_MAX, -G_MAX, G_MAX, G_MAX, 1e-3]])
    gw = np.concatenate([gw, [G_MAX * (1 - 1e-6), G_MAX * (1 - 1e-3), 0.0, -G_MAX * 0.999, 1e-3 * (1 - 1e-9)]])
    assert np.all(np.abs(gw) <= np.abs(gl))
    lead = 2.0 * np.sin((gl + gw) / 2.0) * np.sin((gl - gw) / 2.0)  # cos gw - cos gl
    lead = np.where(np.abs(gl) == np.abs(gw), 0.0, np.abs(lead))
    bound = HALF_LOW * (gl * gl - gw * gw)
    assert np.all(bound <= lead * (1.0 + 1e-12) + 1e-300)
    # the factor is tight to ~4e-4 at the limit, so a bound of 0.5 would fail
    assert np.any(0.5 * (gl * gl - gw * gw) > lead * (1.0 + 1e-9))


def test_nearer_candidate_wins_within_a_ring():
    """Same sin(theta_r) and z_r: the larger cos-distance is the smaller |g|."""
    rng = np.random.default_rng(4)
    n = 100_000
    ss = rng.uniform(1e-6, 1.0, n)
    zz = rng.uniform(-1.0, 1.0, n)
    g0 = rng.uniform(-G_MAX, G_MAX, n)
    g1 = g0 - rng.uniform(1e-4, 2 * G_MAX, n)
    d0 = ss * np.cos(g0) + zz
    d1 = ss * np.cos(g1) + zz
    decided = np.abs(d0 - d1) > 1e-12
    assert np.array_equal((np.abs(g1) < np.abs(g0))[decided], (d1 > d0)[decided])
