"""The inequalities behind the fast paths' index certification.

Bilinear (and the nearest candidates' slots): the footprint kernel takes a
rotated beam pixel's z from a composite matrix (the reference rotation of
the three unit vectors, its rows turned by -lam about z) and its longitude
as lam + atan(t) by a five-term series, instead of two sequential Rodrigues
steps and atan2 (tod_rotations.py:33-134, numba_healpy.py:358-499).  It takes
the integer part of a ring or slot coordinate only when the fractional part
is at least 1e-9 (1e-6 for the polar-cap ring formula) from an integer; the
tests below measure, against the oracle's binary64 restatement of the
reference, that the coordinates the fast path computes differ by less than
a hundredth of those margins: < 1e-11 for belt ring and slot coordinates, <
1e-8 for polar ring coordinates, at nside 2048 and 8192.

The fast path evaluates the cos-distance sin(theta) sin(theta_r) cos(g) +
z z_r only for each ring's nearer candidate (smaller |g|), with a Taylor
cosine through g^8 for |g| <= 0.05, and bounds the nearer candidate's lead
over the farther one from below by 0.4995 sin(theta) sin(theta_r) (g_l^2 -
g_w^2).  Both facts are checked here against numpy's cosine and the exact
product form cos a - cos b = 2 sin((b + a) / 2) sin((b - a) / 2).
"""

import math

import numpy as np
import pytest

import oracle as O

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


# ------------------------------------------------------- bilinear coordinates

LAM_N = 512  # reference longitudes lam_j = 2 pi j / 512
MAX_TAN = 0.03  # largest |tan| of the longitude offset the fast path takes


def _atan_small(t):
    t2 = t * t
    p = t2 * (1.0 / 9.0) - 1.0 / 7.0
    p = t2 * p + 1.0 / 5.0
    p = t2 * p - 1.0 / 3.0
    return (t * t2) * p + t


def _fast_and_reference(n_samples, n_pix, seed, colat_range):
    """Per (sample, beam pixel) the fast path's z and longitude offset from
    lam, and the reference's z and longitude (oracle, binary64), for beam
    pixels within ~1.2 deg of a north-pole beam centre."""
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(math.cos(colat_range[1]), math.cos(colat_range[0]), n_samples))
    ph = rng.uniform(0, 2 * math.pi, n_samples)
    ps = rng.uniform(0, 2 * math.pi, n_samples)
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    params = R.rotation_params_for_pointing(th, ph, ps)
    r = 0.02 * np.sqrt(rng.uniform(0, 1, n_pix))
    a = rng.uniform(0, 2 * math.pi, n_pix)
    vec = np.stack([np.sin(r) * np.cos(a), np.sin(r) * np.sin(a), np.cos(r)], 1).astype(np.float32)
    d = O.rotated_dirs(vec, *params)  # (B, S, 3) reference directions
    cols = O.rotated_dirs(np.eye(3, dtype=np.float32), *params)  # (B, 3 basis vectors, 3)
    # lam: the grid longitude nearest the beam centre's image
    cz = O.rotated_dirs(np.array([[0, 0, 1]], np.float32), *params)[:, 0]
    jl = np.rint(np.mod(np.arctan2(cz[:, 1], cz[:, 0]), 2 * math.pi) * (LAM_N / (2 * math.pi))).astype(int) % LAM_N
    lam = jl * (2 * math.pi / LAM_N)
    cl, sl = np.cos(lam), np.sin(lam)
    X, Y, Z = cols[..., 0], cols[..., 1], cols[..., 2]  # (B, 3): component k of R e_j
    row0 = cl[:, None] * X + sl[:, None] * Y
    row1 = cl[:, None] * Y - sl[:, None] * X
    u = vec.astype(np.float64)
    xp = row0 @ u.T
    yp = row1 @ u.T
    zf = Z @ u.T
    t = yp / xp
    ok = (xp > 0) & (np.abs(t) < MAX_TAN)
    dl = _atan_small(t)
    zr = np.clip(d[..., 2], -1, 1)
    phr = np.mod(np.arctan2(d[..., 1], d[..., 0]), 2 * math.pi)
    return ok, np.clip(zf, -1, 1), dl, lam[:, None], zr, phr


@pytest.mark.parametrize("nside", [2048, 8192])
@pytest.mark.parametrize("colat", [(0.3, 2.84), (0.06, 0.3)], ids=["belt and caps", "near the pole"])
def test_bilinear_coordinates_within_a_hundredth_of_the_margins(nside, colat):
    ok, zf, dl, lam, zr, phr = _fast_and_reference(400, 300, nside, colat)
    assert ok.sum() >= 10_000  # pairs inside the fast path's longitude range
    zf, dl, zr, phr = zf[ok], dl[ok], zr[ok], phr[ok]
    lamb = np.broadcast_to(lam, ok.shape)[ok]
    belt = np.abs(zr) <= 2.0 / 3.0
    # ring coordinates: belt nside (2 - 1.5 z), caps nside sqrt(3 (1 - |z|))
    rb = np.abs(nside * (2 - 1.5 * zf[belt]) - nside * (2 - 1.5 * zr[belt]))
    cap = ~belt & (np.abs(zr) < 1.0)
    rc = np.abs(nside * np.sqrt(3 * (1 - np.abs(zf[cap]))) - nside * np.sqrt(3 * (1 - np.abs(zr[cap]))))
    # slot coordinates on a ring of n = 4 nside pixels (the belt's; cap rings
    # have fewer, so smaller errors): (phi - phi0) / dphi, phi0 = 0
    n = 4 * nside
    dphi = 2 * math.pi / n
    ref_slot = phr / dphi
    fast_slot = lamb / dphi + dl * (1.0 / dphi)
    ds = np.abs(np.mod(fast_slot - ref_slot + n / 2, n) - n / 2)
    print(f"nside {nside}: max |d ring| belt {rb.max() if rb.size else 0:.2e}, caps "
          f"{rc.max() if rc.size else 0:.2e}; max |d slot| {ds.max():.2e} ({ok.sum()} pairs)")
    assert (rb.max() if rb.size else 0.0) < 1e-11
    assert (rc.max() if rc.size else 0.0) < 1e-8
    assert ds.max() < 1e-11
